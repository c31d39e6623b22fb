// Library core: errors, device init, prefix sums, map transpose, device
// sparsity builder, pattern expansion, greedy colouring and row batching.
#include <algorithm>
#include <climits>
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"

namespace op2hip {

static thread_local char g_error[1024];

void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_error, sizeof(g_error), fmt, ap);
  va_end(ap);
}

// ---------------------------------------------------------------------------
// small utility kernels

__global__ void fill_i32(int32_t *p, int64_t n, int32_t v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

static int fill(int32_t *p, int64_t n, int32_t v, hipStream_t s) {
  if (n <= 0) return OP2HIP_OK;
  const uint32_t grid = std::min<uint32_t>(ceil_div(n, 256), 16384);
  fill_i32<<<grid, 256, 0, s>>>(p, n, v);
  OP2HIP_LAUNCHED();
  return OP2HIP_OK;
}

// ---------------------------------------------------------------------------
// exclusive scan: tiles of kScanTile items, one 256-thread block per tile.

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

__device__ __forceinline__ int64_t block_inclusive_scan(int64_t v, int64_t *lds) {
  const int t = threadIdx.x;
  lds[t] = v;
  __syncthreads();
  for (int off = 1; off < kScanThreads; off <<= 1) {
    const int64_t add = t >= off ? lds[t - off] : 0;
    __syncthreads();
    lds[t] += add;
    __syncthreads();
  }
  return lds[t];
}

// out[i+1] = tile-local inclusive prefix; tile_sums[b] = tile total.
__global__ void scan_tiles(int64_t n, const int32_t *__restrict__ in,
                           int32_t *__restrict__ out, int64_t *__restrict__ tile_sums) {
  __shared__ int64_t lds[kScanThreads];
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  int64_t v[kScanItems];
  int64_t run = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const int64_t i = base + k;
    run += (i < n) ? in[i] : 0;
    v[k] = run;
  }
  const int64_t incl = block_inclusive_scan(run, lds);
  const int64_t excl = incl - run;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const int64_t i = base + k;
    if (i < n) out[i + 1] = static_cast<int32_t>(excl + v[k]);
  }
  if (threadIdx.x == kScanThreads - 1) tile_sums[blockIdx.x] = incl;
}

// Single block: exclusive scan of tile_sums in place, total to *total.
__global__ void scan_tile_sums(int64_t ntiles, int64_t *__restrict__ tile_sums,
                               int64_t *__restrict__ total) {
  __shared__ int64_t lds[kScanThreads];
  int64_t carry = 0;
  for (int64_t b = 0; b < ntiles; b += kScanThreads) {
    const int64_t i = b + threadIdx.x;
    const int64_t v = i < ntiles ? tile_sums[i] : 0;
    const int64_t incl = block_inclusive_scan(v, lds);
    if (i < ntiles) tile_sums[i] = carry + incl - v;
    __syncthreads();
    carry += lds[kScanThreads - 1];
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = carry;
}

__global__ void scan_add_offsets(int64_t n, int32_t *__restrict__ out,
                                 const int64_t *__restrict__ tile_sums) {
  const int64_t base = (int64_t)blockIdx.x * kScanTile;
  const int64_t off = tile_sums[blockIdx.x];
  for (int k = threadIdx.x; k < kScanTile; k += kScanThreads) {
    const int64_t i = base + k;
    if (i < n) out[i + 1] = static_cast<int32_t>(out[i + 1] + off);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) out[0] = 0;
}

static size_t scan_bytes(int64_t n) {
  const int64_t ntiles = std::max<int64_t>(1, (n + kScanTile - 1) / kScanTile);
  return (size_t)(ntiles + 2) * sizeof(int64_t);
}

static int scan(int64_t n, const int32_t *in, int32_t *out, int64_t *total_host,
                void *work, size_t work_bytes, hipStream_t s) {
  OP2HIP_REQUIRE(n >= 0, "scan: negative length");
  if (work_bytes < scan_bytes(n)) {
    set_error("scan: workspace %zu < %zu bytes", work_bytes, scan_bytes(n));
    return OP2HIP_ERR_WORKSPACE;
  }
  const int64_t ntiles = std::max<int64_t>(1, (n + kScanTile - 1) / kScanTile);
  int64_t *tile_sums = static_cast<int64_t *>(work);
  int64_t *total = tile_sums + ntiles;
  scan_tiles<<<(uint32_t)ntiles, kScanThreads, 0, s>>>(n, in, out, tile_sums);
  OP2HIP_LAUNCHED();
  scan_tile_sums<<<1, kScanThreads, 0, s>>>(ntiles, tile_sums, total);
  OP2HIP_LAUNCHED();
  scan_add_offsets<<<(uint32_t)ntiles, kScanThreads, 0, s>>>(n, out, tile_sums);
  OP2HIP_LAUNCHED();
  int64_t t = 0;
  OP2HIP_CHECK(hipMemcpyAsync(&t, total, sizeof(t), hipMemcpyDeviceToHost, s));
  OP2HIP_CHECK(hipStreamSynchronize(s));
  if (total_host) *total_host = t;
  if (t > INT32_MAX) {
    set_error("scan: total %lld exceeds int32", (long long)t);
    return OP2HIP_ERR_OVERFLOW;
  }
  return OP2HIP_OK;
}

// ---------------------------------------------------------------------------
// map transpose

__global__ void transpose_count(int32_t nrows, const int32_t *__restrict__ map,
                                int64_t nent, int32_t *__restrict__ counts,
                                int32_t *__restrict__ bad) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nent;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t r = map[i];
    if (r < 0) continue;
    if (r >= nrows) { atomicAdd(bad, 1); continue; }
    atomicAdd(&counts[r], 1);
  }
}

__global__ void transpose_fill(int32_t nrows, const int32_t *__restrict__ map,
                               int64_t nent, int32_t *__restrict__ cursor,
                               int32_t *__restrict__ inc_ent) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nent;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t r = map[i];
    if (r < 0 || r >= nrows) continue;
    const int32_t pos = atomicAdd(&cursor[r], 1);
    inc_ent[pos] = static_cast<int32_t>(i);
  }
}

// Sort each row's incidence list ascending (lists are short: one thread each).
__global__ void transpose_sort(int32_t nrows, const int32_t *__restrict__ inc_ptr,
                               int32_t *__restrict__ inc_ent) {
  for (int32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < nrows;
       r += gridDim.x * blockDim.x) {
    const int32_t lo = inc_ptr[r], hi = inc_ptr[r + 1];
    for (int32_t a = lo + 1; a < hi; ++a) {
      const int32_t v = inc_ent[a];
      int32_t b = a - 1;
      while (b >= lo && inc_ent[b] > v) {
        inc_ent[b + 1] = inc_ent[b];
        --b;
      }
      inc_ent[b + 1] = v;
    }
  }
}

static size_t transpose_bytes(int32_t nrows) {
  // counts/cursor (nrows int32) + flag + scan workspace, 256-byte aligned parts
  size_t a = ((size_t)std::max(nrows, 1) * 4 + 255) & ~(size_t)255;
  return a + 256 + scan_bytes(nrows);
}

// ---------------------------------------------------------------------------
// sparsity rows: one 64-thread block per block row

struct PairTable {
  const int32_t *cmap[OP2HIP_MAX_PAIRS];
  const int32_t *inc_ptr[OP2HIP_MAX_PAIRS];
  const int32_t *inc_ent[OP2HIP_MAX_PAIRS];
  int32_t rarity[OP2HIP_MAX_PAIRS];
  int32_t carity[OP2HIP_MAX_PAIRS];
  int32_t npairs;
};

constexpr int kRowCap = 2048;

__device__ __forceinline__ int32_t candidate(const PairTable &pt, int32_t r, int64_t k,
                                             const int32_t *lo, const int32_t *cnt,
                                             int32_t diag_col) {
  for (int p = 0; p < pt.npairs; ++p) {
    const int64_t n = (int64_t)cnt[p] * pt.carity[p];
    if (k < n) {
      const int32_t inc = lo[p] + (int32_t)(k / pt.carity[p]);
      const int32_t j = (int32_t)(k % pt.carity[p]);
      const int32_t e = pt.inc_ent[p][inc] / pt.rarity[p];
      return pt.cmap[p][(int64_t)e * pt.carity[p] + j];
    }
    k -= n;
  }
  return diag_col;  // the diagonal candidate (or -1)
}

__device__ __forceinline__ int32_t block64_sum(int32_t v, int32_t *lds) {
  // 64-thread block reduction through LDS
  lds[threadIdx.x] = v;
  __syncthreads();
  for (int off = 32; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) lds[threadIdx.x] += lds[threadIdx.x + off];
    __syncthreads();
  }
  const int32_t s = lds[0];
  __syncthreads();
  return s;
}

__device__ __forceinline__ int32_t block64_min(int32_t v, int32_t *lds) {
  lds[threadIdx.x] = v;
  __syncthreads();
  for (int off = 32; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) lds[threadIdx.x] = min(lds[threadIdx.x], lds[threadIdx.x + off]);
    __syncthreads();
  }
  const int32_t s = lds[0];
  __syncthreads();
  return s;
}

template <bool kFill>
__global__ void __launch_bounds__(64)
sparsity_rows(int32_t nrows, int32_t ncols, PairTable pt, int32_t set_diag,
              const int32_t *__restrict__ rowptr, int32_t *__restrict__ out) {
  __shared__ int32_t cand[kRowCap];
  __shared__ uint8_t first[kRowCap];
  __shared__ int32_t red[64];
  __shared__ int32_t lo[OP2HIP_MAX_PAIRS], cnt[OP2HIP_MAX_PAIRS];
  for (int32_t r = blockIdx.x; r < nrows; r += gridDim.x) {
    if (threadIdx.x < (unsigned)pt.npairs) {
      const int p = threadIdx.x;
      lo[p] = pt.inc_ptr[p][r];
      cnt[p] = pt.inc_ptr[p][r + 1] - lo[p];
    }
    __syncthreads();
    int64_t m = 0;
    for (int p = 0; p < pt.npairs; ++p) m += (int64_t)cnt[p] * pt.carity[p];
    const bool has_diag = set_diag && r < ncols;
    const int32_t diag_col = has_diag ? r : -1;
    m += has_diag ? 1 : 0;
    const int32_t base = kFill ? rowptr[r] : 0;
    if (m <= kRowCap) {
      for (int64_t k = threadIdx.x; k < m; k += 64)
        cand[k] = candidate(pt, r, k, lo, cnt, diag_col);
      __syncthreads();
      int32_t mine = 0;
      for (int64_t k = threadIdx.x; k < m; k += 64) {
        const int32_t c = cand[k];
        bool f = c >= 0;
        for (int64_t t = 0; f && t < k; ++t) f = cand[t] != c;
        first[k] = f;
        mine += f;
      }
      __syncthreads();
      if (!kFill) {
        const int32_t total = block64_sum(mine, red);
        if (threadIdx.x == 0) out[r] = total;
      } else {
        for (int64_t k = threadIdx.x; k < m; k += 64) {
          if (!first[k]) continue;
          const int32_t c = cand[k];
          int32_t rank = 0;
          for (int64_t t = 0; t < m; ++t) rank += (first[t] && cand[t] < c);
          out[base + rank] = c;
        }
      }
      __syncthreads();
    } else {
      // long row: repeated minimum extraction straight from the maps
      int32_t last = -1, count = 0;
      for (;;) {
        int32_t best = INT32_MAX;
        for (int64_t k = threadIdx.x; k < m; k += 64) {
          const int32_t c = candidate(pt, r, k, lo, cnt, diag_col);
          if (c > last && c < best) best = c;
        }
        best = block64_min(best, red);
        if (best == INT32_MAX) break;
        if (kFill && threadIdx.x == 0) out[base + count] = best;
        ++count;
        last = best;
      }
      if (!kFill && threadIdx.x == 0) out[r] = count;
      __syncthreads();
    }
  }
}

__global__ void expand_rowptr(int32_t nrows, const int32_t *__restrict__ rowptr,
                              int32_t rdim, int32_t cdim, int32_t *__restrict__ out) {
  for (int32_t r = blockIdx.x * blockDim.x + threadIdx.x; r <= nrows;
       r += gridDim.x * blockDim.x) {
    if (r == nrows) {
      out[(int64_t)nrows * rdim] = rowptr[nrows] * rdim * cdim;
      continue;
    }
    const int32_t len = rowptr[r + 1] - rowptr[r];
    for (int ri = 0; ri < rdim; ++ri)
      out[(int64_t)r * rdim + ri] = rowptr[r] * rdim * cdim + ri * len * cdim;
  }
}

__global__ void expand_colidx(int32_t nrows, const int32_t *__restrict__ rowptr,
                              const int32_t *__restrict__ colidx, int32_t rdim,
                              int32_t cdim, const int32_t *__restrict__ rowptr_out,
                              int32_t *__restrict__ out) {
  for (int32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < nrows;
       r += gridDim.x * blockDim.x) {
    const int32_t lo = rowptr[r], len = rowptr[r + 1] - lo;
    for (int ri = 0; ri < rdim; ++ri) {
      int32_t *o = out + rowptr_out[(int64_t)r * rdim + ri];
      for (int32_t s = 0; s < len; ++s)
        for (int cj = 0; cj < cdim; ++cj) o[s * cdim + cj] = colidx[lo + s] * cdim + cj;
    }
  }
}

static int make_pairs(const op2hip_map_pair *pairs, int32_t npairs,
                      const int32_t *const *inc_ptr, const int32_t *const *inc_ent,
                      PairTable *pt) {
  OP2HIP_REQUIRE(npairs >= 0 && npairs <= OP2HIP_MAX_PAIRS,
                 "sparsity: %d map pairs (max %d)", npairs, OP2HIP_MAX_PAIRS);
  std::memset(pt, 0, sizeof(*pt));
  pt->npairs = npairs;
  for (int p = 0; p < npairs; ++p) {
    OP2HIP_REQUIRE(pairs[p].rarity > 0 && pairs[p].carity > 0, "sparsity: bad arity");
    OP2HIP_REQUIRE(inc_ptr && inc_ent && inc_ptr[p] && inc_ent[p],
                   "sparsity: missing row incidences of pair %d", p);
    pt->cmap[p] = pairs[p].cmap;
    pt->inc_ptr[p] = inc_ptr[p];
    pt->inc_ent[p] = inc_ent[p];
    pt->rarity[p] = pairs[p].rarity;
    pt->carity[p] = pairs[p].carity;
  }
  return OP2HIP_OK;
}

}  // namespace op2hip

using namespace op2hip;

extern "C" {

int op2hip_abi_version(void) { return OP2HIP_ABI_VERSION; }

const char *op2hip_last_error(void) { return g_error; }

int op2hip_init(int32_t device) {
  int n = 0;
  OP2HIP_CHECK(hipGetDeviceCount(&n));
  OP2HIP_REQUIRE(device >= 0 && device < n, "device %d not present (%d devices)", device, n);
  OP2HIP_CHECK(hipSetDevice(device));
  hipDeviceProp_t prop;
  OP2HIP_CHECK(hipGetDeviceProperties(&prop, device));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
    set_error("device %d is %s; this library is built for gfx950 only", device,
              prop.gcnArchName);
    return OP2HIP_ERR_ARG;
  }
  // the no-allocation entry point's scratch address on this device
  void *scratch = nullptr;
  return noalloc_scratch_address(device, &scratch);
}

int op2hip_scan_workspace(int64_t n, size_t *bytes) {
  OP2HIP_REQUIRE(bytes, "scan_workspace: NULL");
  *bytes = scan_bytes(n);
  return OP2HIP_OK;
}

int op2hip_exclusive_scan(int64_t n, const int32_t *in, int32_t *out, int64_t *total_host,
                          void *work, size_t work_bytes, void *stream) {
  return scan(n, in, out, total_host, work, work_bytes, stream_of(stream));
}

int op2hip_transpose_workspace(int32_t nrows, size_t *bytes) {
  OP2HIP_REQUIRE(bytes && nrows >= 0, "transpose_workspace: bad arguments");
  *bytes = transpose_bytes(nrows);
  return OP2HIP_OK;
}

int op2hip_transpose_count(int32_t nrows, const int32_t *map, int32_t arity,
                           int32_t iter_size, int32_t *inc_ptr, int64_t *total_host,
                           void *work, size_t work_bytes, void *stream) {
  hipStream_t s = stream_of(stream);
  OP2HIP_REQUIRE(nrows >= 0 && arity > 0 && iter_size >= 0 && inc_ptr,
                 "transpose_count: bad arguments");
  OP2HIP_REQUIRE(iter_size == 0 || map, "transpose_count: NULL map");
  if (work_bytes < transpose_bytes(nrows)) {
    set_error("transpose: workspace %zu < %zu", work_bytes, transpose_bytes(nrows));
    return OP2HIP_ERR_WORKSPACE;
  }
  char *w = static_cast<char *>(work);
  const size_t a = ((size_t)std::max(nrows, 1) * 4 + 255) & ~(size_t)255;
  int32_t *counts = reinterpret_cast<int32_t *>(w);
  int32_t *bad = reinterpret_cast<int32_t *>(w + a);
  void *scan_work = w + a + 256;
  int rc;
  if ((rc = fill(counts, nrows, 0, s))) return rc;
  if ((rc = fill(bad, 1, 0, s))) return rc;
  const int64_t nent = (int64_t)iter_size * arity;
  if (nent > 0) {
    transpose_count<<<std::min<uint32_t>(ceil_div(nent, 256), 16384), 256, 0, s>>>(
        nrows, map, nent, counts, bad);
    OP2HIP_LAUNCHED();
  }
  if ((rc = scan(nrows, counts, inc_ptr, total_host, scan_work, work_bytes - a - 256, s)))
    return rc;
  int32_t nbad = 0;
  OP2HIP_CHECK(hipMemcpyAsync(&nbad, bad, sizeof(nbad), hipMemcpyDeviceToHost, s));
  OP2HIP_CHECK(hipStreamSynchronize(s));
  if (nbad) {
    set_error("transpose: %d map entries >= %d (target set size)", nbad, nrows);
    return OP2HIP_ERR_ARG;
  }
  return OP2HIP_OK;
}

int op2hip_transpose_fill(int32_t nrows, const int32_t *map, int32_t arity,
                          int32_t iter_size, const int32_t *inc_ptr, int32_t *inc_ent,
                          void *work, size_t work_bytes, void *stream) {
  hipStream_t s = stream_of(stream);
  OP2HIP_REQUIRE(nrows >= 0 && arity > 0 && iter_size >= 0 && inc_ptr,
                 "transpose_fill: bad arguments");
  if (work_bytes < transpose_bytes(nrows)) {
    set_error("transpose: workspace %zu < %zu", work_bytes, transpose_bytes(nrows));
    return OP2HIP_ERR_WORKSPACE;
  }
  int32_t *cursor = static_cast<int32_t *>(work);
  if (nrows > 0)
    OP2HIP_CHECK(hipMemcpyAsync(cursor, inc_ptr, sizeof(int32_t) * nrows,
                                hipMemcpyDeviceToDevice, s));
  const int64_t nent = (int64_t)iter_size * arity;
  if (nent > 0) {
    transpose_fill<<<std::min<uint32_t>(ceil_div(nent, 256), 16384), 256, 0, s>>>(
        nrows, map, nent, cursor, inc_ent);
    OP2HIP_LAUNCHED();
  }
  if (nrows > 0) {
    transpose_sort<<<std::min<uint32_t>(ceil_div(nrows, 256), 16384), 256, 0, s>>>(
        nrows, inc_ptr, inc_ent);
    OP2HIP_LAUNCHED();
  }
  return OP2HIP_OK;
}

int op2hip_sparsity_count(int32_t nrows, int32_t ncols, const op2hip_map_pair *pairs,
                          int32_t npairs, const int32_t *const *inc_ptr,
                          const int32_t *const *inc_ent, int32_t set_diag,
                          int32_t *rowptr, int64_t *nnz_host, void *work,
                          size_t work_bytes, void *stream) {
  hipStream_t s = stream_of(stream);
  OP2HIP_REQUIRE(nrows >= 0 && ncols >= 0 && rowptr, "sparsity_count: bad arguments");
  PairTable pt;
  int rc;
  if ((rc = make_pairs(pairs, npairs, inc_ptr, inc_ent, &pt))) return rc;
  const size_t a = ((size_t)std::max(nrows, 1) * 4 + 255) & ~(size_t)255;
  if (work_bytes < a + scan_bytes(nrows)) {
    set_error("sparsity: workspace %zu < %zu", work_bytes, a + scan_bytes(nrows));
    return OP2HIP_ERR_WORKSPACE;
  }
  int32_t *counts = static_cast<int32_t *>(work);
  if (nrows > 0) {
    sparsity_rows<false><<<std::min<int32_t>(nrows, 65536), 64, 0, s>>>(
        nrows, ncols, pt, set_diag, nullptr, counts);
    OP2HIP_LAUNCHED();
  }
  return scan(nrows, counts, rowptr, nnz_host, static_cast<char *>(work) + a,
              work_bytes - a, s);
}

int op2hip_sparsity_fill(int32_t nrows, int32_t ncols, const op2hip_map_pair *pairs,
                         int32_t npairs, const int32_t *const *inc_ptr,
                         const int32_t *const *inc_ent, int32_t set_diag,
                         const int32_t *rowptr, int32_t *colidx, void *stream) {
  hipStream_t s = stream_of(stream);
  OP2HIP_REQUIRE(nrows >= 0 && ncols >= 0 && rowptr, "sparsity_fill: bad arguments");
  PairTable pt;
  int rc;
  if ((rc = make_pairs(pairs, npairs, inc_ptr, inc_ent, &pt))) return rc;
  if (nrows > 0) {
    sparsity_rows<true><<<std::min<int32_t>(nrows, 65536), 64, 0, s>>>(
        nrows, ncols, pt, set_diag, rowptr, colidx);
    OP2HIP_LAUNCHED();
  }
  return OP2HIP_OK;
}

int op2hip_pattern_expand(int32_t nrows, const int32_t *rowptr, const int32_t *colidx,
                          int32_t rdim, int32_t cdim, int32_t *rowptr_out,
                          int32_t *colidx_out, void *stream) {
  hipStream_t s = stream_of(stream);
  OP2HIP_REQUIRE(nrows >= 0 && rdim > 0 && cdim > 0, "pattern_expand: bad arguments");
  expand_rowptr<<<std::min<uint32_t>(ceil_div((int64_t)nrows + 1, 256), 16384), 256, 0, s>>>(
      nrows, rowptr, rdim, cdim, rowptr_out);
  OP2HIP_LAUNCHED();
  if (nrows > 0) {
    expand_colidx<<<std::min<uint32_t>(ceil_div(nrows, 256), 16384), 256, 0, s>>>(
        nrows, rowptr, colidx, rdim, cdim, rowptr_out, colidx_out);
    OP2HIP_LAUNCHED();
  }
  return OP2HIP_OK;
}

int op2hip_colour_greedy(int32_t nelem, int32_t nmaps, const int32_t *const *maps,
                         const int32_t *arity, const int32_t *target_id,
                         const int32_t *target_size, int32_t ntargets,
                         int32_t *colour, int32_t *ncolours) {
  OP2HIP_REQUIRE(nelem >= 0 && nmaps >= 0 && ntargets >= 0 && colour && ncolours,
                 "colour_greedy: bad arguments");
  for (int m = 0; m < nmaps; ++m)
    OP2HIP_REQUIRE(target_id[m] >= 0 && target_id[m] < ntargets,
                   "colour_greedy: map %d has target %d", m, target_id[m]);
  std::vector<std::vector<uint32_t>> masks(ntargets);
  for (int t = 0; t < ntargets; ++t) masks[t].assign(std::max(target_size[t], 1), 0u);
  std::fill(colour, colour + nelem, -1);
  int32_t left = nelem, base = 0, highest = 0;
  while (left > 0) {
    for (auto &m : masks) std::fill(m.begin(), m.end(), 0u);
    for (int32_t e = 0; e < nelem; ++e) {
      if (colour[e] >= 0) continue;
      uint32_t used = 0;
      for (int m = 0; m < nmaps; ++m) {
        const int32_t *row = maps[m] + (int64_t)e * arity[m];
        for (int i = 0; i < arity[m]; ++i)
          if (row[i] >= 0) {
            if (row[i] >= target_size[target_id[m]]) {
              set_error("colour_greedy: map %d entry %d out of range", m, row[i]);
              return OP2HIP_ERR_ARG;
            }
            used |= masks[target_id[m]][row[i]];
          }
      }
      if (used == 0xFFFFFFFFu) continue;  // try again in the next pass
      const int c = __builtin_ctz(~used);
      colour[e] = base + c;
      highest = std::max(highest, base + c + 1);
      --left;
      for (int m = 0; m < nmaps; ++m) {
        const int32_t *row = maps[m] + (int64_t)e * arity[m];
        for (int i = 0; i < arity[m]; ++i)
          if (row[i] >= 0) masks[target_id[m]][row[i]] |= (1u << c);
      }
    }
    base += 32;
  }
  *ncolours = highest;
  return OP2HIP_OK;
}

int op2hip_plan_tiles(int32_t start, int32_t end, int32_t nmaps, const int32_t *const *maps,
                      const int32_t *arity, const int32_t *target_size, int32_t tile_size,
                      int32_t node_cap, int32_t *ntiles_out, int32_t *tile_ptr,
                      int32_t *tile_elems, int32_t *const *node_ptr, int32_t *const *nodes,
                      int32_t *const *nexcl, uint16_t *const *local, int32_t *max_nodes) {
  OP2HIP_REQUIRE(end >= start && start >= 0 && nmaps >= 1 && nmaps <= 2 && tile_size > 0 &&
                     ntiles_out && tile_ptr && tile_elems && node_ptr && nodes && nexcl && local &&
                     max_nodes,
                 "plan_tiles: bad arguments");
  OP2HIP_REQUIRE(node_cap <= 0 || node_cap >= arity[0],
                 "plan_tiles: a node cap of %d cannot hold one element (arity %d)", node_cap,
                 arity[0]);
  for (int m = 0; m < nmaps; ++m)
    OP2HIP_REQUIRE(maps[m] && arity[m] > 0 && (int64_t)tile_size * arity[m] <= 65535,
                   "plan_tiles: map %d: arity %d, tile size %d (local ids are 16-bit)", m, arity[m],
                   tile_size);
  const int32_t n = end - start;
  // elements around each node of map 0 (the adjacency the regions grow over)
  const int32_t a0 = arity[0], nt0 = std::max(target_size[0], 1);
  std::vector<int32_t> adj_ptr(nt0 + 1, 0), adj;
  for (int32_t k = 0; k < n; ++k)
    for (int i = 0; i < a0; ++i) {
      const int32_t v = maps[0][(int64_t)(start + k) * a0 + i];
      OP2HIP_REQUIRE(v >= 0 && v < target_size[0], "plan_tiles: map 0 entry %d out of range", v);
      ++adj_ptr[v + 1];
    }
  for (int32_t v = 0; v < nt0; ++v) adj_ptr[v + 1] += adj_ptr[v];
  adj.resize(adj_ptr[nt0]);
  {
    std::vector<int32_t> fillp(adj_ptr.begin(), adj_ptr.end() - 1);
    for (int32_t k = 0; k < n; ++k)
      for (int i = 0; i < a0; ++i) adj[fillp[maps[0][(int64_t)(start + k) * a0 + i]]++] = k;
  }
  // regions: breadth first from the lowest unassigned element; elements queued
  // but not taken when the tile is full are released for later tiles
  // (and, with node_cap > 0, at most node_cap distinct nodes of map 0)
  std::vector<uint8_t> taken(n, 0), queued(n, 0);
  std::vector<int32_t> queue, mark(nt0, -1);
  queue.reserve(n);
  int32_t seed = 0, pos = 0, ntiles = 0;
  tile_ptr[0] = 0;
  while (pos < n) {
    const int32_t t = ntiles++;
    queue.clear();
    size_t head = 0;
    const int32_t want = std::min(tile_size, n - pos);
    int32_t got = 0, tile_nodes = 0;
    while (got < want) {
      if (head == queue.size()) {
        if (got > 0 && node_cap > 0) break;  // region exhausted: start a new tile
        while (taken[seed] || queued[seed]) ++seed;
        queued[seed] = 1;
        queue.push_back(seed);
      }
      const int32_t k = queue[head++];
      if (node_cap > 0) {
        int32_t fresh = 0;
        for (int i = 0; i < a0; ++i) fresh += mark[maps[0][(int64_t)(start + k) * a0 + i]] != t;
        if (got > 0 && tile_nodes + fresh > node_cap) {
          queued[k] = 0;
          continue;
        }
        for (int i = 0; i < a0; ++i) {
          int32_t &mk = mark[maps[0][(int64_t)(start + k) * a0 + i]];
          tile_nodes += mk != t;
          mk = t;
        }
      }
      taken[k] = 1;
      tile_elems[pos++] = start + k;
      ++got;
      for (int i = 0; i < a0 && got < want; ++i) {
        const int32_t v = maps[0][(int64_t)(start + k) * a0 + i];
        for (int32_t q = adj_ptr[v]; q < adj_ptr[v + 1]; ++q) {
          const int32_t k2 = adj[q];
          if (!taken[k2] && !queued[k2]) {
            queued[k2] = 1;
            queue.push_back(k2);
          }
        }
      }
    }
    for (size_t q = head; q < queue.size(); ++q) queued[queue[q]] = 0;
    tile_ptr[t + 1] = pos;
  }
  *ntiles_out = ntiles;
  // per map: the distinct nodes of each tile, those of no other tile first
  for (int m = 0; m < nmaps; ++m) {
    const int32_t am = arity[m], ntm = std::max(target_size[m], 1);
    std::vector<int32_t> ntile(ntm, 0), stamp(ntm, -1), slot(ntm, 0);
    for (int32_t t = 0; t < ntiles; ++t)
      for (int32_t p = tile_ptr[t]; p < tile_ptr[t + 1]; ++p)
        for (int i = 0; i < am; ++i) {
          const int32_t v = maps[m][(int64_t)tile_elems[p] * am + i];
          OP2HIP_REQUIRE(v >= 0 && v < target_size[m], "plan_tiles: map %d entry %d out of range",
                         m, v);
          if (stamp[v] != t) {
            stamp[v] = t;
            ++ntile[v];
          }
        }
    std::fill(stamp.begin(), stamp.end(), -1);
    std::vector<int32_t> own, shared;
    int32_t total = 0, mx = 0;
    node_ptr[m][0] = 0;
    for (int32_t t = 0; t < ntiles; ++t) {
      own.clear();
      shared.clear();
      for (int32_t p = tile_ptr[t]; p < tile_ptr[t + 1]; ++p)
        for (int i = 0; i < am; ++i) {
          const int32_t v = maps[m][(int64_t)tile_elems[p] * am + i];
          if (stamp[v] != t) {
            stamp[v] = t;
            (ntile[v] == 1 ? own : shared).push_back(v);
          }
        }
      std::sort(own.begin(), own.end());
      std::sort(shared.begin(), shared.end());
      int32_t c = 0;
      for (int32_t v : own) { nodes[m][total + c] = v; slot[v] = c++; }
      for (int32_t v : shared) { nodes[m][total + c] = v; slot[v] = c++; }
      nexcl[m][t] = (int32_t)own.size();
      for (int32_t p = tile_ptr[t]; p < tile_ptr[t + 1]; ++p)
        for (int i = 0; i < am; ++i)
          local[m][(int64_t)p * am + i] =
              static_cast<uint16_t>(slot[maps[m][(int64_t)tile_elems[p] * am + i]]);
      total += c;
      mx = std::max(mx, c);
      node_ptr[m][t + 1] = total;
    }
    max_nodes[m] = mx;
  }
  return OP2HIP_OK;
}

int op2hip_plan_tile_reduce(int32_t ntiles, const int32_t *node_ptr, const int32_t *nodes,
                            const int32_t *nexcl, int32_t nnodes, int32_t *shared_off,
                            int32_t *nshared, int32_t *shared_nodes, int32_t *shared_ptr,
                            int32_t *shared_slots) {
  OP2HIP_REQUIRE(ntiles >= 0 && nnodes >= 0 && node_ptr && nodes && nexcl && shared_off &&
                     nshared && shared_nodes && shared_ptr && shared_slots,
                 "plan_tile_reduce: bad arguments");
  // partial slots: the shared nodes of each tile, tiles in order
  shared_off[0] = 0;
  for (int32_t b = 0; b < ntiles; ++b) {
    const int32_t cnt = node_ptr[b + 1] - node_ptr[b];
    OP2HIP_REQUIRE(cnt >= 0 && nexcl[b] >= 0 && nexcl[b] <= cnt,
                   "plan_tile_reduce: tile %d has %d nodes, %d exclusive", b, cnt, nexcl[b]);
    shared_off[b + 1] = shared_off[b] + (cnt - nexcl[b]);
  }
  const int32_t nslots = shared_off[ntiles];
  // counting sort of the slots by node id (stable: tile order within a node)
  std::vector<int32_t> count((size_t)nnodes + 1, 0);
  for (int32_t b = 0; b < ntiles; ++b)
    for (int32_t k = node_ptr[b] + nexcl[b]; k < node_ptr[b + 1]; ++k) {
      const int32_t v = nodes[k];
      OP2HIP_REQUIRE(v >= 0 && v < nnodes, "plan_tile_reduce: node %d out of range", v);
      ++count[(size_t)v + 1];
    }
  int32_t ns = 0;
  std::vector<int32_t> first((size_t)nnodes, -1);
  for (int32_t v = 0; v < nnodes; ++v)
    if (count[(size_t)v + 1] > 0) {
      first[v] = ns;
      shared_nodes[ns] = v;
      ++ns;
    }
  shared_ptr[0] = 0;
  for (int32_t s = 0; s < ns; ++s) shared_ptr[s + 1] = shared_ptr[s] + count[(size_t)shared_nodes[s] + 1];
  std::vector<int32_t> fill(shared_ptr, shared_ptr + ns);
  for (int32_t b = 0; b < ntiles; ++b)
    for (int32_t k = node_ptr[b] + nexcl[b]; k < node_ptr[b + 1]; ++k) {
      const int32_t s = first[nodes[k]];
      shared_slots[fill[s]++] = shared_off[b] + (k - node_ptr[b] - nexcl[b]);
    }
  OP2HIP_REQUIRE(shared_ptr[ns] == nslots, "plan_tile_reduce: internal slot count mismatch");
  *nshared = ns;
  return OP2HIP_OK;
}

int op2hip_plan_row_batches(int32_t nrows, const int32_t *inc_ptr, const int32_t *rowptr,
                            int32_t max_inc, int32_t max_nnz, int32_t *batch_rows,
                            int32_t *nbatches, int32_t *max_batch_inc,
                            int32_t *max_batch_nnz, int32_t *max_batch_rows) {
  OP2HIP_REQUIRE(nrows >= 0 && inc_ptr && rowptr && batch_rows && nbatches &&
                     max_inc > 0 && max_nnz > 0,
                 "plan_row_batches: bad arguments");
  int32_t nb = 0, mi = 0, mn = 0, mr = 0;
  int32_t r = 0;
  batch_rows[0] = 0;
  while (r < nrows) {
    const int32_t r0 = r;
    int32_t inc = 0, nnz = 0;
    while (r < nrows) {
      const int32_t di = inc_ptr[r + 1] - inc_ptr[r];
      const int32_t dn = rowptr[r + 1] - rowptr[r];
      if (r > r0 && (inc + di > max_inc || nnz + dn > max_nnz)) break;
      inc += di;
      nnz += dn;
      ++r;
    }
    batch_rows[++nb] = r;
    mi = std::max(mi, inc);
    mn = std::max(mn, nnz);
    mr = std::max(mr, r - r0);
  }
  *nbatches = nb;
  if (max_batch_inc) *max_batch_inc = mi;
  if (max_batch_nnz) *max_batch_nnz = mn;
  if (max_batch_rows) *max_batch_rows = mr;
  return OP2HIP_OK;
}

}  // extern "C"

#include "rowgather.hpp"

namespace op2hip {

template <int kCar>
int plan_rowlist(int32_t nrows, const int32_t *rowptr, const int32_t *colidx,
                 const int32_t *inc_ptr, const int32_t *inc_ent, const int32_t *cmap,
                 int32_t rarity, int32_t max_inc, int32_t max_nnz, int32_t *row_order,
                 int32_t *nbatches, BatchDesc *desc, int32_t *ent, SlotEntry<kCar> *table,
                 int64_t *missing, int32_t *mi_out, int32_t *mn_out, int32_t *mr_out) {
  // rows by their first (lowest) incident element: incidences are sorted
  // by element within a row, so that is the row's first incidence
  std::vector<int32_t> idx(nrows);
  std::vector<int32_t> key(nrows);
  for (int32_t r = 0; r < nrows; ++r) {
    idx[r] = r;
    key[r] = inc_ptr[r + 1] > inc_ptr[r] ? inc_ent[inc_ptr[r]] / rarity : INT32_MAX;
  }
  std::stable_sort(idx.begin(), idx.end(), [&](int32_t x, int32_t y) { return key[x] < key[y]; });
  int32_t nb = 0, mi = 0, mn = 0, mr = 0;
  int64_t pos = 0, miss = 0;
  int32_t k = 0;
  while (k < nrows) {
    const int32_t b0 = k;
    int32_t inc = 0, nnz = 0;
    while (k < nrows) {
      const int32_t r = idx[k];
      const int32_t di = inc_ptr[r + 1] - inc_ptr[r], dn = rowptr[r + 1] - rowptr[r];
      if (k > b0 && (inc + di > max_inc || nnz + dn > max_nnz)) break;
      inc += di;
      nnz += dn;
      ++k;
    }
    if (nnz > 65535 || k - b0 > 65535) {
      set_error("row list: a batch of %d blocks / %d rows exceeds 16-bit offsets", nnz, k - b0);
      return OP2HIP_ERR_ARG;
    }
    BatchDesc d;
    d.r0 = b0;
    d.nr = k - b0;
    d.z0 = 0;
    d.nz = nnz;
    d.i0 = (int32_t)pos;
    d.ni = inc;
    d.pad0 = d.pad1 = 0;
    desc[nb++] = d;
    int32_t zoff = 0;
    for (int32_t q = b0; q < k; ++q) {
      const int32_t r = idx[q];
      row_order[2 * q] = r;
      row_order[2 * q + 1] = zoff;
      const int32_t zlo = rowptr[r], zhi = rowptr[r + 1];
      for (int32_t t = inc_ptr[r]; t < inc_ptr[r + 1]; ++t) {
        const int32_t e = inc_ent[t] / rarity;
        SlotEntry<kCar> s;
        s.row_off = static_cast<uint16_t>(zoff);
        s.lr = static_cast<uint16_t>(q - b0);
        for (int j = 0; j < kCar; ++j) {
          const int32_t c = cmap[(int64_t)e * kCar + j];
          int32_t p = -1;
          if (c >= 0) {
            const int32_t *hit = std::lower_bound(colidx + zlo, colidx + zhi, c);
            if (hit != colidx + zhi && *hit == c) p = int32_t(hit - colidx);
            else ++miss;
          }
          s.slot[j] = (p < 0 || p - zlo > 254) ? uint8_t(255) : static_cast<uint8_t>(p - zlo);
        }
        ent[pos] = inc_ent[t];
        table[pos++] = s;
      }
      zoff += zhi - zlo;
    }
    mi = std::max(mi, inc);
    mn = std::max(mn, nnz);
    mr = std::max(mr, k - b0);
  }
  *nbatches = nb;
  *missing = miss;
  *mi_out = mi;
  *mn_out = mn;
  *mr_out = mr;
  return OP2HIP_OK;
}

template <int kCar>
int plan_cells(int32_t nbatches, const BatchDesc *D, const SlotEntry<kCar> *table, int32_t rarity,
               const int32_t *inc_ent, int32_t *cell_ptr, int32_t *cells, uint8_t *gcount,
               uint16_t *order, int32_t *contrib_ptr, uint16_t *contrib) {
  std::vector<int32_t> es, uniq, cnt, pos, first;
  std::vector<int32_t> idx;
  int64_t cp = 0, gp = 0;
  cell_ptr[0] = 0;
  contrib_ptr[0] = 0;
  for (int32_t b = 0; b < nbatches; ++b) {
    const BatchDesc &d = D[b];
    es.resize(d.ni);
    for (int32_t t = 0; t < d.ni; ++t) es[t] = inc_ent[d.i0 + t] / rarity;
    uniq.assign(es.begin(), es.end());
    std::sort(uniq.begin(), uniq.end());
    uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
    if (uniq.size() > 256) {
      set_error("cell gather: batch %d has %zu elements (at most 256)", b, uniq.size());
      return OP2HIP_ERR_ARG;
    }
    for (size_t k = 0; k < uniq.size(); ++k) cells[cp + (int64_t)k] = uniq[k];
    cnt.assign(d.nz, 0);
    for (int32_t t = 0; t < d.ni; ++t)
      for (int j = 0; j < kCar; ++j)
        if (table[d.i0 + t].slot[j] != 255) ++cnt[table[d.i0 + t].row_off + table[d.i0 + t].slot[j]];
    idx.resize(d.nz);
    for (int32_t z = 0; z < d.nz; ++z) {
      if (cnt[z] > 255) {
        set_error("cell gather: a block takes %d element blocks (at most 255)", cnt[z]);
        return OP2HIP_ERR_ARG;
      }
      gcount[(int64_t)d.z0 + z] = static_cast<uint8_t>(cnt[z]);
      idx[z] = z;
    }
    std::stable_sort(idx.begin(), idx.end(), [&](int32_t x, int32_t y) { return cnt[x] > cnt[y]; });
    first.assign(d.nz, 0);
    int32_t acc = 0;
    for (int32_t w = 0; w < d.nz; ++w) {
      order[(int64_t)d.z0 + w] = static_cast<uint16_t>(idx[w]);
      first[idx[w]] = acc;
      acc += cnt[idx[w]];
    }
    pos.assign(first.begin(), first.end());
    for (int32_t t = 0; t < d.ni; ++t) {
      const int32_t ent = inc_ent[d.i0 + t];
      const int32_t e = ent / rarity, i = ent - e * rarity;
      const int32_t u = int32_t(std::lower_bound(uniq.begin(), uniq.end(), e) - uniq.begin());
      for (int j = 0; j < kCar; ++j) {
        const uint8_t s = table[d.i0 + t].slot[j];
        if (s == 255) continue;
        contrib[gp + pos[table[d.i0 + t].row_off + s]++] =
            static_cast<uint16_t>((u << 8) | (i << 4) | j);
      }
    }
    cp += (int64_t)uniq.size();
    gp += acc;
    if (cp > INT32_MAX || gp > INT32_MAX) {
      set_error("cell gather: plan exceeds 2^31 entries");
      return OP2HIP_ERR_OVERFLOW;
    }
    cell_ptr[b + 1] = (int32_t)cp;
    contrib_ptr[b + 1] = (int32_t)gp;
  }
  return OP2HIP_OK;
}

}  // namespace op2hip

extern "C" {

int op2hip_slot_table_bytes(int32_t carity, size_t *entry_bytes) {
  OP2HIP_REQUIRE(entry_bytes, "slot_table_bytes: NULL");
  switch (carity) {
    case 3: *entry_bytes = sizeof(SlotEntry<3>); return OP2HIP_OK;
    case 4: *entry_bytes = sizeof(SlotEntry<4>); return OP2HIP_OK;
    case 10: *entry_bytes = sizeof(SlotEntry<10>); return OP2HIP_OK;
    default:
      set_error("slot tables exist for column arities 3, 4 and 10 (got %d)", carity);
      return OP2HIP_ERR_ARG;
  }
}

int op2hip_build_slot_table(int32_t nbatches, const int32_t *batch_rows, const int32_t *rowptr,
                            const int32_t *colidx, const int32_t *inc_ptr, const int32_t *inc_ent,
                            const int32_t *cmap, int32_t rarity, int32_t carity, void *table,
                            void *batch_desc, int32_t *bad, void *stream) {
  hipStream_t s = stream_of(stream);
  OP2HIP_REQUIRE(nbatches >= 0 && batch_rows && rowptr && colidx && inc_ptr && inc_ent && cmap &&
                     table && batch_desc && rarity > 0 && bad,
                 "build_slot_table: bad arguments");
  int rc = fill(bad, 2, 0, s);
  if (rc) return rc;
  if (nbatches > 0) {
    auto *d = static_cast<BatchDesc *>(batch_desc);
    switch (carity) {
      case 3:
        build_slot_table<3><<<nbatches, 64, 0, s>>>(nbatches, batch_rows, rowptr, colidx, inc_ptr,
                                                     inc_ent, cmap, rarity,
                                                     static_cast<SlotEntry<3> *>(table), d, bad);
        break;
      case 4:
        build_slot_table<4><<<nbatches, 64, 0, s>>>(nbatches, batch_rows, rowptr, colidx, inc_ptr,
                                                     inc_ent, cmap, rarity,
                                                     static_cast<SlotEntry<4> *>(table), d, bad);
        break;
      case 10:
        build_slot_table<10><<<nbatches, 64, 0, s>>>(nbatches, batch_rows, rowptr, colidx, inc_ptr,
                                                      inc_ent, cmap, rarity,
                                                      static_cast<SlotEntry<10> *>(table), d, bad);
        break;
      default:
        set_error("slot tables exist for column arities 3, 4 and 10 (got %d)", carity);
        return OP2HIP_ERR_ARG;
    }
    OP2HIP_LAUNCHED();
  }
  return OP2HIP_OK;
}

int op2hip_plan_row_list(int32_t nrows, const int32_t *rowptr, const int32_t *colidx,
                         const int32_t *inc_ptr, const int32_t *inc_ent, const int32_t *cmap,
                         int32_t rarity, int32_t carity, int32_t max_inc, int32_t max_nnz,
                         int32_t *row_order, int32_t *nbatches, void *desc, int32_t *ent,
                         void *table, int64_t *missing, int32_t *max_inc_out,
                         int32_t *max_nnz_out, int32_t *max_rows_out) {
  OP2HIP_REQUIRE(nrows >= 0 && rowptr && colidx && inc_ptr && inc_ent && cmap && rarity > 0 &&
                     max_inc > 0 && max_nnz > 0 && row_order && nbatches && desc && ent && table &&
                     missing && max_inc_out && max_nnz_out && max_rows_out,
                 "plan_row_list: bad arguments");
  auto *D = static_cast<BatchDesc *>(desc);
  switch (carity) {
    case 3:
      return plan_rowlist<3>(nrows, rowptr, colidx, inc_ptr, inc_ent, cmap, rarity, max_inc,
                             max_nnz, row_order, nbatches, D, ent,
                             static_cast<SlotEntry<3> *>(table), missing, max_inc_out,
                             max_nnz_out, max_rows_out);
    case 4:
      return plan_rowlist<4>(nrows, rowptr, colidx, inc_ptr, inc_ent, cmap, rarity, max_inc,
                             max_nnz, row_order, nbatches, D, ent,
                             static_cast<SlotEntry<4> *>(table), missing, max_inc_out,
                             max_nnz_out, max_rows_out);
    case 10:
      return plan_rowlist<10>(nrows, rowptr, colidx, inc_ptr, inc_ent, cmap, rarity, max_inc,
                              max_nnz, row_order, nbatches, D, ent,
                              static_cast<SlotEntry<10> *>(table), missing, max_inc_out,
                              max_nnz_out, max_rows_out);
    default:
      set_error("row lists exist for column arities 3, 4 and 10 (got %d)", carity);
      return OP2HIP_ERR_ARG;
  }
}

int op2hip_plan_cell_gather(int32_t nbatches, const void *batch_desc, const void *table,
                            int32_t rarity, int32_t carity, const int32_t *inc_ent,
                            int32_t *cell_ptr, int32_t *cells, uint8_t *gcount, uint16_t *order,
                            int32_t *contrib_ptr, uint16_t *contrib) {
  OP2HIP_REQUIRE(nbatches >= 0 && batch_desc && table && inc_ent && cell_ptr && cells && gcount &&
                     order && contrib_ptr && contrib && rarity > 0 && rarity <= 16,
                 "plan_cell_gather: bad arguments");
  const auto *D = static_cast<const BatchDesc *>(batch_desc);
  switch (carity) {
    case 3:
      return plan_cells<3>(nbatches, D, static_cast<const SlotEntry<3> *>(table), rarity, inc_ent,
                           cell_ptr, cells, gcount, order, contrib_ptr, contrib);
    case 4:
      return plan_cells<4>(nbatches, D, static_cast<const SlotEntry<4> *>(table), rarity, inc_ent,
                           cell_ptr, cells, gcount, order, contrib_ptr, contrib);
    case 10:
      return plan_cells<10>(nbatches, D, static_cast<const SlotEntry<10> *>(table), rarity,
                            inc_ent, cell_ptr, cells, gcount, order, contrib_ptr, contrib);
    default:
      set_error("cell gather plans exist for column arities 3, 4 and 10 (got %d)", carity);
      return OP2HIP_ERR_ARG;
  }
}

int op2hip_gather_table_count(int32_t nbatches, int32_t batch_max_nnz, const void *batch_desc,
                              const void *table, int32_t carity, uint8_t *gcount,
                              int32_t *batch_total, int32_t *over, void *stream) {
  hipStream_t s = stream_of(stream);
  OP2HIP_REQUIRE(nbatches >= 0 && batch_desc && table && gcount && batch_total && over,
                 "gather_table_count: bad arguments");
  OP2HIP_REQUIRE(batch_max_nnz <= kGatherMaxBlocks,
                 "gather tables take batches of at most %d blocks (got %d)", kGatherMaxBlocks,
                 batch_max_nnz);
  int rc = fill(over, 1, 0, s);
  if (rc || nbatches == 0) return rc;
  const auto *d = static_cast<const BatchDesc *>(batch_desc);
  switch (carity) {
    case 3:
      gather_count_kernel<3><<<nbatches, 256, 0, s>>>(
          nbatches, d, static_cast<const SlotEntry<3> *>(table), gcount, batch_total, over);
      break;
    case 4:
      gather_count_kernel<4><<<nbatches, 256, 0, s>>>(
          nbatches, d, static_cast<const SlotEntry<4> *>(table), gcount, batch_total, over);
      break;
    case 10:
      gather_count_kernel<10><<<nbatches, 256, 0, s>>>(
          nbatches, d, static_cast<const SlotEntry<10> *>(table), gcount, batch_total, over);
      break;
    default:
      set_error("gather tables exist for column arities 3, 4 and 10 (got %d)", carity);
      return OP2HIP_ERR_ARG;
  }
  OP2HIP_LAUNCHED();
  return OP2HIP_OK;
}

int op2hip_gather_table_fill(int32_t nbatches, int32_t batch_max_nnz, void *batch_desc,
                             const void *table, int32_t carity, const int32_t *batch_off,
                             const int32_t *batch_total, const uint8_t *gcount, uint16_t *glist,
                             void *stream) {
  hipStream_t s = stream_of(stream);
  OP2HIP_REQUIRE(nbatches >= 0 && batch_desc && table && batch_off && batch_total && gcount &&
                     glist,
                 "gather_table_fill: bad arguments");
  OP2HIP_REQUIRE(batch_max_nnz <= kGatherMaxBlocks,
                 "gather tables take batches of at most %d blocks (got %d)", kGatherMaxBlocks,
                 batch_max_nnz);
  if (nbatches == 0) return OP2HIP_OK;
  auto *d = static_cast<BatchDesc *>(batch_desc);
  switch (carity) {
    case 3:
      gather_fill_kernel<3><<<nbatches, 256, 0, s>>>(nbatches, d,
                                                      static_cast<const SlotEntry<3> *>(table),
                                                      batch_off, batch_total, gcount, glist);
      break;
    case 4:
      gather_fill_kernel<4><<<nbatches, 256, 0, s>>>(nbatches, d,
                                                      static_cast<const SlotEntry<4> *>(table),
                                                      batch_off, batch_total, gcount, glist);
      break;
    case 10:
      gather_fill_kernel<10><<<nbatches, 256, 0, s>>>(nbatches, d,
                                                       static_cast<const SlotEntry<10> *>(table),
                                                       batch_off, batch_total, gcount, glist);
      break;
    default:
      set_error("gather tables exist for column arities 3, 4 and 10 (got %d)", carity);
      return OP2HIP_ERR_ARG;
  }
  OP2HIP_LAUNCHED();
  return OP2HIP_OK;
}

}  // extern "C"
