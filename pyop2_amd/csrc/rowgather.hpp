// Owner-computes ("row gather") Mat INC assembly.
//
// Every workgroup owns one batch of consecutive block rows of the matrix.
// It walks the row incidences (element e, local row i) of its rows -- the
// transpose of the row map, sorted by element -- computes the row block i of
// the element matrix of e, looks each column block up in the row's sorted
// colidx (staged in LDS) and accumulates it into an LDS copy of the rows'
// values.  The rows then receive one coalesced read-modify-write.  No two
// workgroups touch the same row, so there are no global atomics and the
// values array streams through HBM once.  An optional Dat INC on the same
// rows (the residual R = A u of the same loop) is accumulated alongside.
#pragma once

#include "launch.hpp"

namespace op2hip {

// dst[k] += src[k] for k < n (dst in global memory, src in LDS), with kDepth
// independent global loads in flight per thread before the first add: a
// plain loop would wait for every load in turn and serialise the HBM latency.
template <typename T, int kDepth = 8>
__device__ __forceinline__ void flush_add(T *__restrict__ dst, const T *src, int32_t n) {
  const int32_t step = blockDim.x * kDepth;
  for (int32_t k0 = threadIdx.x; k0 < n; k0 += step) {
    T v[kDepth];
#pragma unroll
    for (int u = 0; u < kDepth; ++u) {
      const int32_t k = k0 + u * (int32_t)blockDim.x;
      v[u] = k < n ? dst[k] : T(0);
    }
#pragma unroll
    for (int u = 0; u < kDepth; ++u) {
      const int32_t k = k0 + u * (int32_t)blockDim.x;
      if (k < n) dst[k] = v[u] + src[k];
    }
  }
}

// dst[k] = src[k] for k < n (global -> LDS), kDepth loads in flight.
template <typename T, int kDepth = 4>
__device__ __forceinline__ void stage_copy(T *dst, const T *__restrict__ src, int32_t n) {
  const int32_t step = blockDim.x * kDepth;
  for (int32_t k0 = threadIdx.x; k0 < n; k0 += step) {
    T v[kDepth];
#pragma unroll
    for (int u = 0; u < kDepth; ++u) {
      const int32_t k = k0 + u * (int32_t)blockDim.x;
      v[u] = k < n ? src[k] : T(0);
    }
#pragma unroll
    for (int u = 0; u < kDepth; ++u) {
      const int32_t k = k0 + u * (int32_t)blockDim.x;
      if (k < n) dst[k] = v[u];
    }
  }
}

// Elem must provide:
//   static constexpr int kRar, kCar, kRd, kCd;      (arities, block shape)
//   static constexpr bool kHasRowDat;               (fused Dat INC on rows)
//   using T;                                        (value type)
//   __device__ void row(int32_t e, int i, Emit&&)   (calls emit(j, col, blk, rowdat))
template <class Elem>
struct RowGatherLayout {
  using T = double;   // accumulators are double for float and double elements
  static constexpr int kBs = Elem::kRd * Elem::kCd;
  static size_t bytes(int32_t max_nnz, int32_t max_rows) {
    size_t b = (size_t)max_nnz * kBs * sizeof(T);          // acc
    b += (size_t)(Elem::kHasRowDat ? max_rows * Elem::kRd : 0) * sizeof(T);
    b = (b + 15) & ~(size_t)15;
    b += (size_t)max_nnz * sizeof(int32_t);                 // colidx
    b += (size_t)(max_rows + 1) * 2 * sizeof(int32_t);      // row / incidence offsets
    return b;
  }
};

// Accumulator type of an element: float elements accumulate in double (LDS
// float atomics are slow on gfx950, and double halves the rounding error).
template <typename T>
struct AccumOf {
  using type = double;
};

template <class Elem, int kPre>
__global__ void __launch_bounds__(256)
rowgather_kernel(int32_t start, int32_t end, Elem elem, op2hip_mat A,
                 const int32_t *__restrict__ batch_rows, int32_t nbatches,
                 typename Elem::T *__restrict__ rowdat) {
  using T = typename Elem::T;
  using Acc = typename AccumOf<T>::type;
  constexpr int kBs = Elem::kRd * Elem::kCd;
  extern __shared__ __align__(16) unsigned char smem[];
  const int32_t b = xcd_contiguous(blockIdx.x, nbatches);
  const int32_t r0 = batch_rows[b], r1 = batch_rows[b + 1];
  const int32_t nr = r1 - r0;
  const int32_t z0 = A.rowptr[r0], nz = A.rowptr[r1] - z0;
  const int32_t i0 = A.inc_ptr[r0], ni = A.inc_ptr[r1] - i0;
  T *__restrict__ vals = static_cast<T *>(A.values) + (size_t)z0 * kBs;
  const int32_t nval = nz * kBs;

  // The rows' current values are fetched first, into registers: their HBM
  // latency then overlaps the whole element phase instead of following it.
  T old[kPre];
#pragma unroll
  for (int u = 0; u < kPre; ++u) {
    const int32_t k = threadIdx.x + u * (int32_t)blockDim.x;
    old[u] = k < nval ? vals[k] : T(0);
  }

  Acc *acc = reinterpret_cast<Acc *>(smem);
  Acc *racc = acc + (size_t)nz * kBs;
  size_t off = ((size_t)nz * kBs + (Elem::kHasRowDat ? (size_t)nr * Elem::kRd : 0)) * sizeof(Acc);
  off = (off + 15) & ~(size_t)15;
  int32_t *scol = reinterpret_cast<int32_t *>(smem + off);
  int32_t *srow = scol + nz;          // nr+1 row offsets into scol
  int32_t *sinc = srow + nr + 1;      // nr+1 incidence offsets

  for (int32_t k = threadIdx.x; k < nval; k += blockDim.x) acc[k] = Acc(0);
  if (Elem::kHasRowDat)
    for (int32_t k = threadIdx.x; k < nr * Elem::kRd; k += blockDim.x) racc[k] = Acc(0);
  stage_copy(scol, A.colidx + z0, nz);
  for (int32_t k = threadIdx.x; k <= nr; k += blockDim.x) {
    srow[k] = A.rowptr[r0 + k] - z0;
    sinc[k] = A.inc_ptr[r0 + k] - i0;
  }
  __syncthreads();

  for (int32_t t = threadIdx.x; t < ni; t += blockDim.x) {
    const int32_t ent = A.inc_ent[i0 + t];
    const int32_t e = ent / Elem::kRar;
    if (e < start || e >= end) continue;
    const int i = ent - e * Elem::kRar;
    // local row of incidence t: last lr with sinc[lr] <= t
    int32_t lo = 0, hi = nr - 1;
    while (lo < hi) {
      const int32_t mid = (lo + hi + 1) >> 1;
      if (sinc[mid] <= t) lo = mid; else hi = mid - 1;
    }
    const int32_t lr = lo;
    const int32_t c0 = srow[lr], c1 = srow[lr + 1];
    T rsum[Elem::kRd];
#pragma unroll
    for (int a = 0; a < Elem::kRd; ++a) rsum[a] = T(0);
    elem.row(e, i, [&](int j, int32_t col, const T *blk) {
      if (col < 0) return;
      const int32_t slot = find_slot(scol, c0, c1, col);
      if (slot < 0) {
        atomicAdd(A.status, 1);
        return;
      }
#pragma unroll
      for (int k = 0; k < kBs; ++k) atomic_add(&acc[slot * kBs + k], Acc(blk[k]));
    }, rsum);
    if (Elem::kHasRowDat) {
#pragma unroll
      for (int a = 0; a < Elem::kRd; ++a) atomic_add(&racc[lr * Elem::kRd + a], Acc(rsum[a]));
    }
  }
  __syncthreads();

#pragma unroll
  for (int u = 0; u < kPre; ++u) {
    const int32_t k = threadIdx.x + u * (int32_t)blockDim.x;
    if (k < nval) vals[k] = old[u] + T(acc[k]);
  }
  // a batch larger than kPre * blockDim values (one very long row) finishes here
  for (int32_t k = threadIdx.x + kPre * (int32_t)blockDim.x; k < nval; k += blockDim.x)
    vals[k] += T(acc[k]);
  if (Elem::kHasRowDat) {
    T *rd = rowdat + (size_t)r0 * Elem::kRd;
    for (int32_t k = threadIdx.x; k < nr * Elem::kRd; k += blockDim.x) rd[k] += T(racc[k]);
  }
}

template <class Elem>
int launch_rowgather(int32_t start, int32_t end, const Elem &elem, const op2hip_mat *A,
                     const op2hip_exec *ex, typename Elem::T *rowdat, hipStream_t s) {
  OP2HIP_REQUIRE(ex && ex->batch_rows && ex->nbatches >= 0,
                 "row-gather launch without a row-batch plan");
  OP2HIP_REQUIRE(A->inc_ptr && A->inc_ent, "row-gather launch without row incidences");
  OP2HIP_REQUIRE(!ex->subset, "row-gather assembly does not iterate Subsets");
  OP2HIP_REQUIRE(A->rdim == Elem::kRd && A->cdim == Elem::kCd,
                 "matrix block %dx%d, kernel assembles %dx%d", A->rdim, A->cdim,
                 Elem::kRd, Elem::kCd);
  if (end <= start || ex->nbatches == 0) return OP2HIP_OK;
  const size_t lds = RowGatherLayout<Elem>::bytes(ex->batch_max_nnz, ex->batch_max_rows);
  OP2HIP_REQUIRE(lds <= 64 * 1024, "row batch needs %zu bytes of LDS (max 65536)", lds);
  const int bs = block_size_of(ex, 256);
  OP2HIP_REQUIRE(bs > 0 && bs <= 256 && bs % 64 == 0, "row-gather block size %d", bs);
  // eight prefetched old values per thread: deeper prefetch raises the
  // register count past two workgroups per SIMD and loses more than it hides
  rowgather_kernel<Elem, 8><<<ex->nbatches, bs, lds, s>>>(start, end, elem, *A, ex->batch_rows,
                                                           ex->nbatches, rowdat);
  OP2HIP_LAUNCHED();
  return OP2HIP_OK;
}

}  // namespace op2hip

namespace op2hip {

// Lane-per-block variant: the kCar lanes of a group share one incidence
// (element e, local row i) and each computes one column block j of row block
// i of the element matrix (the group computes e's geometry redundantly, in
// lock-step, at no extra instruction cost).  A wave holds 64 / kCar groups.
// Compared with one incidence per thread this exposes kCar times more
// parallelism and keeps every lane's dependency chain short.
//
// Elem must additionally provide:
//   struct Geom;  __device__ Geom geometry(int32_t e) const;
//   __device__ int32_t col(int32_t e, int j) const;
//   __device__ void block(const Geom&, int i, int j, T *blk) const;
//   __device__ void rowdat(int32_t col, const T *blk, T *rs) const;   (fused)
template <class Elem>
__global__ void __launch_bounds__(256)
rowgather_lanes_kernel(int32_t start, int32_t end, Elem elem, op2hip_mat A,
                       const int32_t *__restrict__ batch_rows, int32_t nbatches,
                       typename Elem::T *__restrict__ rowdat) {
  using T = typename Elem::T;
  constexpr int kBs = Elem::kRd * Elem::kCd;
  constexpr int kGroups = 64 / Elem::kCar;     // incidences per wave and round
  extern __shared__ __align__(16) unsigned char smem[];
  const int32_t b = xcd_contiguous(blockIdx.x, nbatches);
  const int32_t r0 = batch_rows[b], r1 = batch_rows[b + 1];
  const int32_t nr = r1 - r0;
  const int32_t z0 = A.rowptr[r0], nz = A.rowptr[r1] - z0;
  const int32_t i0 = A.inc_ptr[r0], ni = A.inc_ptr[r1] - i0;

  T *acc = reinterpret_cast<T *>(smem);
  T *racc = acc + (size_t)nz * kBs;
  size_t off = ((size_t)nz * kBs + (Elem::kHasRowDat ? (size_t)nr * Elem::kRd : 0)) * sizeof(T);
  off = (off + 15) & ~(size_t)15;
  int32_t *scol = reinterpret_cast<int32_t *>(smem + off);
  int32_t *srow = scol + nz;
  int32_t *sinc = srow + nr + 1;

  for (int32_t k = threadIdx.x; k < nz * kBs; k += blockDim.x) acc[k] = T(0);
  if (Elem::kHasRowDat)
    for (int32_t k = threadIdx.x; k < nr * Elem::kRd; k += blockDim.x) racc[k] = T(0);
  stage_copy(scol, A.colidx + z0, nz);
  for (int32_t k = threadIdx.x; k <= nr; k += blockDim.x) {
    srow[k] = A.rowptr[r0 + k] - z0;
    sinc[k] = A.inc_ptr[r0 + k] - i0;
  }
  __syncthreads();

  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int nwaves = blockDim.x >> 6;
  const int grp = lane / Elem::kCar;
  const int j = lane - grp * Elem::kCar;
  for (int32_t base = wave * kGroups; base < ni; base += nwaves * kGroups) {
    const int32_t t = base + grp;
    if (grp >= kGroups || t >= ni) continue;
    const int32_t ent = A.inc_ent[i0 + t];
    const int32_t e = ent / Elem::kRar;
    if (e < start || e >= end) continue;
    const int i = ent - e * Elem::kRar;
    int32_t lo = 0, hi = nr - 1;
    while (lo < hi) {
      const int32_t mid = (lo + hi + 1) >> 1;
      if (sinc[mid] <= t) lo = mid; else hi = mid - 1;
    }
    const int32_t lr = lo;
    const int32_t col = elem.col(e, j);
    if (col < 0) continue;
    const typename Elem::Geom g = elem.geometry(e);
    T blk[kBs];
    elem.block(g, i, j, blk);
    const int32_t slot = find_slot(scol, srow[lr], srow[lr + 1], col);
    if (slot < 0) {
      atomicAdd(A.status, 1);
      continue;
    }
#pragma unroll
    for (int k = 0; k < kBs; ++k) atomic_add(&acc[slot * kBs + k], blk[k]);
    if constexpr (Elem::kHasRowDat) {
      T rs[Elem::kRd];
#pragma unroll
      for (int a = 0; a < Elem::kRd; ++a) rs[a] = T(0);
      elem.rowdat(col, blk, rs);
#pragma unroll
      for (int a = 0; a < Elem::kRd; ++a) atomic_add(&racc[lr * Elem::kRd + a], rs[a]);
    }
  }
  __syncthreads();

  flush_add(static_cast<T *>(A.values) + (size_t)z0 * kBs, acc, nz * kBs);
  if (Elem::kHasRowDat) flush_add(rowdat + (size_t)r0 * Elem::kRd, racc, nr * Elem::kRd);
}

template <class Elem>
int launch_rowgather_lanes(int32_t start, int32_t end, const Elem &elem, const op2hip_mat *A,
                           const op2hip_exec *ex, typename Elem::T *rowdat, hipStream_t s) {
  OP2HIP_REQUIRE(ex && ex->batch_rows && ex->nbatches >= 0,
                 "row-gather launch without a row-batch plan");
  OP2HIP_REQUIRE(A->inc_ptr && A->inc_ent, "row-gather launch without row incidences");
  OP2HIP_REQUIRE(!ex->subset, "row-gather assembly does not iterate Subsets");
  OP2HIP_REQUIRE(A->rdim == Elem::kRd && A->cdim == Elem::kCd,
                 "matrix block %dx%d, kernel assembles %dx%d", A->rdim, A->cdim,
                 Elem::kRd, Elem::kCd);
  if (end <= start || ex->nbatches == 0) return OP2HIP_OK;
  const size_t lds = RowGatherLayout<Elem>::bytes(ex->batch_max_nnz, ex->batch_max_rows);
  OP2HIP_REQUIRE(lds <= 64 * 1024, "row batch needs %zu bytes of LDS (max 65536)", lds);
  const int bs = block_size_of(ex, 256);
  OP2HIP_REQUIRE(bs > 0 && bs <= 256 && bs % 64 == 0, "row-gather block size %d", bs);
  rowgather_lanes_kernel<Elem><<<ex->nbatches, bs, lds, s>>>(start, end, elem, *A,
                                                              ex->batch_rows, ex->nbatches, rowdat);
  OP2HIP_LAUNCHED();
  return OP2HIP_OK;
}

}  // namespace op2hip

namespace op2hip {

// ---------------------------------------------------------------------------
// Slot-table row gather.  A one-off plan step stores, per row incidence,
// where its contributions land in the batch accumulator: the row's block
// offset inside the batch, the row's index inside the batch and, per column
// block j, the block's position inside the row.  The assembly kernel then
// needs neither the pattern (rowptr / colidx) nor any search: LDS holds only
// the accumulator, and each workgroup's header is one 32-byte descriptor.

template <int kCar>
struct alignas(8) SlotEntry {
  uint16_t row_off;          // block offset of the row inside the batch
  uint16_t lr;               // row index inside the batch
  uint8_t slot[kCar];        // column block j at row_off + slot[j]; 255 = none
};

struct BatchDesc {
  int32_t r0, nr, z0, nz, i0, ni, pad0, pad1;
};

// A slot entry read once per launch: non-temporal 8-byte loads, so the
// streamed table does not evict the gathered coordinates and maps.
template <int kCar>
__device__ __forceinline__ SlotEntry<kCar> load_slot_entry_nt(const SlotEntry<kCar> *p) {
  static_assert(sizeof(SlotEntry<kCar>) % 8 == 0, "slot entries are whole 8-byte words");
  constexpr int kW = sizeof(SlotEntry<kCar>) / 8;
  const uint64_t *q = reinterpret_cast<const uint64_t *>(p);
  uint64_t w[kW];
#pragma unroll
  for (int k = 0; k < kW; ++k) w[k] = __builtin_nontemporal_load(q + k);
  SlotEntry<kCar> out;
  __builtin_memcpy(&out, w, sizeof(out));
  return out;
}

template <int kCar>
__global__ void build_slot_table(int32_t nbatches, const int32_t *__restrict__ batch_rows,
                                 const int32_t *__restrict__ rowptr,
                                 const int32_t *__restrict__ colidx,
                                 const int32_t *__restrict__ inc_ptr,
                                 const int32_t *__restrict__ inc_ent,
                                 const int32_t *__restrict__ cmap, int32_t rarity,
                                 SlotEntry<kCar> *__restrict__ table,
                                 BatchDesc *__restrict__ desc, int32_t *__restrict__ bad) {
  const int32_t b = blockIdx.x;
  if (b >= nbatches) return;
  const int32_t r0 = batch_rows[b], r1 = batch_rows[b + 1];
  if (threadIdx.x == 0) {
    BatchDesc d;
    d.r0 = r0;
    d.nr = r1 - r0;
    d.z0 = rowptr[r0];
    d.nz = rowptr[r1] - rowptr[r0];
    d.i0 = inc_ptr[r0];
    d.ni = inc_ptr[r1] - inc_ptr[r0];
    d.pad0 = d.pad1 = 0;
    desc[b] = d;
  }
  for (int32_t r = r0; r < r1; ++r) {
    const int32_t zlo = rowptr[r], zhi = rowptr[r + 1];
    for (int32_t t = inc_ptr[r] + threadIdx.x; t < inc_ptr[r + 1]; t += blockDim.x) {
      const int32_t ent = inc_ent[t];
      const int32_t e = ent / rarity;
      SlotEntry<kCar> s;
      s.row_off = static_cast<uint16_t>(zlo - rowptr[r0]);
      s.lr = static_cast<uint16_t>(r - r0);
      bool repeated = false;
#pragma unroll
      for (int j = 0; j < kCar; ++j) {
        const int32_t c = cmap[(int64_t)e * kCar + j];
        const int32_t p = c < 0 ? -1 : find_slot(colidx, zlo, zhi, c);
        if (c >= 0 && p < 0) atomicAdd(bad, 1);
        s.slot[j] = (p < 0 || p - zlo > 254) ? uint8_t(255) : static_cast<uint8_t>(p - zlo);
#pragma unroll
        for (int jj = 0; jj < j; ++jj) repeated |= s.slot[j] != 255 && s.slot[jj] == s.slot[j];
      }
      if (repeated) atomicAdd(bad + 1, 1);
      table[t] = s;
    }
  }
}

// Gather table of the staged row gather, built from a slot table.  Output
// block z of a batch (its z-th Mat block) receives the element blocks
// (t, j) -- row incidence t of the batch, column block j -- whose slot is z;
// gather_count[z0 + z] is their number and the batch's gather list holds
// t * kCar + j for z = 0, 1, ... in turn, each run in incidence order.
constexpr int kGatherMaxBlocks = 1024;   // blocks of a batch the plan kernels take

template <int kCar>
__global__ void __launch_bounds__(256)
gather_count_kernel(int32_t nbatches, const BatchDesc *__restrict__ desc,
                    const SlotEntry<kCar> *__restrict__ table, uint8_t *__restrict__ gcount,
                    int32_t *__restrict__ batch_total, int32_t *__restrict__ over) {
  __shared__ int32_t cnt[kGatherMaxBlocks];
  __shared__ int32_t total;
  const int32_t b = blockIdx.x;
  if (b >= nbatches) return;
  const BatchDesc d = desc[b];
  for (int32_t z = threadIdx.x; z < d.nz; z += blockDim.x) cnt[z] = 0;
  if (threadIdx.x == 0) total = 0;
  __syncthreads();
  for (int32_t q = threadIdx.x; q < d.ni * kCar; q += blockDim.x) {
    const int32_t t = q / kCar, j = q - t * kCar;
    const SlotEntry<kCar> se = table[d.i0 + t];
    if (se.slot[j] != 255) atomicAdd(&cnt[se.row_off + se.slot[j]], 1);
  }
  __syncthreads();
  int32_t mine = 0;
  for (int32_t z = threadIdx.x; z < d.nz; z += blockDim.x) {
    const int32_t c = cnt[z];
    if (c > 255) atomicAdd(over, 1);
    gcount[(size_t)d.z0 + z] = static_cast<uint8_t>(c > 255 ? 255 : c);
    mine += c;
  }
  atomicAdd(&total, mine);
  __syncthreads();
  if (threadIdx.x == 0) batch_total[b] = total;
}

template <int kCar>
__global__ void __launch_bounds__(256)
gather_fill_kernel(int32_t nbatches, BatchDesc *__restrict__ desc,
                   const SlotEntry<kCar> *__restrict__ table, const int32_t *__restrict__ batch_off,
                   const int32_t *__restrict__ batch_total, const uint8_t *__restrict__ gcount,
                   uint16_t *__restrict__ glist) {
  __shared__ int32_t first[kGatherMaxBlocks];
  __shared__ int32_t cursor[kGatherMaxBlocks];
  const int32_t b = blockIdx.x;
  if (b >= nbatches) return;
  const BatchDesc d = desc[b];
  const int32_t off = batch_off[b];
  if (threadIdx.x == 0) {
    desc[b].pad0 = off;
    desc[b].pad1 = batch_total[b];
    int32_t acc = 0;
    for (int32_t z = 0; z < d.nz; ++z) {
      first[z] = acc;
      cursor[z] = acc;
      acc += gcount[(size_t)d.z0 + z];
    }
  }
  __syncthreads();
  uint16_t *list = glist + off;
  for (int32_t q = threadIdx.x; q < d.ni * kCar; q += blockDim.x) {
    const int32_t t = q / kCar, j = q - t * kCar;
    const SlotEntry<kCar> se = table[d.i0 + t];
    if (se.slot[j] != 255) {
      const int32_t p = atomicAdd(&cursor[se.row_off + se.slot[j]], 1);
      list[p] = static_cast<uint16_t>(q);
    }
  }
  __syncthreads();
  // incidence order within each block's run (runs are short: insertion sort)
  for (int32_t z = threadIdx.x; z < d.nz; z += blockDim.x) {
    uint16_t *run = list + first[z];
    const int32_t n = gcount[(size_t)d.z0 + z];
    for (int32_t x = 1; x < n; ++x) {
      const uint16_t v = run[x];
      int32_t y = x - 1;
      while (y >= 0 && run[y] > v) {
        run[y + 1] = run[y];
        --y;
      }
      run[y + 1] = v;
    }
  }
}

// Incidence handled by thread slot t of a batch of ni incidences: a stride
// coprime to ni spreads a wave's lanes over the batch's rows, so lanes rarely
// add into the same block in one LDS instruction (consecutive incidences of a
// row all hit that row's diagonal block).  Stride 37 or 41 (both prime), or
// 1 when ni is a multiple of both.
__device__ __forceinline__ int32_t spread_incidence(int32_t t, int32_t ni) {
  const int32_t p = (ni % 37) ? 37 : ((ni % 41) ? 41 : 1);
  return (int32_t)(((int64_t)t * p) % ni);
}

// Whether an element spreads its incidences over the lanes (Elem::kSpread):
// worth it when the element work is large (P2), not when neighbouring lanes
// gain more from sharing their gathers (P1).
template <class Elem, class = void>
struct SpreadOf {
  static constexpr bool value = false;
};
template <class Elem>
struct SpreadOf<Elem, std::void_t<decltype(Elem::kSpread)>> {
  static constexpr bool value = Elem::kSpread;
};

// Lanes per row incidence in the slot-table gather (Elem::kHalves = 2 with a
// row_half<H> member splits an incidence's column blocks between two lanes).
template <class Elem, class = void>
struct HalvesOf {
  static constexpr int value = 1;
};
template <class Elem>
struct HalvesOf<Elem, std::void_t<decltype(Elem::kHalves)>> {
  static constexpr int value = Elem::kHalves;
};

// Elements whose row evaluation splits into a gather of the element's vertex
// coordinates and arithmetic on them (Elem::kSplitGather = true) provide
//   static constexpr int kNv, kGdim;          vertices per element, coordinates
//   const T *X;                                coordinates, kGdim per vertex
//   void vertices(int32_t e, int32_t (&v)[kNv]) const;
//   template <class Emit> void row_at(const T (&xv)[kNv][kGdim], int i, Emit &&) const;
// where row_at emits every column block of row i (column id argument unused).
template <class Elem, class = void>
struct SplitGatherOf {
  static constexpr bool value = false;
};
template <class Elem>
struct SplitGatherOf<Elem, std::void_t<decltype(Elem::kSplitGather)>> {
  static constexpr bool value = Elem::kSplitGather;
};

// Row incidences per thread and round of a split-gather element
// (Elem::kRound, default 2).
template <class Elem, class = void>
struct RoundOf {
  static constexpr int value = 2;
};
template <class Elem>
struct RoundOf<Elem, std::void_t<decltype(Elem::kRound)>> {
  static constexpr int value = Elem::kRound;
};

// One round of a split-gather batch: kR row incidences per thread.  load()
// issues the incidence entries, slot entries (both streamed, non-temporal)
// and vertex ids of all of them before any element is evaluated, so the
// dependent loads of one incidence do not wait behind the arithmetic of the
// previous one; eval() gathers each incidence's coordinates, evaluates its
// block row and adds it into the batch's LDS accumulators.
template <class Elem>
struct SplitRound {
  using T = typename Elem::T;
  static constexpr int kR = RoundOf<Elem>::value;
  static constexpr int kBs = Elem::kRd * Elem::kCd;
  int32_t ent[kR], vv[kR][Elem::kNv];
  SlotEntry<Elem::kCar> se[kR];
  bool ok[kR];

  __device__ __forceinline__ void load(const Elem &elem, int32_t base, int32_t ni,
                                       const int32_t *__restrict__ ents,
                                       const SlotEntry<Elem::kCar> *__restrict__ table,
                                       int32_t start, int32_t end) {
    const int32_t bd = (int32_t)blockDim.x;
#pragma unroll
    for (int u = 0; u < kR; ++u) {
      const int32_t t0 = base + u * bd;
      ok[u] = t0 < ni;
      const int32_t t = SpreadOf<Elem>::value ? spread_incidence(ok[u] ? t0 : 0, ni) : t0;
      ent[u] = ok[u] ? __builtin_nontemporal_load(&ents[t]) : 0;
      if (ok[u]) se[u] = load_slot_entry_nt(&table[t]);
    }
#pragma unroll
    for (int u = 0; u < kR; ++u) {
      const int32_t e = ent[u] / Elem::kRar;
      ok[u] = ok[u] && e >= start && e < end;
      if (ok[u]) {
        elem.vertices(e, vv[u]);
      } else {
#pragma unroll
        for (int m = 0; m < Elem::kNv; ++m) vv[u][m] = 0;
      }
    }
  }

  __device__ __forceinline__ void eval(const Elem &elem, double *acc) const {
#pragma unroll
    for (int u = 0; u < kR; ++u) {
      if (!ok[u]) continue;
      T xv[Elem::kNv][Elem::kGdim];
#pragma unroll
      for (int m = 0; m < Elem::kNv; ++m)
#pragma unroll
        for (int k = 0; k < Elem::kGdim; ++k)
          xv[m][k] = elem.X[(size_t)Elem::kGdim * vv[u][m] + k];
      const int32_t e = ent[u] / Elem::kRar;
      const int i = ent[u] - e * Elem::kRar;
      double *row = acc + (size_t)se[u].row_off * kBs;
      const SlotEntry<Elem::kCar> &s = se[u];
      elem.row_at(xv, i, [&](int j, int32_t, const T *blk) {
        // slot 255 is a negative column id, which adds nothing (PETSc drops
        // negative indices): a column outside the pattern was refused when
        // the table was built
        const int sl = s.slot[j];
        if (sl == 255) return;
#pragma unroll
        for (int k = 0; k < kBs; ++k) atomic_add(&row[sl * kBs + k], double(blk[k]));
      });
    }
  }
};

// Slot-table batch of a split-gather element, in rounds of SplitRound.
// float64 batches start their LDS accumulators from the old matrix values
// -- all loads of the batch are issued before the first round waits, and the
// batch is written back from LDS with no load after the element work --
// while float32 batches keep the old values of their first kPre * blockDim
// entries in registers (their LDS accumulators, in float64, start from zero).
template <class Elem, int kPre>
__device__ __forceinline__ void slots_batch_split(const BatchDesc &d, int32_t start, int32_t end,
                                                  const Elem &elem, const op2hip_mat &A,
                                                  const SlotEntry<Elem::kCar> *__restrict__ table) {
  using T = typename Elem::T;
  using R = SplitRound<Elem>;
  constexpr int kBs = R::kBs;
  constexpr bool kInit = sizeof(T) == 8;
  constexpr int kInitUnroll = 24;
  extern __shared__ __align__(16) unsigned char smem[];
  T *__restrict__ vals = static_cast<T *>(A.values) + (size_t)d.z0 * kBs;
  const int32_t nval = d.nz * kBs;
  double *acc = reinterpret_cast<double *>(smem);
  const int32_t bd = (int32_t)blockDim.x;

  R round;
  T old[kInit ? 1 : kPre];
  if constexpr (!kInit) {
#pragma unroll
    for (int u = 0; u < kPre; ++u) {
      const int32_t k = threadIdx.x + u * bd;
      old[u] = k < nval ? __builtin_nontemporal_load(&vals[k]) : T(0);
    }
  }
  round.load(elem, threadIdx.x, d.ni, A.inc_ent + d.i0, table + d.i0, start, end);
  if constexpr (kInit) {
    for (int32_t k0 = threadIdx.x; k0 < nval; k0 += kInitUnroll * bd) {
      T v[kInitUnroll];
#pragma unroll
      for (int u = 0; u < kInitUnroll; ++u) {
        const int32_t k = k0 + u * bd;
        v[u] = k < nval ? __builtin_nontemporal_load(&vals[k]) : T(0);
      }
#pragma unroll
      for (int u = 0; u < kInitUnroll; ++u) {
        const int32_t k = k0 + u * bd;
        if (k < nval) acc[k] = double(v[u]);
      }
    }
  } else {
    for (int32_t k = threadIdx.x; k < nval; k += bd) acc[k] = 0.0;
  }
  __syncthreads();

  for (int32_t base = threadIdx.x; base < d.ni; base += R::kR * bd) {
    if (base != (int32_t)threadIdx.x)
      round.load(elem, base, d.ni, A.inc_ent + d.i0, table + d.i0, start, end);
    round.eval(elem, acc);
  }
  __syncthreads();

  if constexpr (kInit) {
    for (int32_t k = threadIdx.x; k < nval; k += bd) __builtin_nontemporal_store(T(acc[k]), &vals[k]);
  } else {
#pragma unroll
    for (int u = 0; u < kPre; ++u) {
      const int32_t k = threadIdx.x + u * bd;
      if (k < nval) __builtin_nontemporal_store(old[u] + T(acc[k]), &vals[k]);
    }
    for (int32_t k0 = threadIdx.x + kPre * bd; k0 < nval; k0 += 8 * bd) {
      T v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int32_t k = k0 + u * bd;
        v[u] = k < nval ? vals[k] : T(0);
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int32_t k = k0 + u * bd;
        if (k < nval) vals[k] = v[u] + T(acc[k]);
      }
    }
  }
}

// One batch of the slot-table row gather (the workgroup's share of a launch,
// or one work item of a batched launch).
template <class Elem, int kPre>
__device__ __forceinline__ void slots_batch(const BatchDesc &d, int32_t start, int32_t end,
                                            const Elem &elem, const op2hip_mat &A,
                                            const SlotEntry<Elem::kCar> *__restrict__ table,
                                            typename Elem::T *__restrict__ rowdat) {
  using T = typename Elem::T;
  using Acc = typename AccumOf<T>::type;
  constexpr int kBs = Elem::kRd * Elem::kCd;
  extern __shared__ __align__(16) unsigned char smem[];
  // the batch's values are read once and written once: non-temporal
  T *__restrict__ vals = static_cast<T *>(A.values) + (size_t)d.z0 * kBs;
  const int32_t nval = d.nz * kBs;
  if constexpr (SplitGatherOf<Elem>::value && !Elem::kHasRowDat) {
    slots_batch_split<Elem, kPre>(d, start, end, elem, A, table);
    return;
  }

  T old[kPre];
#pragma unroll
  for (int u = 0; u < kPre; ++u) {
    const int32_t k = threadIdx.x + u * (int32_t)blockDim.x;
    old[u] = k < nval ? __builtin_nontemporal_load(&vals[k]) : T(0);
  }
  Acc *acc = reinterpret_cast<Acc *>(smem);
  Acc *racc = acc + nval;
  for (int32_t k = threadIdx.x; k < nval; k += blockDim.x) acc[k] = Acc(0);
  if (Elem::kHasRowDat)
    for (int32_t k = threadIdx.x; k < d.nr * Elem::kRd; k += blockDim.x) racc[k] = Acc(0);
  __syncthreads();

  constexpr int kH = HalvesOf<Elem>::value;
  for (int32_t tau = threadIdx.x; tau < kH * d.ni; tau += blockDim.x) {
    // task tau: incidence tau % ni, half tau / ni (uniform over most waves)
    const int h = kH == 1 ? 0 : int(tau >= d.ni);
    const int32_t t0 = tau - h * d.ni;
    const int32_t t = SpreadOf<Elem>::value ? spread_incidence(t0, d.ni) : t0;
    const int32_t ent = A.inc_ent[d.i0 + t];
    const SlotEntry<Elem::kCar> se = table[d.i0 + t];
    const int32_t e = ent / Elem::kRar;
    if (e < start || e >= end) continue;
    const int i = ent - e * Elem::kRar;
    Acc *row = acc + (size_t)se.row_off * kBs;
    T rsum[Elem::kRd];
#pragma unroll
    for (int a = 0; a < Elem::kRd; ++a) rsum[a] = T(0);
    auto emit = [&](int j, int32_t col, const T *blk) {
      const int s = se.slot[j];
      if (col < 0) return;
      if (s == 255) {
        atomicAdd(A.status, 1);
        return;
      }
#pragma unroll
      for (int k = 0; k < kBs; ++k) atomic_add(&row[s * kBs + k], Acc(blk[k]));
    };
    if constexpr (kH == 1) {
      elem.row(e, i, emit, rsum);
    } else {
      if (h == 0) elem.template row_half<0>(e, i, emit, rsum);
      else elem.template row_half<1>(e, i, emit, rsum);
    }
    if (Elem::kHasRowDat) {
#pragma unroll
      for (int a = 0; a < Elem::kRd; ++a) atomic_add(&racc[se.lr * Elem::kRd + a], Acc(rsum[a]));
    }
  }
  __syncthreads();

#pragma unroll
  for (int u = 0; u < kPre; ++u) {
    const int32_t k = threadIdx.x + u * (int32_t)blockDim.x;
    if (k < nval) __builtin_nontemporal_store(old[u] + T(acc[k]), &vals[k]);
  }
  for (int32_t k0 = threadIdx.x + kPre * (int32_t)blockDim.x; k0 < nval; k0 += 8 * (int32_t)blockDim.x) {
    T v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int32_t k = k0 + u * (int32_t)blockDim.x;
      v[u] = k < nval ? vals[k] : T(0);
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int32_t k = k0 + u * (int32_t)blockDim.x;
      if (k < nval) vals[k] = v[u] + T(acc[k]);
    }
  }
  if (Elem::kHasRowDat) {
    T *rd = rowdat + (size_t)d.r0 * Elem::kRd;
    for (int32_t k = threadIdx.x; k < d.nr * Elem::kRd; k += blockDim.x) rd[k] += T(racc[k]);
  }
}

// How the slot-table kernel of an element maps workgroups to batches
// (Elem::kXcdChunk): 0 (the default) hands each XCD a contiguous range of
// batches, kChunk > 0 deals chunks of kChunk consecutive batches to the
// XCDs in turn.
template <class Elem, class = void>
struct XcdChunkOf {
  static constexpr int32_t value = 0;
};
template <class Elem>
struct XcdChunkOf<Elem, std::void_t<decltype(Elem::kXcdChunk)>> {
  static constexpr int32_t value = Elem::kXcdChunk;
};

template <class Elem, int kPre>
__global__ void __launch_bounds__(256)
rowgather_slots_kernel(int32_t start, int32_t end, Elem elem, op2hip_mat A,
                       const BatchDesc *__restrict__ descs, int32_t nbatches,
                       const SlotEntry<Elem::kCar> *__restrict__ table,
                       typename Elem::T *__restrict__ rowdat) {
  constexpr int32_t kChunk = XcdChunkOf<Elem>::value;
  int32_t b;
  if constexpr (kChunk > 0)
    b = xcd_dealt<kChunk>(blockIdx.x, nbatches);
  else
    b = xcd_contiguous(blockIdx.x, nbatches);
  slots_batch<Elem, kPre>(descs[b], start, end, elem, A, table, rowdat);
}

// Deepest old-value prefetch of an element (Elem::kMaxPrefetch when given):
// the register count grows by two (double) per prefetched value.
template <class Elem, class = void>
struct MaxPrefetch {
  static constexpr int64_t value = 24;
};
template <class Elem>
struct MaxPrefetch<Elem, std::void_t<decltype(Elem::kMaxPrefetch)>> {
  static constexpr int64_t value = Elem::kMaxPrefetch;
};

template <class Elem>
int launch_rowgather_slots(int32_t start, int32_t end, const Elem &elem, const op2hip_mat *A,
                           const op2hip_exec *ex, typename Elem::T *rowdat, hipStream_t s) {
  OP2HIP_REQUIRE(ex && ex->batch_desc && ex->slot_table && ex->nbatches >= 0,
                 "slot-table row gather without its plan");
  OP2HIP_REQUIRE(A->inc_ent, "row-gather launch without row incidences");
  OP2HIP_REQUIRE(!ex->subset, "row-gather assembly does not iterate Subsets");
  OP2HIP_REQUIRE(A->rdim == Elem::kRd && A->cdim == Elem::kCd,
                 "matrix block %dx%d, kernel assembles %dx%d", A->rdim, A->cdim,
                 Elem::kRd, Elem::kCd);
  if (end <= start || ex->nbatches == 0) return OP2HIP_OK;
  const size_t lds = (size_t)ex->batch_max_nnz * Elem::kRd * Elem::kCd * sizeof(double) +
                     (Elem::kHasRowDat ? (size_t)ex->batch_max_rows * Elem::kRd * sizeof(double) : 0);
  OP2HIP_REQUIRE(lds <= 64 * 1024, "row batch needs %zu bytes of LDS (max 65536)", lds);
  const int bs = block_size_of(ex, 256);
  OP2HIP_REQUIRE(bs > 0 && bs <= 256 && bs % 64 == 0, "row-gather block size %d", bs);
  // old values of the whole batch are loaded into registers before the
  // element work starts (up to 24 per thread, which keeps the register count
  // below the LDS-bound occupancy of these kernels)
  const auto *descs = static_cast<const BatchDesc *>(ex->batch_desc);
  const auto *table = static_cast<const SlotEntry<Elem::kCar> *>(ex->slot_table);
  const int64_t per_thread = std::min<int64_t>(
      ((int64_t)ex->batch_max_nnz * Elem::kRd * Elem::kCd + bs - 1) / bs, MaxPrefetch<Elem>::value);
  if constexpr (SplitGatherOf<Elem>::value && !Elem::kHasRowDat) {
    (void)per_thread;
    rowgather_slots_kernel<Elem, 8><<<ex->nbatches, bs, lds, s>>>(start, end, elem, *A, descs,
                                                                   ex->nbatches, table, rowdat);
  } else {
    if (per_thread <= 8)
      rowgather_slots_kernel<Elem, 8><<<ex->nbatches, bs, lds, s>>>(start, end, elem, *A, descs,
                                                                     ex->nbatches, table, rowdat);
    else if (per_thread <= 16)
      rowgather_slots_kernel<Elem, 16><<<ex->nbatches, bs, lds, s>>>(start, end, elem, *A, descs,
                                                                      ex->nbatches, table, rowdat);
    else
      rowgather_slots_kernel<Elem, 24><<<ex->nbatches, bs, lds, s>>>(start, end, elem, *A, descs,
                                                                      ex->nbatches, table, rowdat);
  }
  OP2HIP_LAUNCHED();
  return OP2HIP_OK;
}

// ---------------------------------------------------------------------------
// Staged row gather: no atomics.  A batch holds at most blockDim / kGroups row
// incidences.  Phase 1 computes the element block rows of the batch's
// incidences into LDS; the element's column blocks are split into kGroups
// groups and thread group g (whole waves) computes group g's blocks for every
// incidence, so each wave runs one group's code.  Phase 2 gives every output
// block of the batch to one thread, which adds the element blocks its gather
// list names (in incidence = element order) to the old value and writes the
// block once.  No value depends on scheduling.
//
// Elem must additionally provide
//   static constexpr int kGroups;
//   template <int G, class Emit> void row_group(int32_t e, int i, Emit &&, T *rsum) const;
// where row_group<G> emits the column blocks of group G of row i of e.

template <class Elem, int G = 0, class Emit>
__device__ __forceinline__ void row_group_of(const Elem &elem, int g, int32_t e, int i,
                                             Emit &&emit, typename Elem::T *rsum) {
  if constexpr (G < Elem::kGroups) {
    if (g == G) elem.template row_group<G>(e, i, emit, rsum);
    else row_group_of<Elem, G + 1>(elem, g, e, i, emit, rsum);
  }
}

// LDS of the staged kernel: element block rows | row sums (fused) | gather
// list of the batch | wave totals of the block scan | row incidence offsets.
template <class Elem>
struct StagedLayout {
  using T = typename Elem::T;
  static constexpr int kBs = Elem::kRd * Elem::kCd;
  __host__ __device__ static size_t rowbuf(int per_pass) {
    return (size_t)per_pass * Elem::kCar * kBs * sizeof(T);
  }
  __host__ __device__ static size_t rsbuf(int per_pass) {
    return Elem::kHasRowDat ? (size_t)Elem::kGroups * per_pass * Elem::kRd * sizeof(T) : 0;
  }
  __host__ __device__ static size_t list_at(int per_pass) {
    return (rowbuf(per_pass) + rsbuf(per_pass) + 15) & ~(size_t)15;
  }
  __host__ __device__ static size_t wtot_at(int per_pass) {
    return (list_at(per_pass) + (size_t)per_pass * Elem::kCar * sizeof(uint16_t) + 15) & ~(size_t)15;
  }
  __host__ __device__ static size_t inc_at(int per_pass, int kOut) {
    return wtot_at(per_pass) + (size_t)kOut * 4 * sizeof(int32_t);
  }
  static size_t bytes(int per_pass, int kOut, int32_t max_rows) {
    return inc_at(per_pass, kOut) + (Elem::kHasRowDat ? (size_t)(max_rows + 1) * sizeof(int32_t) : 0);
  }
};

template <class Elem, int kOut>
__global__ void __launch_bounds__(256)
rowgather_staged_kernel(int32_t start, int32_t end, Elem elem, op2hip_mat A,
                        const BatchDesc *__restrict__ descs, int32_t nbatches,
                        const uint8_t *__restrict__ gcount, const uint16_t *__restrict__ glist,
                        typename Elem::T *__restrict__ rowdat) {
  using T = typename Elem::T;
  using Acc = typename AccumOf<T>::type;
  using L = StagedLayout<Elem>;
  constexpr int kBs = Elem::kRd * Elem::kCd;
  constexpr int kCar = Elem::kCar;
  extern __shared__ __align__(16) unsigned char smem[];
  const int per_pass = blockDim.x / Elem::kGroups;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
  const BatchDesc d = descs[xcd_contiguous(blockIdx.x, nbatches)];
  T *__restrict__ vals = static_cast<T *>(A.values) + (size_t)d.z0 * kBs;

  T *rowbuf = reinterpret_cast<T *>(smem);
  T *rsbuf = rowbuf + (size_t)per_pass * kCar * kBs;
  uint16_t *slist = reinterpret_cast<uint16_t *>(smem + L::list_at(per_pass));
  int32_t *wtot = reinterpret_cast<int32_t *>(smem + L::wtot_at(per_pass));
  int32_t *sinc = reinterpret_cast<int32_t *>(smem + L::inc_at(per_pass, kOut));

  // loads in flight during phase 1: old values and gather counts of this
  // thread's output blocks, the batch's gather list (pad1 <= ni * kCar)
  T old[kOut][kBs];
  int32_t cnt[kOut];
#pragma unroll
  for (int o = 0; o < kOut; ++o) {
    const int32_t z = threadIdx.x + o * (int32_t)blockDim.x;
    cnt[o] = z < d.nz ? int32_t(gcount[(size_t)d.z0 + z]) : 0;
#pragma unroll
    for (int k = 0; k < kBs; ++k) old[o][k] = z < d.nz ? vals[(size_t)z * kBs + k] : T(0);
  }
  for (int32_t k = threadIdx.x; k < d.pad1; k += blockDim.x) slist[k] = glist[(size_t)d.pad0 + k];
  if (Elem::kHasRowDat)
    for (int32_t k = threadIdx.x; k <= d.nr; k += blockDim.x) sinc[k] = A.inc_ptr[d.r0 + k] - d.i0;

  // phase 1: element block rows into LDS
  const int g = threadIdx.x / per_pass;
  const int t = threadIdx.x - g * per_pass;
  if (t < d.ni) {
    const int32_t ent = A.inc_ent[d.i0 + t];
    const int32_t e = ent / Elem::kRar;
    const int i = ent - e * Elem::kRar;
    T *rb = rowbuf + (size_t)t * kCar * kBs;
    T rsum[Elem::kRd];
#pragma unroll
    for (int a = 0; a < Elem::kRd; ++a) rsum[a] = T(0);
    auto emit = [&](int j, int32_t, const T *blk) {
#pragma unroll
      for (int k = 0; k < kBs; ++k) rb[j * kBs + k] = blk[k];
    };
    if (e >= start && e < end) {
      row_group_of(elem, g, e, i, emit, rsum);
    } else if (g == 0) {
      // outside the iteration range: the incidence adds zero blocks
      for (int k = 0; k < kCar * kBs; ++k) rb[k] = T(0);
    }
    if (Elem::kHasRowDat) {
#pragma unroll
      for (int a = 0; a < Elem::kRd; ++a) rsbuf[((size_t)g * per_pass + t) * Elem::kRd + a] = rsum[a];
    }
  }

  // exclusive scan of the gather counts in block order z = tid + o * blockDim
  int32_t incl[kOut];
#pragma unroll
  for (int o = 0; o < kOut; ++o) {
    int32_t x = cnt[o];
#pragma unroll
    for (int dd = 1; dd < 64; dd <<= 1) {
      const int32_t y = __shfl_up(x, dd, 64);
      if (lane >= dd) x += y;
    }
    incl[o] = x;
    if (lane == 63) wtot[o * 4 + wave] = x;
  }
  __syncthreads();

  // phase 2: one thread per output block
  int32_t base = 0;
#pragma unroll
  for (int o = 0; o < kOut; ++o) {
    int32_t before = 0, all = 0;
    for (int w = 0; w < nwaves; ++w) {
      const int32_t v = wtot[o * 4 + w];
      before += w < wave ? v : 0;
      all += v;
    }
    const int32_t z = threadIdx.x + o * (int32_t)blockDim.x;
    const int32_t first = base + before + incl[o] - cnt[o];
    base += all;
    if (z >= d.nz) continue;
    Acc acc[kBs];
#pragma unroll
    for (int k = 0; k < kBs; ++k) acc[k] = Acc(0);
    for (int32_t c = 0; c < cnt[o]; ++c) {
      const T *blk = rowbuf + (size_t)slist[first + c] * kBs;
#pragma unroll
      for (int k = 0; k < kBs; ++k) acc[k] += Acc(blk[k]);
    }
#pragma unroll
    for (int k = 0; k < kBs; ++k) vals[(size_t)z * kBs + k] = old[o][k] + T(acc[k]);
  }
  if (Elem::kHasRowDat) {
    for (int32_t lr = threadIdx.x; lr < d.nr; lr += blockDim.x) {
      Acc r[Elem::kRd];
#pragma unroll
      for (int a = 0; a < Elem::kRd; ++a) r[a] = Acc(0);
      for (int32_t u = sinc[lr]; u < sinc[lr + 1]; ++u)
#pragma unroll
        for (int gg = 0; gg < Elem::kGroups; ++gg)
#pragma unroll
          for (int a = 0; a < Elem::kRd; ++a) r[a] += Acc(rsbuf[((size_t)gg * per_pass + u) * Elem::kRd + a]);
      T *rd = rowdat + (size_t)(d.r0 + lr) * Elem::kRd;
#pragma unroll
      for (int a = 0; a < Elem::kRd; ++a) rd[a] += T(r[a]);
    }
  }
}

template <class Elem>
int launch_rowgather_staged(int32_t start, int32_t end, const Elem &elem, const op2hip_mat *A,
                            const op2hip_exec *ex, typename Elem::T *rowdat, hipStream_t s) {
  OP2HIP_REQUIRE(ex && ex->batch_desc && ex->gather_count && ex->gather_list && ex->nbatches >= 0,
                 "staged row gather without its plan (batch descriptors and gather table)");
  OP2HIP_REQUIRE(A->inc_ent && A->inc_ptr, "row-gather launch without row incidences");
  OP2HIP_REQUIRE(!ex->subset, "row-gather assembly does not iterate Subsets");
  OP2HIP_REQUIRE(A->rdim == Elem::kRd && A->cdim == Elem::kCd,
                 "matrix block %dx%d, kernel assembles %dx%d", A->rdim, A->cdim,
                 Elem::kRd, Elem::kCd);
  if (end <= start || ex->nbatches == 0) return OP2HIP_OK;
  const int bs = block_size_of(ex, 256);
  OP2HIP_REQUIRE(bs > 0 && bs <= 256 && bs % (64 * Elem::kGroups) == 0,
                 "staged row gather: block size %d must be a multiple of %d", bs, 64 * Elem::kGroups);
  const int per_pass = bs / Elem::kGroups;
  OP2HIP_REQUIRE(ex->batch_max_inc <= per_pass,
                 "staged row gather: a batch has %d row incidences, at most %d fit", ex->batch_max_inc,
                 per_pass);
  const int32_t nout = (ex->batch_max_nnz + bs - 1) / bs;
  OP2HIP_REQUIRE(nout <= 4, "staged row gather: %d blocks in a batch, at most %d",
                 ex->batch_max_nnz, 4 * bs);
  const int kout = nout <= 1 ? 1 : (nout <= 2 ? 2 : 4);
  const size_t lds = StagedLayout<Elem>::bytes(per_pass, kout, ex->batch_max_rows);
  OP2HIP_REQUIRE(lds <= 64 * 1024, "row batch needs %zu bytes of LDS (max 65536)", lds);
  const auto *descs = static_cast<const BatchDesc *>(ex->batch_desc);
  if (kout == 1)
    rowgather_staged_kernel<Elem, 1><<<ex->nbatches, bs, lds, s>>>(
        start, end, elem, *A, descs, ex->nbatches, ex->gather_count, ex->gather_list, rowdat);
  else if (kout == 2)
    rowgather_staged_kernel<Elem, 2><<<ex->nbatches, bs, lds, s>>>(
        start, end, elem, *A, descs, ex->nbatches, ex->gather_count, ex->gather_list, rowdat);
  else
    rowgather_staged_kernel<Elem, 4><<<ex->nbatches, bs, lds, s>>>(
        start, end, elem, *A, descs, ex->nbatches, ex->gather_count, ex->gather_list, rowdat);
  OP2HIP_LAUNCHED();
  return OP2HIP_OK;
}

// ---------------------------------------------------------------------------
// Cell-gather row gather: no atomics, one element geometry per batch.  The
// batch's distinct elements (op2hip_plan_cell_gather) have their geometry
// computed once into LDS (Elem::kGeo doubles each); every output block of the
// batch belongs to one thread, which evaluates each element block (e, i, j)
// that lands on it from the staged geometry and adds them up in element
// order; the block is written once.  Blocks are handed to threads by
// contribution count (descending), so the few heavy blocks (vertex
// diagonals) share waves instead of stalling many.
//
// Elem must additionally provide
//   static constexpr int kGeo;                          // even
//   void geometry(int32_t e, double *geo) const;        // kGeo doubles
//   static void no_geometry(double *geo);               // element adds zero blocks
//   void block(const double *geo, int i, int j, T *blk) const;  // runtime i, j
//   void rowdat(int32_t col, const T *blk, Acc *rs) const;      // fused rows only

template <class Elem>
struct CellsLayout {
  // region A holds the staged geometry, then (after the blocks are summed) the
  // batch's new block values for the coalesced flush
  __host__ __device__ static size_t region_a(int32_t max_cells, int32_t max_nnz) {
    const size_t g = (size_t)max_cells * Elem::kGeo * sizeof(double);
    const size_t r = (size_t)max_nnz * Elem::kRd * Elem::kCd * sizeof(double);
    return ((g > r ? g : r) + 15) & ~(size_t)15;
  }
  static size_t bytes(int32_t max_cells, int32_t max_nnz, int32_t max_contrib, int32_t max_rows,
                      int kOut) {
    size_t b = region_a(max_cells, max_nnz);
    b += (((size_t)max_contrib * sizeof(uint16_t)) + 15) & ~(size_t)15;
    b += (size_t)kOut * 4 * sizeof(int32_t);
    if (Elem::kHasRowDat)
      b += (size_t)(max_rows + 2) * sizeof(int32_t) + (size_t)max_rows * Elem::kRd * sizeof(double);
    return b;
  }
};

template <class Elem, int kOut, int kPre>
__global__ void __launch_bounds__(256)
rowgather_cells_kernel(int32_t start, int32_t end, Elem elem, op2hip_mat A,
                       const BatchDesc *__restrict__ descs, int32_t nbatches,
                       const int32_t *__restrict__ cell_ptr, const int32_t *__restrict__ cells,
                       const uint8_t *__restrict__ gcount, const uint16_t *__restrict__ order,
                       const int32_t *__restrict__ contrib_ptr,
                       const uint16_t *__restrict__ contrib, int32_t max_cells, int32_t max_nnz,
                       int32_t max_contrib, typename Elem::T *__restrict__ rowdat) {
  using T = typename Elem::T;
  using Acc = typename AccumOf<T>::type;
  constexpr int kBs = Elem::kRd * Elem::kCd;
  extern __shared__ __align__(16) unsigned char smem[];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
  const int32_t b = xcd_contiguous(blockIdx.x, nbatches);
  const BatchDesc d = descs[b];
  const int32_t c0 = cell_ptr[b], nc = cell_ptr[b + 1] - c0;
  const int32_t g0 = contrib_ptr[b], ng = contrib_ptr[b + 1] - g0;
  T *__restrict__ vals = static_cast<T *>(A.values) + (size_t)d.z0 * kBs;
  const int32_t nval = d.nz * kBs;

  double *geo = reinterpret_cast<double *>(smem);
  double *res = geo;  // region A, reused after the sums
  size_t off = CellsLayout<Elem>::region_a(max_cells, max_nnz);
  uint16_t *clist = reinterpret_cast<uint16_t *>(smem + off);
  off += (((size_t)max_contrib * sizeof(uint16_t)) + 15) & ~(size_t)15;
  int32_t *wtot = reinterpret_cast<int32_t *>(smem + off);
  off += (size_t)kOut * 4 * sizeof(int32_t);
  int32_t *srow = reinterpret_cast<int32_t *>(smem + off);
  double *racc = reinterpret_cast<double *>(
      smem + off + (((size_t)(d.nr + 1) * sizeof(int32_t) + 7) & ~(size_t)7));

  // old values of the batch, coalesced, in flight during the element work
  T old[kPre];
#pragma unroll
  for (int u = 0; u < kPre; ++u) {
    const int32_t k = threadIdx.x + u * (int32_t)blockDim.x;
    old[u] = k < nval ? vals[k] : T(0);
  }
  int32_t z[kOut], cnt[kOut];
#pragma unroll
  for (int o = 0; o < kOut; ++o) {
    const int32_t w = threadIdx.x + o * (int32_t)blockDim.x;
    z[o] = w < d.nz ? int32_t(order[(size_t)d.z0 + w]) : 0;
    cnt[o] = w < d.nz ? int32_t(gcount[(size_t)d.z0 + z[o]]) : 0;
  }
  for (int32_t k = threadIdx.x; k < ng; k += blockDim.x) clist[k] = contrib[(size_t)g0 + k];
  if (Elem::kHasRowDat) {
    for (int32_t k = threadIdx.x; k <= d.nr; k += blockDim.x) srow[k] = A.rowptr[d.r0 + k] - d.z0;
    for (int32_t k = threadIdx.x; k < d.nr * Elem::kRd; k += blockDim.x) racc[k] = 0.0;
  }
  for (int32_t u = threadIdx.x; u < nc; u += blockDim.x) {
    const int32_t e = cells[(size_t)c0 + u];
    if (e >= start && e < end) elem.geometry(e, geo + (size_t)u * Elem::kGeo);
    else Elem::no_geometry(geo + (size_t)u * Elem::kGeo);
  }

  // exclusive scan of the counts in thread order
  int32_t incl[kOut];
#pragma unroll
  for (int o = 0; o < kOut; ++o) {
    int32_t x = cnt[o];
#pragma unroll
    for (int dd = 1; dd < 64; dd <<= 1) {
      const int32_t y = __shfl_up(x, dd, 64);
      if (lane >= dd) x += y;
    }
    incl[o] = x;
    if (lane == 63) wtot[o * 4 + wave] = x;
  }
  __syncthreads();

  Acc acc[kOut][kBs];
  int32_t base = 0;
#pragma unroll
  for (int o = 0; o < kOut; ++o) {
    int32_t before = 0, all = 0;
    for (int w = 0; w < nwaves; ++w) {
      const int32_t v = wtot[o * 4 + w];
      before += w < wave ? v : 0;
      all += v;
    }
    const int32_t first = base + before + incl[o] - cnt[o];
    base += all;
#pragma unroll
    for (int k = 0; k < kBs; ++k) acc[o][k] = Acc(0);
    for (int32_t c = 0; c < cnt[o]; ++c) {
      const uint32_t code = clist[first + c];
      T blk[kBs];
      elem.block(geo + (size_t)(code >> 8) * Elem::kGeo, int((code >> 4) & 15u), int(code & 15u),
                 blk);
#pragma unroll
      for (int k = 0; k < kBs; ++k) acc[o][k] += Acc(blk[k]);
    }
  }
  __syncthreads();  // geometry no longer needed: region A takes the sums

#pragma unroll
  for (int o = 0; o < kOut; ++o) {
    const int32_t w = threadIdx.x + o * (int32_t)blockDim.x;
    if (w >= d.nz) continue;
#pragma unroll
    for (int k = 0; k < kBs; ++k) res[(size_t)z[o] * kBs + k] = acc[o][k];
    if (Elem::kHasRowDat) {
      // the row's residual gets this block's new contribution times u[col]
      int32_t lo = 0, hi = d.nr;
      while (hi - lo > 1) {
        const int32_t mid = (lo + hi) >> 1;
        if (srow[mid] <= z[o]) lo = mid; else hi = mid;
      }
      T add[kBs];
#pragma unroll
      for (int k = 0; k < kBs; ++k) add[k] = T(acc[o][k]);
      Acc rs[Elem::kRd];
#pragma unroll
      for (int a = 0; a < Elem::kRd; ++a) rs[a] = Acc(0);
      elem.rowdat(A.colidx[(size_t)d.z0 + z[o]], add, rs);
#pragma unroll
      for (int a = 0; a < Elem::kRd; ++a) atomic_add(&racc[lo * Elem::kRd + a], rs[a]);
    }
  }
  __syncthreads();

#pragma unroll
  for (int u = 0; u < kPre; ++u) {
    const int32_t k = threadIdx.x + u * (int32_t)blockDim.x;
    if (k < nval) vals[k] = old[u] + T(res[k]);
  }
  for (int32_t k0 = threadIdx.x + kPre * (int32_t)blockDim.x; k0 < nval; k0 += 8 * (int32_t)blockDim.x) {
    T v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int32_t k = k0 + u * (int32_t)blockDim.x;
      v[u] = k < nval ? vals[k] : T(0);
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int32_t k = k0 + u * (int32_t)blockDim.x;
      if (k < nval) vals[k] = v[u] + T(res[k]);
    }
  }
  if (Elem::kHasRowDat) {
    T *rd = rowdat + (size_t)d.r0 * Elem::kRd;
    for (int32_t k = threadIdx.x; k < d.nr * Elem::kRd; k += blockDim.x) rd[k] += T(racc[k]);
  }
}

template <class Elem>
int launch_rowgather_cells(int32_t start, int32_t end, const Elem &elem, const op2hip_mat *A,
                           const op2hip_exec *ex, typename Elem::T *rowdat, hipStream_t s) {
  OP2HIP_REQUIRE(ex && ex->batch_desc && ex->cell_ptr && ex->cells && ex->gather_count &&
                     ex->block_order && ex->contrib_ptr && ex->contrib && ex->nbatches >= 0,
                 "cell-gather row gather without its plan");
  OP2HIP_REQUIRE(!ex->subset, "row-gather assembly does not iterate Subsets");
  OP2HIP_REQUIRE(A->rdim == Elem::kRd && A->cdim == Elem::kCd,
                 "matrix block %dx%d, kernel assembles %dx%d", A->rdim, A->cdim,
                 Elem::kRd, Elem::kCd);
  if (end <= start || ex->nbatches == 0) return OP2HIP_OK;
  const int bs = block_size_of(ex, 256);
  OP2HIP_REQUIRE(bs > 0 && bs <= 256 && bs % 64 == 0, "row-gather block size %d", bs);
  OP2HIP_REQUIRE(ex->max_batch_cells > 0 && ex->max_batch_cells <= 256 && ex->max_batch_contrib >= 0,
                 "cell-gather plan: bad maxima");
  const int32_t nout = (ex->batch_max_nnz + bs - 1) / bs;
  OP2HIP_REQUIRE(nout <= 4, "cell-gather row gather: %d blocks in a batch, at most %d",
                 ex->batch_max_nnz, 4 * bs);
  const int kout = nout <= 1 ? 1 : (nout <= 2 ? 2 : 4);
  const size_t lds = CellsLayout<Elem>::bytes(ex->max_batch_cells, ex->batch_max_nnz,
                                              ex->max_batch_contrib, ex->batch_max_rows, kout);
  OP2HIP_REQUIRE(lds <= 64 * 1024, "row batch needs %zu bytes of LDS (max 65536)", lds);
  const auto *descs = static_cast<const BatchDesc *>(ex->batch_desc);
  // prefetch depth: the batch's values per thread, at most 24
  const int64_t per_thread = ((int64_t)ex->batch_max_nnz * Elem::kRd * Elem::kCd + bs - 1) / bs;
#define OP2HIP_CELLS(K, P)                                                                     \
  rowgather_cells_kernel<Elem, K, P><<<ex->nbatches, bs, lds, s>>>(                            \
      start, end, elem, *A, descs, ex->nbatches, ex->cell_ptr, ex->cells, ex->gather_count,    \
      ex->block_order, ex->contrib_ptr, ex->contrib, ex->max_batch_cells, ex->batch_max_nnz,   \
      ex->max_batch_contrib, rowdat)
  if (per_thread <= 8) {
    if (kout == 1) OP2HIP_CELLS(1, 8);
    else if (kout == 2) OP2HIP_CELLS(2, 8);
    else OP2HIP_CELLS(4, 8);
  } else if (per_thread <= 16) {
    if (kout == 1) OP2HIP_CELLS(1, 16);
    else if (kout == 2) OP2HIP_CELLS(2, 16);
    else OP2HIP_CELLS(4, 16);
  } else {
    if (kout == 1) OP2HIP_CELLS(1, 24);
    else if (kout == 2) OP2HIP_CELLS(2, 24);
    else OP2HIP_CELLS(4, 24);
  }
#undef OP2HIP_CELLS
  OP2HIP_LAUNCHED();
  return OP2HIP_OK;
}

// ---------------------------------------------------------------------------
// Row-list row gather: the slot-table row gather over batches of rows taken
// in an element-local order (op2hip_plan_row_list) rather than consecutive
// row ids.  With a scattered row numbering (a permuted mesh) a batch of
// consecutive rows touches elements all over the mesh and every coordinate
// gather misses the cache; rows ordered by their first element keep a
// batch's elements, and so its gathers, local.  The price is that the
// batch's rows are scattered in the Mat: the old values are gathered and the
// new ones written row segment by row segment.

template <class Elem>
struct RowListLayout {
  static size_t bytes(int32_t max_nnz, int32_t max_rows) {
    size_t b = (size_t)max_nnz * Elem::kRd * Elem::kCd * sizeof(double);
    b += (Elem::kHasRowDat ? (size_t)max_rows * Elem::kRd : 0) * sizeof(double);
    b += (size_t)(max_rows + 1) * sizeof(int32_t) * 3;
    return b;
  }
};

// Position in the Mat's values of batch value k (binary search of the row
// segments staged in LDS: seg[lr] first block of row lr in the batch,
// gbase[lr] its first block in the Mat; seg[nr] = nz).
template <int kBs>
__device__ __forceinline__ size_t rowlist_index(int32_t k, const int32_t *seg,
                                                const int32_t *gbase, int32_t nr) {
  const int32_t q = k / kBs, comp = k - q * kBs;
  int32_t lo = 0, hi = nr;
  while (hi - lo > 1) {
    const int32_t mid = (lo + hi) >> 1;
    if (seg[mid] <= q) lo = mid; else hi = mid;
  }
  return (size_t)(gbase[lo] + q - seg[lo]) * kBs + comp;
}

template <class Elem, int kPre>
__global__ void __launch_bounds__(256)
rowgather_rowlist_kernel(int32_t start, int32_t end, Elem elem, op2hip_mat A,
                         const BatchDesc *__restrict__ descs, int32_t nbatches,
                         const int2 *__restrict__ row_order, const int32_t *__restrict__ bent,
                         const SlotEntry<Elem::kCar> *__restrict__ table,
                         const int32_t *__restrict__ rbase, const uint64_t *__restrict__ packed,
                         typename Elem::T *__restrict__ rowdat) {
  using T = typename Elem::T;
  using Acc = typename AccumOf<T>::type;
  constexpr int kBs = Elem::kRd * Elem::kCd;
  extern __shared__ __align__(16) unsigned char smem[];
  const BatchDesc d = descs[xcd_contiguous(blockIdx.x, nbatches)];
  T *__restrict__ vals = static_cast<T *>(A.values);
  const int32_t nval = d.nz * kBs;
  Acc *acc = reinterpret_cast<Acc *>(smem);
  Acc *racc = acc + nval;
  int32_t *seg = reinterpret_cast<int32_t *>(racc + (Elem::kHasRowDat ? d.nr * Elem::kRd : 0));
  int32_t *gbase = seg + d.nr + 1;
  int32_t *rows = gbase + d.nr + 1;

  // the plan's row bases are contiguous per batch (no scattered rowptr read)
  for (int32_t lr = threadIdx.x; lr < d.nr; lr += blockDim.x) {
    const int2 p = row_order[d.r0 + lr];
    seg[lr] = p.y;
    gbase[lr] = rbase ? rbase[d.r0 + lr] : A.rowptr[p.x];
    rows[lr] = p.x;
  }
  if (threadIdx.x == 0) seg[d.nr] = d.nz;
  for (int32_t k = threadIdx.x; k < nval; k += blockDim.x) acc[k] = Acc(0);
  if (Elem::kHasRowDat)
    for (int32_t k = threadIdx.x; k < d.nr * Elem::kRd; k += blockDim.x) racc[k] = Acc(0);
  __syncthreads();

  T old[kPre];
  size_t gidx[kPre];
#pragma unroll
  for (int u = 0; u < kPre; ++u) {
    const int32_t k = threadIdx.x + u * (int32_t)blockDim.x;
    gidx[u] = k < nval ? rowlist_index<kBs>(k, seg, gbase, d.nr) : 0;
    old[u] = k < nval ? vals[gidx[u]] : T(0);
  }

  for (int32_t t0 = threadIdx.x; t0 < d.ni; t0 += blockDim.x) {
    const int32_t t = SpreadOf<Elem>::value ? spread_incidence(t0, d.ni) : t0;
    int32_t ent;
    SlotEntry<Elem::kCar> se;
    if constexpr (Elem::kCar == 3 && !Elem::kHasRowDat) {
      // the packed words carry no row index (se.lr), which only a row Dat reads
      if (packed) {
        // one word: incidence, row offset and three 5-bit slots (31: none)
        const uint64_t w = packed[d.i0 + t];
        ent = (int32_t)(uint32_t)w;
        se.row_off = (uint16_t)(w >> 32);
        se.lr = 0;
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          const uint32_t s = (uint32_t)(w >> (48 + 5 * j)) & 31u;
          se.slot[j] = s == 31u ? uint8_t(255) : uint8_t(s);
        }
      } else {
        ent = bent[d.i0 + t];
        se = table[d.i0 + t];
      }
    } else {
      ent = bent[d.i0 + t];
      se = table[d.i0 + t];
    }
    const int32_t e = ent / Elem::kRar;
    if (e < start || e >= end) continue;
    const int i = ent - e * Elem::kRar;
    Acc *row = acc + (size_t)se.row_off * kBs;
    T rsum[Elem::kRd];
#pragma unroll
    for (int a = 0; a < Elem::kRd; ++a) rsum[a] = T(0);
    elem.row(e, i, [&](int j, int32_t col, const T *blk) {
      const int s = se.slot[j];
      if (col < 0) return;
      if (s == 255) {
        atomicAdd(A.status, 1);
        return;
      }
#pragma unroll
      for (int k = 0; k < kBs; ++k) atomic_add(&row[s * kBs + k], Acc(blk[k]));
    }, rsum);
    if (Elem::kHasRowDat) {
#pragma unroll
      for (int a = 0; a < Elem::kRd; ++a) atomic_add(&racc[se.lr * Elem::kRd + a], Acc(rsum[a]));
    }
  }
  __syncthreads();

#pragma unroll
  for (int u = 0; u < kPre; ++u) {
    const int32_t k = threadIdx.x + u * (int32_t)blockDim.x;
    if (k < nval) vals[gidx[u]] = old[u] + T(acc[k]);
  }
  for (int32_t k = threadIdx.x + kPre * (int32_t)blockDim.x; k < nval; k += blockDim.x) {
    const size_t g = rowlist_index<kBs>(k, seg, gbase, d.nr);
    vals[g] += T(acc[k]);
  }
  if (Elem::kHasRowDat) {
    for (int32_t k = threadIdx.x; k < d.nr * Elem::kRd; k += blockDim.x) {
      const int32_t lr = k / Elem::kRd, a = k - lr * Elem::kRd;
      rowdat[(size_t)rows[lr] * Elem::kRd + a] += T(racc[k]);
    }
  }
}

template <class Elem>
int launch_rowgather_rowlist(int32_t start, int32_t end, const Elem &elem, const op2hip_mat *A,
                             const op2hip_exec *ex, typename Elem::T *rowdat, hipStream_t s) {
  OP2HIP_REQUIRE(ex && ex->batch_desc && ex->slot_table && ex->row_order && ex->batch_ent &&
                     ex->nbatches >= 0,
                 "row-list row gather without its plan");
  OP2HIP_REQUIRE(!ex->subset, "row-gather assembly does not iterate Subsets");
  OP2HIP_REQUIRE(A->rdim == Elem::kRd && A->cdim == Elem::kCd,
                 "matrix block %dx%d, kernel assembles %dx%d", A->rdim, A->cdim,
                 Elem::kRd, Elem::kCd);
  if (end <= start || ex->nbatches == 0) return OP2HIP_OK;
  const size_t lds = RowListLayout<Elem>::bytes(ex->batch_max_nnz, ex->batch_max_rows);
  OP2HIP_REQUIRE(lds <= 64 * 1024, "row batch needs %zu bytes of LDS (max 65536)", lds);
  const int bs = block_size_of(ex, 256);
  OP2HIP_REQUIRE(bs > 0 && bs <= 256 && bs % 64 == 0, "row-gather block size %d", bs);
  const auto *descs = static_cast<const BatchDesc *>(ex->batch_desc);
  const auto *ro = reinterpret_cast<const int2 *>(ex->row_order);
  const auto *table = static_cast<const SlotEntry<Elem::kCar> *>(ex->slot_table);
  const int64_t per_thread = ((int64_t)ex->batch_max_nnz * Elem::kRd * Elem::kCd + bs - 1) / bs;
  if (per_thread <= 8)
    rowgather_rowlist_kernel<Elem, 8><<<ex->nbatches, bs, lds, s>>>(
        start, end, elem, *A, descs, ex->nbatches, ro, ex->batch_ent, table, ex->row_base,
        Elem::kCar == 3 && Elem::kRar == 3 && !Elem::kHasRowDat ? ex->batch_packed : nullptr, rowdat);
  else
    rowgather_rowlist_kernel<Elem, 16><<<ex->nbatches, bs, lds, s>>>(
        start, end, elem, *A, descs, ex->nbatches, ro, ex->batch_ent, table, ex->row_base,
        Elem::kCar == 3 && Elem::kRar == 3 && !Elem::kHasRowDat ? ex->batch_packed : nullptr, rowdat);
  OP2HIP_LAUNCHED();
  return OP2HIP_OK;
}

}  // namespace op2hip
