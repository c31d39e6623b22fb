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
#pragma unroll
      for (int j = 0; j < kCar; ++j) {
        const int32_t c = cmap[(int64_t)e * kCar + j];
        const int32_t p = c < 0 ? -1 : find_slot(colidx, zlo, zhi, c);
        if (c >= 0 && p < 0) atomicAdd(bad, 1);
        s.slot[j] = (p < 0 || p - zlo > 254) ? uint8_t(255) : static_cast<uint8_t>(p - zlo);
      }
      table[t] = s;
    }
  }
}

template <class Elem, int kPre>
__global__ void __launch_bounds__(256)
rowgather_slots_kernel(int32_t start, int32_t end, Elem elem, op2hip_mat A,
                       const BatchDesc *__restrict__ descs, int32_t nbatches,
                       const SlotEntry<Elem::kCar> *__restrict__ table,
                       typename Elem::T *__restrict__ rowdat) {
  using T = typename Elem::T;
  using Acc = typename AccumOf<T>::type;
  constexpr int kBs = Elem::kRd * Elem::kCd;
  extern __shared__ __align__(16) unsigned char smem[];
  const BatchDesc d = descs[xcd_contiguous(blockIdx.x, nbatches)];
  T *__restrict__ vals = static_cast<T *>(A.values) + (size_t)d.z0 * kBs;
  const int32_t nval = d.nz * kBs;

  T old[kPre];
#pragma unroll
  for (int u = 0; u < kPre; ++u) {
    const int32_t k = threadIdx.x + u * (int32_t)blockDim.x;
    old[u] = k < nval ? vals[k] : T(0);
  }
  Acc *acc = reinterpret_cast<Acc *>(smem);
  Acc *racc = acc + nval;
  for (int32_t k = threadIdx.x; k < nval; k += blockDim.x) acc[k] = Acc(0);
  if (Elem::kHasRowDat)
    for (int32_t k = threadIdx.x; k < d.nr * Elem::kRd; k += blockDim.x) racc[k] = Acc(0);
  __syncthreads();

  for (int32_t t = threadIdx.x; t < d.ni; t += blockDim.x) {
    const int32_t ent = A.inc_ent[d.i0 + t];
    const SlotEntry<Elem::kCar> se = table[d.i0 + t];
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
    if (k < nval) vals[k] = old[u] + T(acc[k]);
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
  rowgather_slots_kernel<Elem, 8><<<ex->nbatches, bs, lds, s>>>(
      start, end, elem, *A, static_cast<const BatchDesc *>(ex->batch_desc), ex->nbatches,
      static_cast<const SlotEntry<Elem::kCar> *>(ex->slot_table), rowdat);
  OP2HIP_LAUNCHED();
  return OP2HIP_OK;
}

}  // namespace op2hip
