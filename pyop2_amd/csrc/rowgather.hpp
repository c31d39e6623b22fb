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

// Elem must provide:
//   static constexpr int kRar, kCar, kRd, kCd;      (arities, block shape)
//   static constexpr bool kHasRowDat;               (fused Dat INC on rows)
//   using T;                                        (value type)
//   __device__ void row(int32_t e, int i, Emit&&)   (calls emit(j, col, blk, rowdat))
template <class Elem>
struct RowGatherLayout {
  using T = typename Elem::T;
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

template <class Elem>
__global__ void __launch_bounds__(256)
rowgather_kernel(int32_t start, int32_t end, Elem elem, op2hip_mat A,
                 const int32_t *__restrict__ batch_rows, int32_t nbatches,
                 typename Elem::T *__restrict__ rowdat) {
  using T = typename Elem::T;
  constexpr int kBs = Elem::kRd * Elem::kCd;
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
  int32_t *srow = scol + nz;          // nr+1 row offsets into scol
  int32_t *sinc = srow + nr + 1;      // nr+1 incidence offsets

  for (int32_t k = threadIdx.x; k < nz * kBs; k += blockDim.x) acc[k] = T(0);
  if (Elem::kHasRowDat)
    for (int32_t k = threadIdx.x; k < nr * Elem::kRd; k += blockDim.x) racc[k] = T(0);
  for (int32_t k = threadIdx.x; k < nz; k += blockDim.x) scol[k] = A.colidx[z0 + k];
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
      for (int k = 0; k < kBs; ++k) atomic_add(&acc[slot * kBs + k], blk[k]);
    }, rsum);
    if (Elem::kHasRowDat) {
#pragma unroll
      for (int a = 0; a < Elem::kRd; ++a) atomic_add(&racc[lr * Elem::kRd + a], rsum[a]);
    }
  }
  __syncthreads();

  T *vals = static_cast<T *>(A.values) + (size_t)z0 * kBs;
  for (int32_t k = threadIdx.x; k < nz * kBs; k += blockDim.x) vals[k] += acc[k];
  if (Elem::kHasRowDat) {
    T *rd = rowdat + (size_t)r0 * Elem::kRd;
    for (int32_t k = threadIdx.x; k < nr * Elem::kRd; k += blockDim.x) rd[k] += racc[k];
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
  rowgather_kernel<Elem><<<ex->nbatches, bs, lds, s>>>(start, end, elem, *A, ex->batch_rows,
                                                        ex->nbatches, rowdat);
  OP2HIP_LAUNCHED();
  return OP2HIP_OK;
}

}  // namespace op2hip
