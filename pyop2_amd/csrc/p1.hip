// P1 triangle element loops: lumped mass, Laplace residual, Laplace matrix.
#include "p1.hpp"

namespace op2hip {

template <bool kAtomic>
__device__ __forceinline__ void scatter_add(double *p, double v) {
  if (kAtomic) atomic_add(p, v); else plain_add(p, v);
}

template <bool kAtomic>
__global__ void __launch_bounds__(256)
p1_mass_kernel(int32_t lo, int32_t hi, const int32_t *__restrict__ list,
               double *__restrict__ M, const double *__restrict__ X,
               const int32_t *__restrict__ c2v) {
  const int32_t t = lo + blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= hi) return;
  const int32_t e = element_of(t, list);
  const int32_t v0 = c2v[3 * e], v1 = c2v[3 * e + 1], v2 = c2v[3 * e + 2];
  const P1Geom g = p1_geom(X, v0, v1, v2);
  const double m = fabs(g.det) / 6.0;
  scatter_add<kAtomic>(&M[v0], m);
  scatter_add<kAtomic>(&M[v1], m);
  scatter_add<kAtomic>(&M[v2], m);
}

template <bool kAtomic>
__global__ void __launch_bounds__(256)
p1_residual_kernel(int32_t lo, int32_t hi, const int32_t *__restrict__ list,
                   double *__restrict__ R, const double *__restrict__ X,
                   const double *__restrict__ u, const int32_t *__restrict__ c2v) {
  const int32_t t = lo + blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= hi) return;
  const int32_t e = element_of(t, list);
  const int32_t v[3] = {c2v[3 * e], c2v[3 * e + 1], c2v[3 * e + 2]};
  const P1Geom g = p1_geom(X, v[0], v[1], v[2]);
  const double uu[3] = {u[v[0]], u[v[1]], u[v[2]]};
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    double acc = 0.0;
#pragma unroll
    for (int j = 0; j < 3; ++j) acc += (g.b[i] * g.b[j] + g.c[i] * g.c[j]) * g.s * uu[j];
    scatter_add<kAtomic>(&R[v[i]], acc);
  }
}

// Cell-centric matrix assembly: slot of each (i, j) by binary search of the
// row's sorted colidx, then an atomic (or, coloured, plain) add.
template <bool kAtomic>
__global__ void __launch_bounds__(256)
p1_matrix_kernel(int32_t lo, int32_t hi, const int32_t *__restrict__ list,
                 op2hip_mat A, const double *__restrict__ X,
                 const int32_t *__restrict__ c2v) {
  const int32_t t = lo + blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= hi) return;
  const int32_t e = element_of(t, list);
  const int32_t v[3] = {c2v[3 * e], c2v[3 * e + 1], c2v[3 * e + 2]};
  const P1Geom g = p1_geom(X, v[0], v[1], v[2]);
  double *vals = static_cast<double *>(A.values);
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    if (v[i] < 0) continue;
    const int32_t rlo = A.rowptr[v[i]], rhi = A.rowptr[v[i] + 1];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      if (v[j] < 0) continue;
      const int32_t slot = find_slot(A.colidx, rlo, rhi, v[j]);
      if (slot < 0) {
        atomicAdd(A.status, 1);
        continue;
      }
      scatter_add<kAtomic>(&vals[slot], (g.b[i] * g.b[j] + g.c[i] * g.c[j]) * g.s);
    }
  }
}

}  // namespace op2hip

using namespace op2hip;

extern "C" {

int op2hip_wrap_p1_lumped_mass(int32_t start, int32_t end, double *M, const double *X,
                               const int32_t *c2v, const op2hip_exec *ex, void *stream) {
  hipStream_t s = stream_of(stream);
  if (ex && ex->strategy == OP2HIP_STAGED) return launch_staged(start, end, P1MassStaged{M, X}, ex, s);
  const int bs = block_size_of(ex, 256);
  return for_each_launch(ex, start, end, [&](int32_t lo, int32_t hi, const int32_t *list,
                                             bool atomic) -> int {
    if (atomic) p1_mass_kernel<true><<<ceil_div(hi - lo, bs), bs, 0, s>>>(lo, hi, list, M, X, c2v);
    else p1_mass_kernel<false><<<ceil_div(hi - lo, bs), bs, 0, s>>>(lo, hi, list, M, X, c2v);
    OP2HIP_LAUNCHED();
    return OP2HIP_OK;
  });
}

int op2hip_wrap_p1_laplace_residual(int32_t start, int32_t end, double *R, const double *X,
                                    const double *u, const int32_t *c2v,
                                    const op2hip_exec *ex, void *stream) {
  hipStream_t s = stream_of(stream);
  if (ex && ex->strategy == OP2HIP_STAGED)
    return launch_staged(start, end, P1ResidualStaged{R, X, u}, ex, s);
  const int bs = block_size_of(ex, 256);
  return for_each_launch(ex, start, end, [&](int32_t lo, int32_t hi, const int32_t *list,
                                             bool atomic) -> int {
    if (atomic)
      p1_residual_kernel<true><<<ceil_div(hi - lo, bs), bs, 0, s>>>(lo, hi, list, R, X, u, c2v);
    else
      p1_residual_kernel<false><<<ceil_div(hi - lo, bs), bs, 0, s>>>(lo, hi, list, R, X, u, c2v);
    OP2HIP_LAUNCHED();
    return OP2HIP_OK;
  });
}

int op2hip_wrap_p1_laplace_matrix(int32_t start, int32_t end, const op2hip_mat *A,
                                  const double *X, const int32_t *c2v,
                                  const op2hip_exec *ex, void *stream) {
  hipStream_t s = stream_of(stream);
  OP2HIP_REQUIRE(A && A->rowptr && A->colidx && A->values && A->status, "p1 matrix: bad Mat");
  OP2HIP_REQUIRE(A->rdim == 1 && A->cdim == 1, "p1 matrix: Mat block must be 1x1");
  if (ex && ex->strategy == OP2HIP_ROWGATHER) {
    P1MatrixLanes elem{X, c2v};
    return launch_rowgather_lanes(start, end, elem, A, ex, static_cast<double *>(nullptr), s);
  }
  if (ex && ex->strategy == OP2HIP_ROWGATHER_SLOTS) {
    P1MatrixRow elem{X, c2v};
    return launch_rowgather_slots(start, end, elem, A, ex, static_cast<double *>(nullptr), s);
  }
  if (ex && ex->strategy == OP2HIP_ROWGATHER_STAGED) {
    P1MatrixRow elem{X, c2v};
    return launch_rowgather_staged(start, end, elem, A, ex, static_cast<double *>(nullptr), s);
  }
  if (ex && ex->strategy == OP2HIP_ROWGATHER_ROWLIST) {
    P1MatrixRow elem{X, c2v};
    return launch_rowgather_rowlist(start, end, elem, A, ex, static_cast<double *>(nullptr), s);
  }
  if (ex && ex->strategy == OP2HIP_ROWGATHER_CELLS) {
    P1MatrixCells elem{X, c2v};
    return launch_rowgather_cells(start, end, elem, A, ex, static_cast<double *>(nullptr), s);
  }
  if (ex && ex->strategy == OP2HIP_ROWGATHER_THREAD) {
    P1MatrixRow elem{X, c2v};
    return launch_rowgather(start, end, elem, A, ex, static_cast<double *>(nullptr), s);
  }
  const int bs = block_size_of(ex, 256);
  const op2hip_mat Av = *A;
  return for_each_launch(ex, start, end, [&](int32_t lo, int32_t hi, const int32_t *list,
                                             bool atomic) -> int {
    if (atomic)
      p1_matrix_kernel<true><<<ceil_div(hi - lo, bs), bs, 0, s>>>(lo, hi, list, Av, X, c2v);
    else
      p1_matrix_kernel<false><<<ceil_div(hi - lo, bs), bs, 0, s>>>(lo, hi, list, Av, X, c2v);
    OP2HIP_LAUNCHED();
    return OP2HIP_OK;
  });
}

}  // extern "C"
