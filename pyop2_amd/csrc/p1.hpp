// P1 triangle element math shared by the single-mesh loops (p1.hip) and the
// batched launch (batch.hip).
//
// Element math (same operations as the C-string kernels in
// pyop2_amd/elements.py): with b_i = y_{i+1} - y_{i+2},
// c_i = x_{i+2} - x_{i+1} (cyclic) and det = 2 |T| (signed),
// K_e[i][j] = (b_i b_j + c_i c_j) / (2 |det|), lumped mass |det| / 6.
#pragma once

#include "rowgather.hpp"
#include "staged.hpp"

namespace op2hip {

struct P1Geom {
  double b[3], c[3], s, det;
};

__device__ __forceinline__ P1Geom p1_geom_of(double2 p0, double2 p1, double2 p2) {
  P1Geom g;
  g.b[0] = p1.y - p2.y;
  g.b[1] = p2.y - p0.y;
  g.b[2] = p0.y - p1.y;
  g.c[0] = p2.x - p1.x;
  g.c[1] = p0.x - p2.x;
  g.c[2] = p1.x - p0.x;
  g.det = (p1.x - p0.x) * (p2.y - p0.y) - (p2.x - p0.x) * (p1.y - p0.y);
  g.s = 1.0 / (2.0 * fabs(g.det));
  return g;
}

__device__ __forceinline__ P1Geom p1_geom(const double *__restrict__ X, int32_t v0,
                                          int32_t v1, int32_t v2) {
  const double2 *X2 = reinterpret_cast<const double2 *>(X);
  return p1_geom_of(X2[v0], X2[v1], X2[v2]);
}

// Staged lumped mass: coordinates staged, M incremented (map cell -> vertex).
struct P1MassStaged {
  using T = double;
  static constexpr int kMaps = 1;
  static constexpr int kAr0 = 3, kIn0 = 2, kOut0 = 1;
  static constexpr int kAr1 = 0, kIn1 = 0, kOut1 = 0;
  double *M;
  const double *X;
  template <int Mp>
  __device__ __forceinline__ void stage(int32_t v, double *dst) const {
    dst[0] = X[2 * (size_t)v];
    dst[1] = X[2 * (size_t)v + 1];
  }
  template <int Mp>
  __host__ __device__ __forceinline__ double *out() const { return M; }
  __device__ __forceinline__ void element(int32_t, const double (&x)[3][2], const double (&)[1][1],
                                          double (&y)[3][1], double (&)[1][1]) const {
    const P1Geom g = p1_geom_of(make_double2(x[0][0], x[0][1]), make_double2(x[1][0], x[1][1]),
                                make_double2(x[2][0], x[2][1]));
    const double m = fabs(g.det) / 6.0;
    y[0][0] = m;
    y[1][0] = m;
    y[2][0] = m;
  }
};

// Staged Laplace residual: (x, y, u) staged per vertex, R incremented.
struct P1ResidualStaged {
  using T = double;
  static constexpr int kMaps = 1;
  static constexpr int kAr0 = 3, kIn0 = 3, kOut0 = 1;
  static constexpr int kAr1 = 0, kIn1 = 0, kOut1 = 0;
  double *R;
  const double *X;
  const double *u;
  template <int Mp>
  __device__ __forceinline__ void stage(int32_t v, double *dst) const {
    dst[0] = X[2 * (size_t)v];
    dst[1] = X[2 * (size_t)v + 1];
    dst[2] = u[v];
  }
  template <int Mp>
  __host__ __device__ __forceinline__ double *out() const { return R; }
  __device__ __forceinline__ void element(int32_t, const double (&x)[3][3], const double (&)[1][1],
                                          double (&y)[3][1], double (&)[1][1]) const {
    const P1Geom g = p1_geom_of(make_double2(x[0][0], x[0][1]), make_double2(x[1][0], x[1][1]),
                                make_double2(x[2][0], x[2][1]));
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      double acc = 0.0;
#pragma unroll
      for (int j = 0; j < 3; ++j) acc += (g.b[i] * g.b[j] + g.c[i] * g.c[j]) * g.s * x[j][2];
      y[i][0] = acc;
    }
  }
};

// Row-gather element: row i of K_e.
struct P1MatrixRow {
  using T = double;
  static constexpr int kRar = 3, kCar = 3, kRd = 1, kCd = 1;
  static constexpr bool kHasRowDat = false;
  static constexpr int kGroups = 1;
  const double *X;
  const int32_t *c2v;
  template <class Emit>
  __device__ __forceinline__ void row(int32_t e, int i, Emit &&emit, double *) const {
    const int32_t v0 = c2v[3 * e], v1 = c2v[3 * e + 1], v2 = c2v[3 * e + 2];
    const P1Geom g = p1_geom(X, v0, v1, v2);
    // select row i with constant indices (a runtime index into the
    // coefficient arrays would place them in scratch memory)
    const double bi = i == 0 ? g.b[0] : (i == 1 ? g.b[1] : g.b[2]);
    const double ci = i == 0 ? g.c[0] : (i == 1 ? g.c[1] : g.c[2]);
    const double k0 = (bi * g.b[0] + ci * g.c[0]) * g.s;
    const double k1 = (bi * g.b[1] + ci * g.c[1]) * g.s;
    const double k2 = (bi * g.b[2] + ci * g.c[2]) * g.s;
    emit(0, v0, &k0);
    emit(1, v1, &k1);
    emit(2, v2, &k2);
  }
  template <int G, class Emit>
  __device__ __forceinline__ void row_group(int32_t e, int i, Emit &&emit, double *rs) const {
    row(e, i, emit, rs);
  }
};

// Cell-gather element: per element (b_0..2, c_0..2, 1 / (2 |det|)), block
// (i, j) = (b_i b_j + c_i c_j) / (2 |det|).
struct P1MatrixCells {
  using T = double;
  static constexpr int kRar = 3, kCar = 3, kRd = 1, kCd = 1;
  static constexpr bool kHasRowDat = false;
  static constexpr int kGeo = 8;
  const double *X;
  const int32_t *c2v;
  __device__ __forceinline__ void geometry(int32_t e, double *geo) const {
    const P1Geom g = p1_geom(X, c2v[3 * e], c2v[3 * e + 1], c2v[3 * e + 2]);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      geo[k] = g.b[k];
      geo[3 + k] = g.c[k];
    }
    geo[6] = g.s;
    geo[7] = 0.0;
  }
  __device__ static void no_geometry(double *geo) {
#pragma unroll
    for (int k = 0; k < kGeo; ++k) geo[k] = 0.0;
  }
  __device__ __forceinline__ void block(const double *geo, int i, int j, double *blk) const {
    blk[0] = (geo[i] * geo[j] + geo[3 + i] * geo[3 + j]) * geo[6];
  }
  __device__ __forceinline__ void rowdat(int32_t, const double *, double *) const {}
};

// Lane-per-block row-gather element: lane j of a group computes K_e[i][j].
struct P1MatrixLanes {
  using T = double;
  using Geom = P1Geom;
  static constexpr int kRar = 3, kCar = 3, kRd = 1, kCd = 1;
  static constexpr bool kHasRowDat = false;
  const double *X;
  const int32_t *c2v;
  __device__ __forceinline__ Geom geometry(int32_t e) const {
    return p1_geom(X, c2v[3 * e], c2v[3 * e + 1], c2v[3 * e + 2]);
  }
  __device__ __forceinline__ int32_t col(int32_t e, int j) const { return c2v[3 * e + j]; }
  __device__ __forceinline__ void block(const Geom &g, int i, int j, double *blk) const {
    const double bi = i == 0 ? g.b[0] : (i == 1 ? g.b[1] : g.b[2]);
    const double ci = i == 0 ? g.c[0] : (i == 1 ? g.c[1] : g.c[2]);
    const double bj = j == 0 ? g.b[0] : (j == 1 ? g.b[1] : g.b[2]);
    const double cj = j == 0 ? g.c[0] : (j == 1 ? g.c[1] : g.c[2]);
    blk[0] = (bi * bj + ci * cj) * g.s;
  }
  __device__ __forceinline__ void rowdat(int32_t, const double *, double *) const {}
};


}  // namespace op2hip
