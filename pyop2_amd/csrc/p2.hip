// P2 tetrahedron, isotropic linear elasticity: residual, BSR(3) element-matrix
// assembly, fused residual + assembly, and the adjoint of the residual loop.
//
// Geometry: J[r][c] = X[c+1][r] - X[0][r], K = J^{-1} (K[m][k] = d xi_m/d x_k),
// barycentric gradients gl[0] = -(K[0]+K[1]+K[2]), gl[m+1] = K[m].
// Basis gradients at quadrature point q: grad phi_i = sum_m D[q][i][m] gl[m]
// with D the derivative of the P2 basis in barycentric coordinates
// (vertex i: (4 L_i - 1) at m = i; edge (a, b): 4 L_b at m = a, 4 L_a at m = b).
// 4-point Gauss rule, weight 1/24, so wdet = |det J| / 24.
//
// Residual: R_i += wdet sum_q sigma(u)_q grad phi_i, evaluated through the
// barycentric form  grad u_q = sum_m w_qm (x) gl_m,  w_qm = sum_j D[q][j][m] u_j,
// and  R_i += wdet sum_q sum_m D[q][i][m] sigma_q gl_m.
// Element matrix block (i, j): with M = sum_q grad phi_i (x) grad phi_j,
// B = wdet (lam M + mu M^T + mu tr(M) I).
#include <mutex>

#include "rowgather.hpp"
#include "staged.hpp"

namespace op2hip {
namespace p2 {

constexpr double kQa = 0.5854101966249685;
constexpr double kQb = 0.1381966011250105;
constexpr double kQw = 1.0 / 24.0;
constexpr int kEa[6] = {0, 0, 0, 1, 1, 2};
constexpr int kEb[6] = {1, 2, 3, 2, 3, 3};

__host__ __device__ constexpr double bary(int q, int m) { return q == m ? kQa : kQb; }

// d phi_i / d lambda_m at quadrature point q.
__host__ __device__ constexpr double dcoef(int q, int i, int m) {
  return i < 4 ? (m == i ? 4.0 * bary(q, i) - 1.0 : 0.0)
               : (m == kEa[i - 4] ? 4.0 * bary(q, kEb[i - 4])
                                  : (m == kEb[i - 4] ? 4.0 * bary(q, kEa[i - 4]) : 0.0));
}

struct DTable {
  double v[4][10][4];
  constexpr DTable() : v{} {
    for (int q = 0; q < 4; ++q)
      for (int i = 0; i < 10; ++i)
        for (int m = 0; m < 4; ++m) v[q][i][m] = dcoef(q, i, m);
  }
};
__constant__ DTable kD = DTable();

template <typename T>
struct Geom {
  T gl[4][3];
  T K[3][3];
  T wdet;
};

// Geometry of the tetrahedron with vertex coordinates xv[v][r].
template <typename T>
__device__ __forceinline__ Geom<T> geometry_of(const T (&xv)[4][3]) {
  T J[3][3];
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int r = 0; r < 3; ++r) J[r][c] = xv[c + 1][r] - xv[0][r];
  const T c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
  const T c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
  const T c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
  const T det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
  const T idet = T(1) / det;
  Geom<T> g;
  g.K[0][0] = c00 * idet;
  g.K[1][0] = c01 * idet;
  g.K[2][0] = c02 * idet;
  g.K[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * idet;
  g.K[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * idet;
  g.K[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * idet;
  g.K[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * idet;
  g.K[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * idet;
  g.K[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * idet;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    g.gl[0][k] = -(g.K[0][k] + g.K[1][k] + g.K[2][k]);
    g.gl[1][k] = g.K[0][k];
    g.gl[2][k] = g.K[1][k];
    g.gl[3][k] = g.K[2][k];
  }
  g.wdet = T(kQw) * fabs(det);
  return g;
}

template <typename T>
__device__ __forceinline__ Geom<T> geometry(const T *__restrict__ X,
                                            const int32_t *__restrict__ c2v, int32_t e) {
  T xv[4][3];
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int32_t n = c2v[4 * e + v];
#pragma unroll
    for (int r = 0; r < 3; ++r) xv[v][r] = X[3 * n + r];
  }
  return geometry_of(xv);
}

// Barycentric-form gradient of a P2 field at quadrature point q:
// grad[b][k] = sum_m (sum_j D[q][j][m] f[j][b]) gl[m][k].
template <typename T, int Q>
__device__ __forceinline__ void field_grad(const T (&f)[10][3], const Geom<T> &g, T (&grad)[3][3]) {
  T w[4][3];
#pragma unroll
  for (int m = 0; m < 4; ++m)
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      T acc = T(0);
#pragma unroll
      for (int j = 0; j < 10; ++j)
        if (dcoef(Q, j, m) != 0.0) acc += T(dcoef(Q, j, m)) * f[j][b];
      w[m][b] = acc;
    }
#pragma unroll
  for (int b = 0; b < 3; ++b)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      T acc = T(0);
#pragma unroll
      for (int m = 0; m < 4; ++m) acc += w[m][b] * g.gl[m][k];
      grad[b][k] = acc;
    }
}

template <typename T>
__device__ __forceinline__ void stress(const T (&gu)[3][3], T lam, T mu, T (&s)[3][3]) {
  const T dv = gu[0][0] + gu[1][1] + gu[2][2];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int k = 0; k < 3; ++k) s[a][k] = mu * (gu[a][k] + gu[k][a]) + (a == k ? lam * dv : T(0));
}

// out[i][a] += sum_m D[q][i][m] (s gl_m)[a]
template <typename T, int Q>
__device__ __forceinline__ void test_gradients(const T (&s)[3][3], const Geom<T> &g, T (&out)[10][3]) {
  T sm[4][3];
#pragma unroll
  for (int m = 0; m < 4; ++m)
#pragma unroll
    for (int a = 0; a < 3; ++a)
      sm[m][a] = s[a][0] * g.gl[m][0] + s[a][1] * g.gl[m][1] + s[a][2] * g.gl[m][2];
#pragma unroll
  for (int i = 0; i < 10; ++i)
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      T acc = out[i][a];
#pragma unroll
      for (int m = 0; m < 4; ++m)
        if (dcoef(Q, i, m) != 0.0) acc += T(dcoef(Q, i, m)) * sm[m][a];
      out[i][a] = acc;
    }
}


// Vertex-stress form of the element residual.  grad u is linear on the
// element, so the stress is linear too: s(L) = sum_n L_n s_n with s_n the
// stress of grad u at vertex n,
//   grad u(vertex n) = 4 (u_n (x) gl_n + sum_{b != n} u_nb (x) gl_b) - P,
//   P = sum_j u_j (x) gl_j  (vertices j).
// The 4-point rule integrates the degree-2 integrand exactly, and its
// moments sum_q L_qn L_qm = 0.2 (1 + [n == m]) turn the sum over the points
// into sums over vertices (S = sum_n s_n):
//   vertex i:     R_i  = 0.2 wdet (4 s_i - S) gl_i,
//   edge (a, b):  R_ab = 0.8 wdet ((S + s_b) gl_a + (S + s_a) gl_b).
// About 1000 flops against 1600 of the point-by-point form, and no array of
// the element's 30 values lives across the whole evaluation: u(j, c) reads
// node j's component c on demand, out(i, v) takes node i's three values.
// Symmetric stresses are kept as (xx, yy, zz, xy, xz, yz).
template <typename T>
__device__ __forceinline__ void sym_stress(const T (&G)[3][3], T lam, T mu, T (&s)[6]) {
  const T lt = lam * (G[0][0] + G[1][1] + G[2][2]);
  const T m2 = mu + mu;
  s[0] = m2 * G[0][0] + lt;
  s[1] = m2 * G[1][1] + lt;
  s[2] = m2 * G[2][2] + lt;
  s[3] = mu * (G[0][1] + G[1][0]);
  s[4] = mu * (G[0][2] + G[2][0]);
  s[5] = mu * (G[1][2] + G[2][1]);
}

template <typename T>
__device__ __forceinline__ void sym_apply(const T (&s)[6], const T (&v)[3], T (&o)[3]) {
  o[0] = s[0] * v[0] + s[3] * v[1] + s[4] * v[2];
  o[1] = s[3] * v[0] + s[1] * v[1] + s[5] * v[2];
  o[2] = s[4] * v[0] + s[5] * v[1] + s[2] * v[2];
}

// the local node of the edge joining vertices n and b (n != b)
constexpr int kEdgeOf[4][4] = {{-1, 4, 5, 6}, {4, -1, 7, 8}, {5, 7, -1, 9}, {6, 8, 9, -1}};

template <typename T, class U, class Out>
__device__ __forceinline__ void residual_vertex_form(const Geom<T> &g, T lam, T mu, U &&u, Out &&out) {
  // pass 1: the stress of P = sum_j u_j (x) gl_j
  T sP[6];
  {
    T P[3][3] = {};
#pragma unroll
    for (int n = 0; n < 4; ++n) {
      const T v0 = u(n, 0), v1 = u(n, 1), v2 = u(n, 2);
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        P[0][k] += v0 * g.gl[n][k];
        P[1][k] += v1 * g.gl[n][k];
        P[2][k] += v2 * g.gl[n][k];
      }
    }
    sym_stress(P, lam, mu, sP);
  }
  // pass 2: the vertex stresses s_n = 4 s(u_n (x) gl_n + sum_b u_nb (x) gl_b) - sP
  T sv[4][6];
#pragma unroll
  for (int n = 0; n < 4; ++n) {
    T G[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const T un = u(n, a);
#pragma unroll
      for (int k = 0; k < 3; ++k) G[a][k] = un * g.gl[n][k];
    }
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      if (b == n) continue;
      const int e = kEdgeOf[n][b];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const T ue = u(e, a);
#pragma unroll
        for (int k = 0; k < 3; ++k) G[a][k] += ue * g.gl[b][k];
      }
    }
    sym_stress(G, T(4) * lam, T(4) * mu, sv[n]);
#pragma unroll
    for (int c = 0; c < 6; ++c) sv[n][c] -= sP[c];
  }
  // pass 3: the ten increments, S = sum_n s_n
  T S[6];
#pragma unroll
  for (int c = 0; c < 6; ++c) S[c] = sv[0][c] + sv[1][c] + sv[2][c] + sv[3][c];
  const T cv = T(0.2) * g.wdet, ce = T(0.8) * g.wdet;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    T M[6], r[3];
#pragma unroll
    for (int c = 0; c < 6; ++c) M[c] = cv * (T(4) * sv[i][c] - S[c]);
    sym_apply(M, g.gl[i], r);
    out(i, r);
  }
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const int a = kEa[k], b = kEb[k];
    T Ma[6], Mb[6], ra[3], rb[3], r[3];
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      Ma[c] = ce * (S[c] + sv[b][c]);
      Mb[c] = ce * (S[c] + sv[a][c]);
    }
    sym_apply(Ma, g.gl[a], ra);
    sym_apply(Mb, g.gl[b], rb);
#pragma unroll
    for (int c = 0; c < 3; ++c) r[c] = ra[c] + rb[c];
    out(4 + k, r);
  }
}

// The element residual in the vertex-stress form (about 1000 flops; the
// point-by-point quadrature form -- field_grad, stress and test_gradients at
// each of the four points, as the adjoint evaluates it -- gives the same
// values to rounding).
template <typename T>
__device__ __forceinline__ void element_residual(const T (&uu)[10][3], const Geom<T> &g, T lam,
                                                 T mu, T (&re)[10][3]) {
  residual_vertex_form(
      g, lam, mu, [&](int j, int c) { return uu[j][c]; },
      [&](int i, const T (&r)[3]) {
#pragma unroll
        for (int c = 0; c < 3; ++c) re[i][c] = r[c];
      });
}

template <typename T, bool kAtomic>
__device__ __forceinline__ void add3(T *p, const T (&v)[3]) {
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (kAtomic) atomic_add(p + a, v[a]); else plain_add(p + a, v[a]);
  }
}

template <typename T, bool kAtomic>
__global__ void __launch_bounds__(256)
residual_kernel(int32_t lo, int32_t hi, const int32_t *__restrict__ list, T *__restrict__ R,
                const T *__restrict__ X, const T *__restrict__ u,
                const int32_t *__restrict__ c2n, const int32_t *__restrict__ c2v, T lam, T mu) {
  const int32_t t = lo + blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= hi) return;
  const int32_t e = element_of(t, list);
  int32_t nd[10];
#pragma unroll
  for (int j = 0; j < 10; ++j) nd[j] = c2n[10 * e + j];
  const Geom<T> g = geometry(X, c2v, e);
  T uu[10][3];
#pragma unroll
  for (int j = 0; j < 10; ++j)
#pragma unroll
    for (int b = 0; b < 3; ++b) uu[j][b] = u[3 * nd[j] + b];
  T re[10][3];
  element_residual(uu, g, lam, mu, re);
#pragma unroll
  for (int i = 0; i < 10; ++i) add3<T, kAtomic>(R + 3 * nd[i], re[i]);
}

// Gradient of basis i at every quadrature point (runtime i, table lookup).
template <typename T>
__device__ __forceinline__ void basis_grads(const Geom<T> &g, int i, T (&gi)[4][3]) {
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const T d0 = T(kD.v[q][i][0]), d1 = T(kD.v[q][i][1]), d2 = T(kD.v[q][i][2]),
            d3 = T(kD.v[q][i][3]);
#pragma unroll
    for (int k = 0; k < 3; ++k)
      gi[q][k] = d0 * g.gl[0][k] + d1 * g.gl[1][k] + d2 * g.gl[2][k] + d3 * g.gl[3][k];
  }
}

// Block row i of the element matrix in closed form.  With gi_q the gradient
// of basis i at point q and L_qn the barycentric coordinates of the points,
// let G_n = sum_q L_qn gi_q and S = sum_q gi_q.  The rule's moments
// sum_q L_qn = 1, sum_q L_qn^2 = 0.4, sum_q L_qn L_qm = 0.2 (n != m) give
//   vertex row i:       4 G_n = (3.2 [n == i] - 0.8) gl_i,   S = 0,
//   edge row i = (a,b): 4 G_n = 3.2 (gl_a + gl_b) + 3.2 ([n == b] gl_a + [n == a] gl_b),
//                       S = 4 (gl_a + gl_b),
// and M_ij = sum_q gi_q (x) gj_q = sum_n (sum_q D[q][j][n] gi_q) (x) gl_n is
//   vertex column j:    M_ij = (4 G_j - S) (x) gl_j,
//   edge column (c,d):  M_ij = 4 G_d (x) gl_c + 4 G_c (x) gl_d.
template <typename T>
struct RowCoef {
  T G4[4][3];  // 4 G_n
  T S[3];
  T lw, mw;    // wdet * lam, wdet * mu
};

template <typename T>
__device__ __forceinline__ T gl_of(const Geom<T> &g, int v, int k) {
  // select among values (a runtime index, or a select among the array's
  // elements as lvalues, would place the geometry in scratch memory)
  const T g0 = g.gl[0][k], g1 = g.gl[1][k], g2 = g.gl[2][k], g3 = g.gl[3][k];
  const T lo = v == 0 ? g0 : g1;
  const T hi = v == 2 ? g2 : g3;
  return v < 2 ? lo : hi;
}

template <typename T>
__device__ __forceinline__ RowCoef<T> row_coefficients(const Geom<T> &g, int i, T lam, T mu) {
  const bool edge = i >= 4;
  const int ei = i - 4;
  const int a = edge ? (ei < 3 ? 0 : (ei < 5 ? 1 : 2)) : i;
  const int b = edge ? (ei < 3 ? ei + 1 : (ei < 5 ? ei - 1 : 3)) : i;
  const T c0 = edge ? T(3.2) : T(-0.8);
  const T c1 = edge ? T(3.2) : T(1.6);
  const T cs = edge ? T(4) : T(0);
  RowCoef<T> r;
  T ga[3], gb[3], w[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    ga[k] = gl_of(g, a, k);
    gb[k] = gl_of(g, b, k);
    w[k] = edge ? ga[k] + gb[k] : ga[k];
    r.S[k] = cs * w[k];
  }
#pragma unroll
  for (int n = 0; n < 4; ++n)
#pragma unroll
    for (int k = 0; k < 3; ++k)
      r.G4[n][k] = c0 * w[k] + c1 * ((n == b ? ga[k] : T(0)) + (n == a ? gb[k] : T(0)));
  r.lw = g.wdet * lam;
  r.mw = g.wdet * mu;
  return r;
}

// Block (i, j) = wdet (lam M + mu M^T + mu tr(M) I) of row i's coefficients.
template <typename T, int J>
__device__ __forceinline__ void closed_block(const RowCoef<T> &r, const Geom<T> &g, T (&B)[9]) {
  T M[3][3];
  if constexpr (J < 4) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const T h = r.G4[J][a] - r.S[a];
#pragma unroll
      for (int b = 0; b < 3; ++b) M[a][b] = h * g.gl[J][b];
    }
  } else {
    constexpr int c = kEa[J - 4], d = kEb[J - 4];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b) M[a][b] = r.G4[d][a] * g.gl[c][b] + r.G4[c][a] * g.gl[d][b];
  }
  const T tr = r.mw * (M[0][0] + M[1][1] + M[2][2]);
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b)
      B[3 * a + b] = r.lw * M[a][b] + r.mw * M[b][a] + (a == b ? tr : T(0));
}

template <typename T, int J, class F>
__device__ __forceinline__ void for_blocks(const RowCoef<T> &r, const Geom<T> &g, F &&f) {
  T B[9];
  closed_block<T, J>(r, g, B);
  f(J, B);
  if constexpr (J + 1 < 10) for_blocks<T, J + 1>(r, g, f);
}

// The blocks j with bit j of kMask set.
template <typename T, unsigned kMask, int J = 0, class F>
__device__ __forceinline__ void for_blocks_in(const RowCoef<T> &r, const Geom<T> &g, F &&f) {
  if constexpr (J < 10) {
    if constexpr (((kMask >> J) & 1u) != 0) {
      T B[9];
      closed_block<T, J>(r, g, B);
      f(J, B);
    }
    for_blocks_in<T, kMask, J + 1>(r, g, f);
  }
}

// Column groups of the staged assembly, balanced by cost (an edge column
// block costs about 4/3 of a vertex one): {0,1,4}, {2,3,5}, {6,7}, {8,9}.
constexpr unsigned kColumnGroups[4] = {0x013u, 0x02Cu, 0x0C0u, 0x300u};

// Cell-centric BSR(3) assembly: block (i, j) -> slot by binary search of row
// c2n[i]'s colidx, 9 atomic (or coloured plain) adds.
template <typename T, bool kAtomic, bool kFused>
__global__ void __launch_bounds__(256)
matrix_kernel(int32_t lo, int32_t hi, const int32_t *__restrict__ list, op2hip_mat A,
              T *__restrict__ R, const T *__restrict__ X, const T *__restrict__ u,
              const int32_t *__restrict__ c2n, const int32_t *__restrict__ c2v, T lam, T mu) {
  const int32_t t = lo + blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= hi) return;
  const int32_t e = element_of(t, list);
  int32_t nd[10];
#pragma unroll
  for (int j = 0; j < 10; ++j) nd[j] = c2n[10 * e + j];
  const Geom<T> g = geometry(X, c2v, e);
  T *vals = static_cast<T *>(A.values);
  for (int i = 0; i < 10; ++i) {
    const int32_t r = nd[i];
    if (r < 0) continue;
    const RowCoef<T> rc = row_coefficients(g, i, lam, mu);
    const int32_t rlo = A.rowptr[r], rhi = A.rowptr[r + 1];
    T rs[3] = {T(0), T(0), T(0)};
    for_blocks<T, 0>(rc, g, [&](int j, const T (&B)[9]) {
      const int32_t c = nd[j];
      if (c < 0) return;
      if (kFused) {
        const T u0 = u[3 * c], u1 = u[3 * c + 1], u2 = u[3 * c + 2];
#pragma unroll
        for (int a = 0; a < 3; ++a) rs[a] += B[3 * a] * u0 + B[3 * a + 1] * u1 + B[3 * a + 2] * u2;
      }
      const int32_t slot = find_slot(A.colidx, rlo, rhi, c);
      if (slot < 0) {
        atomicAdd(A.status, 1);
        return;
      }
#pragma unroll
      for (int k = 0; k < 9; ++k) {
        if (kAtomic) atomic_add(&vals[9 * (size_t)slot + k], B[k]);
        else plain_add(&vals[9 * (size_t)slot + k], B[k]);
      }
    });
    if (kFused) add3<T, kAtomic>(R + 3 * r, rs);
  }
}

// Row-gather element: block row i of the element matrix (and, fused, the
// row's residual contribution sum_j B_ij u_j).
template <typename TT, bool kFused>
struct MatrixRow {
  using T = TT;
  static constexpr int kRar = 10, kCar = 10, kRd = 3, kCd = 3;
  static constexpr bool kHasRowDat = kFused;
  // prefetch depth that keeps the float64 kernels at or below 168 registers
  // (three waves per SIMD, the occupancy their LDS allows)
  static constexpr int kMaxPrefetch = sizeof(TT) == 8 ? (kFused ? 8 : 16) : 24;
  static constexpr bool kSpread = true;
  // split-gather rounds of one incidence per thread for float64: a batch's
  // block budget (at most 640 blocks, about 160 incidences) leaves threads
  // to spare, and the second slot's registers would cost a wave per SIMD
  // (121 registers: four waves per SIMD)
  static constexpr int kRound = sizeof(TT) == 8 ? 1 : 2;
  // chunks of 8 consecutive row batches go to the XCDs in turn: the chip
  // streams one compact window of the matrix values at a time, and each XCD's
  // L2 keeps the elements its 8 neighbouring batches share (float64 at 64^3:
  // 1.66 ms, against 1.72-1.73 ms for single batches dealt in turn and 1.80
  // ms for a contiguous range of batches per XCD)
  static constexpr int32_t kXcdChunk = 8;
  const T *X;
  const T *u;
  const int32_t *c2n;
  const int32_t *c2v;
  T lam, mu;
  // cols: emit the column node ids (the pattern-search row gathers need them;
  // the slot-table ones do not: the hand-written loops take maps without
  // negative entries and the slot table marks a column missing from the
  // pattern), and the fused variant always reads them for u.
  bool cols = true;
  template <class Emit>
  __device__ __forceinline__ void row(int32_t e, int i, Emit &&emit, T *rsum) const {
    int32_t nd[10];
#pragma unroll
    for (int j = 0; j < 10; ++j) nd[j] = (kFused || cols) ? c2n[10 * e + j] : 0;
    const Geom<T> g = geometry(X, c2v, e);
    const RowCoef<T> rc = row_coefficients(g, i, lam, mu);
    for_blocks<T, 0>(rc, g, [&](int j, const T (&B)[9]) {
      const int32_t c = nd[j];
      if (kFused && c >= 0) {
        const T u0 = u[3 * c], u1 = u[3 * c + 1], u2 = u[3 * c + 2];
#pragma unroll
        for (int a = 0; a < 3; ++a)
          rsum[a] += B[3 * a] * u0 + B[3 * a + 1] * u1 + B[3 * a + 2] * u2;
      }
      emit(j, c, B);
    });
  }
  // Two lanes per row incidence in the slot-table gather (float64; float32
  // runs faster with one): half H emits blocks {0,1,4,5,6} or {2,3,7,8,9}
  // (two vertex and three edge columns each).
  static constexpr int kHalves = sizeof(TT) == 8 ? 2 : 1;
  template <int H, class Emit>
  __device__ __forceinline__ void row_half(int32_t e, int i, Emit &&emit, T *rsum) const {
    constexpr unsigned kMask = H == 0 ? 0x073u : 0x38Cu;
    int32_t nd[10];
#pragma unroll
    for (int j = 0; j < 10; ++j)
      if ((kMask >> j) & 1u) nd[j] = (kFused || cols) ? c2n[10 * e + j] : 0;
    const Geom<T> g = geometry(X, c2v, e);
    const RowCoef<T> rc = row_coefficients(g, i, lam, mu);
    for_blocks_in<T, kMask>(rc, g, [&](int j, const T (&B)[9]) {
      const int32_t c = nd[j];
      if (kFused && c >= 0) {
        const T u0 = u[3 * c], u1 = u[3 * c + 1], u2 = u[3 * c + 2];
#pragma unroll
        for (int a = 0; a < 3; ++a)
          rsum[a] += B[3 * a] * u0 + B[3 * a + 1] * u1 + B[3 * a + 2] * u2;
      }
      emit(j, c, B);
    });
  }
  // split gather (slot-table launches of the assembly alone): vertex ids,
  // then every column block of row i from the gathered coordinates
  static constexpr bool kSplitGather = !kFused;
  static constexpr int kNv = 4, kGdim = 3;
  __device__ __forceinline__ void vertices(int32_t e, int32_t (&v)[4]) const {
#pragma unroll
    for (int m = 0; m < 4; ++m) v[m] = c2v[4 * e + m];
  }
  template <class Emit>
  __device__ __forceinline__ void row_at(const T (&xv)[4][3], int i, Emit &&emit) const {
    const Geom<T> g = geometry_of(xv);
    const RowCoef<T> rc = row_coefficients(g, i, lam, mu);
    for_blocks<T, 0>(rc, g, [&](int j, const T (&B)[9]) { emit(j, 0, B); });
  }
  static constexpr int kGroups = 4;
  template <int G, class Emit>
  __device__ __forceinline__ void row_group(int32_t e, int i, Emit &&emit, T *rsum) const {
    int32_t nd[10];
#pragma unroll
    for (int j = 0; j < 10; ++j)
      if ((kColumnGroups[G] >> j) & 1u) nd[j] = (kFused || cols) ? c2n[10 * e + j] : 0;
    const Geom<T> g = geometry(X, c2v, e);
    const RowCoef<T> rc = row_coefficients(g, i, lam, mu);
    for_blocks_in<T, kColumnGroups[G]>(rc, g, [&](int j, const T (&B)[9]) {
      const int32_t c = nd[j];
      if (kFused && c >= 0) {
        const T u0 = u[3 * c], u1 = u[3 * c + 1], u2 = u[3 * c + 2];
#pragma unroll
        for (int a = 0; a < 3; ++a)
          rsum[a] += B[3 * a] * u0 + B[3 * a + 1] * u1 + B[3 * a + 2] * u2;
      }
      emit(j, c, B);
    });
  }
};

// Cell-gather element: per element the barycentric gradients gl[m] (padded
// to 4), wdet * lam and wdet * mu; block (i, j) evaluated from them with the
// closed form of row_coefficients / closed_block for runtime i and j.
template <typename TT, bool kFused>
struct MatrixCells {
  using T = TT;
  static constexpr int kRar = 10, kCar = 10, kRd = 3, kCd = 3;
  static constexpr bool kHasRowDat = kFused;
  static constexpr int kGeo = 18;
  const T *X;
  const T *u;
  const int32_t *c2v;
  T lam, mu;
  __device__ __forceinline__ void geometry(int32_t e, double *geo) const {
    const Geom<T> g = p2::geometry(X, c2v, e);
#pragma unroll
    for (int m = 0; m < 4; ++m) {
#pragma unroll
      for (int k = 0; k < 3; ++k) geo[4 * m + k] = double(g.gl[m][k]);
      geo[4 * m + 3] = 0.0;
    }
    geo[16] = double(g.wdet) * double(lam);
    geo[17] = double(g.wdet) * double(mu);
  }
  __device__ static void no_geometry(double *geo) {
#pragma unroll
    for (int k = 0; k < kGeo; ++k) geo[k] = 0.0;
  }
  // vertices (a, b) of P2 basis i (a = b = i for a vertex)
  __device__ static __forceinline__ void ends(int i, int &a, int &b) {
    const int ei = i - 4;
    a = i < 4 ? i : (ei < 3 ? 0 : (ei < 5 ? 1 : 2));
    b = i < 4 ? i : (ei < 3 ? ei + 1 : (ei < 5 ? ei - 1 : 3));
  }
  __device__ __forceinline__ void block(const double *geo, int i, int j, T *B) const {
    int a, b, c, d;
    ends(i, a, b);
    ends(j, c, d);
    const bool ie = i >= 4, jv = j < 4;
    const double c0 = ie ? 3.2 : -0.8, c1 = ie ? 3.2 : 1.6, cs = ie ? 4.0 : 0.0;
    double P[3], Q[3], gc[3], gd[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const double ga = geo[4 * a + k], gb = geo[4 * b + k];
      const double w = ie ? ga + gb : ga;
      gc[k] = geo[4 * c + k];
      gd[k] = geo[4 * d + k];
      const double g4c = c0 * w + c1 * ((c == b ? ga : 0.0) + (c == a ? gb : 0.0));
      const double g4d = c0 * w + c1 * ((d == b ? ga : 0.0) + (d == a ? gb : 0.0));
      P[k] = g4d - (jv ? cs * w : 0.0);
      Q[k] = jv ? 0.0 : g4c;
    }
    double M[3][3];
#pragma unroll
    for (int x = 0; x < 3; ++x)
#pragma unroll
      for (int y = 0; y < 3; ++y) M[x][y] = P[x] * gc[y] + Q[x] * gd[y];
    const double lw = geo[16], mw = geo[17];
    const double tr = mw * (M[0][0] + M[1][1] + M[2][2]);
#pragma unroll
    for (int x = 0; x < 3; ++x)
#pragma unroll
      for (int y = 0; y < 3; ++y) B[3 * x + y] = T(lw * M[x][y] + mw * M[y][x] + (x == y ? tr : 0.0));
  }
  __device__ __forceinline__ void rowdat(int32_t col, const T *B, double *rs) const {
    const double u0 = u[3 * (size_t)col], u1 = u[3 * (size_t)col + 1], u2 = u[3 * (size_t)col + 2];
#pragma unroll
    for (int x = 0; x < 3; ++x) rs[x] += double(B[3 * x]) * u0 + double(B[3 * x + 1]) * u1 + double(B[3 * x + 2]) * u2;
  }
};

// Lane-per-block row-gather element: lane j of a group computes block (i, j).
template <typename TT, bool kFused>
struct MatrixLanes {
  using T = TT;
  using Geom = p2::Geom<TT>;
  static constexpr int kRar = 10, kCar = 10, kRd = 3, kCd = 3;
  static constexpr bool kHasRowDat = kFused;
  const T *X;
  const T *u;
  const int32_t *c2n;
  const int32_t *c2v;
  T lam, mu;
  __device__ __forceinline__ Geom geometry(int32_t e) const { return p2::geometry(X, c2v, e); }
  __device__ __forceinline__ int32_t col(int32_t e, int j) const { return c2n[10 * e + j]; }
  __device__ __forceinline__ void block(const Geom &g, int i, int j, T *B) const {
    T gi[4][3], gj[4][3];
    basis_grads(g, i, gi);
    basis_grads(g, j, gj);
    T M[3][3] = {};
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) M[a][b] += gi[q][a] * gj[q][b];
    const T tr = mu * (M[0][0] + M[1][1] + M[2][2]);
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b)
        B[3 * a + b] = g.wdet * (lam * M[a][b] + mu * M[b][a] + (a == b ? tr : T(0)));
  }
  __device__ __forceinline__ void rowdat(int32_t c, const T *B, T *rs) const {
    const T u0 = u[3 * c], u1 = u[3 * c + 1], u2 = u[3 * c + 2];
#pragma unroll
    for (int a = 0; a < 3; ++a) rs[a] += B[3 * a] * u0 + B[3 * a + 1] * u1 + B[3 * a + 2] * u2;
  }
};

// Staged residual: map 0 = cell -> P2 node (u staged, R incremented), map 1 =
// cell -> vertex (coordinates staged).
template <typename TT>
struct ResidualStaged {
  using T = TT;
  static constexpr int kMaps = 2;
  static constexpr int kAr0 = 10, kIn0 = 3, kOut0 = 3;
  static constexpr int kAr1 = 4, kIn1 = 3, kOut1 = 0;
  T *R;
  const T *X;
  const T *u;
  T lam, mu;
  template <int M>
  __device__ __forceinline__ void stage(int32_t v, T *dst) const {
    const T *src = (M == 0 ? u : X) + 3 * (size_t)v;
    dst[0] = src[0];
    dst[1] = src[1];
    dst[2] = src[2];
  }
  template <int M>
  __host__ __device__ __forceinline__ T *out() const { return R; }
  __device__ __forceinline__ void element(int32_t, const T (&uu)[10][3], const T (&xv)[4][3],
                                          T (&re)[10][3], T (&)[4][1]) const {
    element_residual(uu, geometry_of(xv), lam, mu, re);
  }
};

// Staged adjoint: map 0 = cell -> P2 node (u and Rbar staged, ubar
// incremented), map 1 = cell -> vertex (X staged, Xbar incremented).
template <typename TT>
struct AdjointStaged {
  using T = TT;
  // reads u and Rbar from the tile's LDS records at each quadrature point:
  // 210 registers (two waves per SIMD) against 256 + 40 accumulation
  // registers (one wave) for the same element taking its 60 inputs as
  // arrays; 230 against 257-262 us at 64^3
  static constexpr bool kLdsElement = true;
  static constexpr int kMaps = 2;
  static constexpr int kAr0 = 10, kIn0 = 6, kOut0 = 3;
  static constexpr int kAr1 = 4, kIn1 = 3, kOut1 = 3;
  T *ubar;
  T *Xbar;
  const T *Rbar;
  const T *X;
  const T *u;
  T lam, mu;
  template <int M>
  __device__ __forceinline__ void stage(int32_t v, T *dst) const {
    if (M == 0) {
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        dst[a] = u[3 * (size_t)v + a];
        dst[3 + a] = Rbar[3 * (size_t)v + a];
      }
    } else {
#pragma unroll
      for (int a = 0; a < 3; ++a) dst[a] = X[3 * (size_t)v + a];
    }
  }
  template <int M>
  __host__ __device__ __forceinline__ T *out() const { return M == 0 ? ubar : Xbar; }
  __device__ __forceinline__ void element_lds(int32_t, const T *in0, const uint16_t (&l0)[10],
                                              const T *in1, const uint16_t (&l1)[4],
                                              T (&ub)[10][3], T (&xb)[4][3]) const;
};

// Adjoint of the residual loop for one element (see header of the entry point).
template <typename T, int Q>
__device__ __forceinline__ void adjoint_q(const T (&uu)[10][3], const T (&ee)[10][3],
                                          const Geom<T> &g, T lam, T mu, T (&ub)[10][3],
                                          T (&P)[3][3]) {
  T gu[3][3], ge[3][3], su[3][3], se[3][3];
  field_grad<T, Q>(uu, g, gu);
  field_grad<T, Q>(ee, g, ge);
  stress(gu, lam, mu, su);
  stress(ge, lam, mu, se);
  test_gradients<T, Q>(se, g, ub);
  T phi = T(0);
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int k = 0; k < 3; ++k) phi += su[a][k] * ge[a][k];
#pragma unroll
  for (int l = 0; l < 3; ++l)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      T acc = (l == k) ? phi : T(0);
#pragma unroll
      for (int b = 0; b < 3; ++b) acc -= gu[b][l] * se[b][k] + ge[b][l] * su[b][k];
      P[l][k] += acc;
    }
}

// Adjoint of one element's residual: ub = K_e ee (the residual's transpose
// applied to ee) and xb = d(ee . R_e)/dX (rows: vertices).
template <typename T>
__device__ __forceinline__ void element_adjoint(const T (&uu)[10][3], const T (&ee)[10][3],
                                                const Geom<T> &g, T lam, T mu, T (&ub)[10][3],
                                                T (&xb)[4][3]) {
#pragma unroll
  for (int i = 0; i < 10; ++i)
#pragma unroll
    for (int a = 0; a < 3; ++a) ub[i][a] = T(0);
  T P[3][3] = {};
  adjoint_q<T, 0>(uu, ee, g, lam, mu, ub, P);
  adjoint_q<T, 1>(uu, ee, g, lam, mu, ub, P);
  adjoint_q<T, 2>(uu, ee, g, lam, mu, ub, P);
  adjoint_q<T, 3>(uu, ee, g, lam, mu, ub, P);
#pragma unroll
  for (int i = 0; i < 10; ++i) {
#pragma unroll
    for (int a = 0; a < 3; ++a) ub[i][a] *= g.wdet;
  }
  // dS/dJ = wdet * P K^T ; X[m+1] gets column m, X[0] minus the row sums.
#pragma unroll
  for (int l = 0; l < 3; ++l) {
    T rowsum = T(0);
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      const T v = g.wdet * (P[l][0] * g.K[m][0] + P[l][1] * g.K[m][1] + P[l][2] * g.K[m][2]);
      xb[m + 1][l] = v;
      rowsum += v;
    }
    xb[0][l] = -rowsum;
  }
}


// Barycentric-form gradient at point Q of a field held in LDS: component c
// of node j at in[l[j] * stride + off + c].
template <typename T, int Q>
__device__ __forceinline__ void field_grad_lds(const T *in, const uint16_t (&l)[10], int stride,
                                               int off, const Geom<T> &g, T (&grad)[3][3]) {
  T w[4][3];
#pragma unroll
  for (int m = 0; m < 4; ++m)
#pragma unroll
    for (int b = 0; b < 3; ++b) w[m][b] = T(0);
#pragma unroll
  for (int j = 0; j < 10; ++j) {
    const T *f = in + (size_t)l[j] * stride + off;
    const T f0 = f[0], f1 = f[1], f2 = f[2];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      if (dcoef(Q, j, m) == 0.0) continue;
      const T c = T(dcoef(Q, j, m));
      w[m][0] += c * f0;
      w[m][1] += c * f1;
      w[m][2] += c * f2;
    }
  }
#pragma unroll
  for (int b = 0; b < 3; ++b)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      T acc = T(0);
#pragma unroll
      for (int m = 0; m < 4; ++m) acc += w[m][b] * g.gl[m][k];
      grad[b][k] = acc;
    }
}

// The shape term of the adjoint at point Q, P += phi I - gu^T se - ge^T su
// (u and Rbar read from the staged records, six values a node).  The
// scheduling barrier keeps the next point's reads after this point's work,
// so the element never holds all its inputs at once.
template <typename T, int Q>
__device__ __forceinline__ void shape_q_lds(const T *in0, const uint16_t (&l0)[10],
                                            const Geom<T> &g, T lam, T mu, T (&P)[3][3]) {
  T gu[3][3], ge[3][3], su[3][3], se[3][3];
  field_grad_lds<T, Q>(in0, l0, 6, 0, g, gu);
  field_grad_lds<T, Q>(in0, l0, 6, 3, g, ge);
  stress(gu, lam, mu, su);
  stress(ge, lam, mu, se);
  T phi = T(0);
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int k = 0; k < 3; ++k) phi += su[a][k] * ge[a][k];
#pragma unroll
  for (int l = 0; l < 3; ++l)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      T acc = (l == k) ? phi : T(0);
#pragma unroll
      for (int b = 0; b < 3; ++b) acc -= gu[b][l] * se[b][k] + ge[b][l] * su[b][k];
      P[l][k] += acc;
    }
  __builtin_amdgcn_sched_barrier(0);
}

// Staged adjoint element: xb from the shape term at the four points, ub =
// K_e Rbar as the residual of Rbar (the residual is linear in u and K_e
// symmetric), both reading the staged records on demand.
template <typename TT>
__device__ __forceinline__ void AdjointStaged<TT>::element_lds(int32_t, const TT *in0,
                                                               const uint16_t (&l0)[10],
                                                               const TT *in1,
                                                               const uint16_t (&l1)[4],
                                                               TT (&ub)[10][3],
                                                               TT (&xb)[4][3]) const {
  TT xv[4][3];
#pragma unroll
  for (int v = 0; v < 4; ++v)
#pragma unroll
    for (int a = 0; a < 3; ++a) xv[v][a] = in1[(size_t)l1[v] * 3 + a];
  const Geom<TT> g = geometry_of(xv);
  TT P[3][3] = {};
  shape_q_lds<TT, 0>(in0, l0, g, lam, mu, P);
  shape_q_lds<TT, 1>(in0, l0, g, lam, mu, P);
  shape_q_lds<TT, 2>(in0, l0, g, lam, mu, P);
  shape_q_lds<TT, 3>(in0, l0, g, lam, mu, P);
  // dS/dJ = wdet * P K^T ; X[m+1] gets column m, X[0] minus the row sums
#pragma unroll
  for (int l = 0; l < 3; ++l) {
    TT rowsum = TT(0);
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      const TT v = g.wdet * (P[l][0] * g.K[m][0] + P[l][1] * g.K[m][1] + P[l][2] * g.K[m][2]);
      xb[m + 1][l] = v;
      rowsum += v;
    }
    xb[0][l] = -rowsum;
  }
  residual_vertex_form(
      g, lam, mu, [&](int j, int c) { return in0[(size_t)l0[j] * 6 + 3 + c]; },
      [&](int i, const TT (&r)[3]) {
#pragma unroll
        for (int c = 0; c < 3; ++c) ub[i][c] = r[c];
      });
}

template <typename T, bool kAtomic>
__global__ void __launch_bounds__(256)
adjoint_kernel(int32_t lo, int32_t hi, const int32_t *__restrict__ list, T *__restrict__ ubar,
               T *__restrict__ Xbar, const T *__restrict__ Rbar, const T *__restrict__ X,
               const T *__restrict__ u, const int32_t *__restrict__ c2n,
               const int32_t *__restrict__ c2v, T lam, T mu) {
  const int32_t t = lo + blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= hi) return;
  const int32_t e = element_of(t, list);
  int32_t nd[10];
#pragma unroll
  for (int j = 0; j < 10; ++j) nd[j] = c2n[10 * e + j];
  const Geom<T> g = geometry(X, c2v, e);
  T uu[10][3], ee[10][3];
#pragma unroll
  for (int j = 0; j < 10; ++j)
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      uu[j][b] = u[3 * nd[j] + b];
      ee[j][b] = Rbar[3 * nd[j] + b];
    }
  T ub[10][3], xb[4][3];
  element_adjoint(uu, ee, g, lam, mu, ub, xb);
#pragma unroll
  for (int i = 0; i < 10; ++i) add3<T, kAtomic>(ubar + 3 * nd[i], ub[i]);
#pragma unroll
  for (int v = 0; v < 4; ++v) add3<T, kAtomic>(Xbar + 3 * c2v[4 * e + v], xb[v]);
}

template <typename T>
int residual(int32_t start, int32_t end, T *R, const T *X, const T *u, const int32_t *c2n,
             const int32_t *c2v, double lam, double mu, const op2hip_exec *ex, hipStream_t s) {
  if (ex && ex->strategy == OP2HIP_STAGED)
    return launch_staged(start, end, ResidualStaged<T>{R, X, u, T(lam), T(mu)}, ex, s);
  const int bs = block_size_of(ex, 256);
  return for_each_launch(ex, start, end, [&](int32_t lo, int32_t hi, const int32_t *list,
                                             bool atomic) -> int {
    if (atomic)
      residual_kernel<T, true><<<ceil_div(hi - lo, bs), bs, 0, s>>>(lo, hi, list, R, X, u, c2n,
                                                                    c2v, T(lam), T(mu));
    else
      residual_kernel<T, false><<<ceil_div(hi - lo, bs), bs, 0, s>>>(lo, hi, list, R, X, u, c2n,
                                                                     c2v, T(lam), T(mu));
    OP2HIP_LAUNCHED();
    return OP2HIP_OK;
  });
}

template <typename T, bool kFused>
int matrix(int32_t start, int32_t end, const op2hip_mat *A, T *R, const T *X, const T *u,
           const int32_t *c2n, const int32_t *c2v, double lam, double mu, const op2hip_exec *ex,
           hipStream_t s) {
  OP2HIP_REQUIRE(A && A->rowptr && A->colidx && A->values && A->status, "p2 matrix: bad Mat");
  OP2HIP_REQUIRE(A->rdim == 3 && A->cdim == 3, "p2 matrix: Mat block must be 3x3 (BSR)");
  if (ex && ex->strategy == OP2HIP_ROWGATHER) {
    MatrixLanes<T, kFused> elem{X, u, c2n, c2v, T(lam), T(mu)};
    return launch_rowgather_lanes(start, end, elem, A, ex, R, s);
  }
  if (ex && ex->strategy == OP2HIP_ROWGATHER_SLOTS) {
    MatrixRow<T, kFused> elem{X, u, c2n, c2v, T(lam), T(mu), false};
    return launch_rowgather_slots(start, end, elem, A, ex, R, s);
  }
  if (ex && ex->strategy == OP2HIP_ROWGATHER_STAGED) {
    MatrixRow<T, kFused> elem{X, u, c2n, c2v, T(lam), T(mu), false};
    return launch_rowgather_staged(start, end, elem, A, ex, R, s);
  }
  if (ex && ex->strategy == OP2HIP_ROWGATHER_ROWLIST) {
    MatrixRow<T, kFused> elem{X, u, c2n, c2v, T(lam), T(mu), false};
    return launch_rowgather_rowlist(start, end, elem, A, ex, R, s);
  }
  if (ex && ex->strategy == OP2HIP_ROWGATHER_CELLS) {
    MatrixCells<T, kFused> elem{X, u, c2v, T(lam), T(mu)};
    return launch_rowgather_cells(start, end, elem, A, ex, R, s);
  }
  if (ex && ex->strategy == OP2HIP_ROWGATHER_THREAD) {
    MatrixRow<T, kFused> elem{X, u, c2n, c2v, T(lam), T(mu)};
    return launch_rowgather(start, end, elem, A, ex, R, s);
  }
  const int bs = block_size_of(ex, 256);
  const op2hip_mat Av = *A;
  return for_each_launch(ex, start, end, [&](int32_t lo, int32_t hi, const int32_t *list,
                                             bool atomic) -> int {
    if (atomic)
      matrix_kernel<T, true, kFused><<<ceil_div(hi - lo, bs), bs, 0, s>>>(
          lo, hi, list, Av, R, X, u, c2n, c2v, T(lam), T(mu));
    else
      matrix_kernel<T, false, kFused><<<ceil_div(hi - lo, bs), bs, 0, s>>>(
          lo, hi, list, Av, R, X, u, c2n, c2v, T(lam), T(mu));
    OP2HIP_LAUNCHED();
    return OP2HIP_OK;
  });
}

}  // namespace p2
}  // namespace op2hip

// ---- the residual without allocation ---------------------------------------
// Partial sums of the nodes that staged tiles share, for
// op2hip_wrap_p2_elasticity_residual_noalloc: a file-scope device array of
// OP2HIP_NOALLOC_SCRATCH_BYTES, which every GPU the library runs on holds
// until the process ends.  Host code never passes the array's name to a
// kernel (in host code it is a stand-in, not a device address): it uses the
// device address looked up once per device.
static __device__ __align__(256) double scratch[OP2HIP_NOALLOC_SCRATCH_BYTES / sizeof(double)];

namespace op2hip {
namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void *g_scratch_ptr[kMaxDevices];
}  // namespace

// Device address of `scratch` on `device` (the current device), looked up
// with hipGetSymbolAddress on the first call per device and kept.
int noalloc_scratch_address(int device, void **out) {
  OP2HIP_REQUIRE(device >= 0 && device < kMaxDevices, "device %d out of range", device);
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error("hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
      return OP2HIP_ERR_HIP;
    }
    g_scratch_ptr[device] = p;
  }
  *out = g_scratch_ptr[device];
  return OP2HIP_OK;
}
}  // namespace op2hip

using namespace op2hip;

extern "C" {

int op2hip_wrap_p2_elasticity_residual_noalloc(int32_t start, int32_t end, double *R,
                                               const double *X, const double *u,
                                               const int32_t *c2n, const int32_t *c2v,
                                               double lam, double mu, const op2hip_exec *ex,
                                               int64_t partial_slots, void *stream) {
  OP2HIP_REQUIRE(ex, "residual_noalloc: no execution plan");
  op2hip_exec local = *ex;
  if (local.strategy == OP2HIP_STAGED && local.tile_shared_off[0]) {
    const int64_t cap = (int64_t)(OP2HIP_NOALLOC_SCRATCH_BYTES / (3 * sizeof(double)));
    OP2HIP_REQUIRE(partial_slots >= 0 && partial_slots <= cap,
                   "residual_noalloc: %lld partial slots of shared tile nodes exceed the "
                   "library's scratch of %lld slots (%u bytes); use "
                   "op2hip_wrap_p2_elasticity_residual with a workspace",
                   (long long)partial_slots, (long long)cap, (unsigned)OP2HIP_NOALLOC_SCRATCH_BYTES);
    int device = 0;
    OP2HIP_CHECK(hipGetDevice(&device));
    void *p = nullptr;
    const int rc = noalloc_scratch_address(device, &p);
    if (rc != OP2HIP_OK) return rc;
    local.partials[0] = static_cast<double *>(p);
    local.partials[1] = nullptr;
  }
  return p2::residual<double>(start, end, R, X, u, c2n, c2v, lam, mu, &local, stream_of(stream));
}

int op2hip_wrap_p2_elasticity_residual(int32_t start, int32_t end, double *R, const double *X,
                                       const double *u, const int32_t *c2n, const int32_t *c2v,
                                       double lam, double mu, const op2hip_exec *ex,
                                       void *stream) {
  return p2::residual<double>(start, end, R, X, u, c2n, c2v, lam, mu, ex, stream_of(stream));
}

int op2hip_wrap_p2_elasticity_residual_f32(int32_t start, int32_t end, float *R, const float *X,
                                           const float *u, const int32_t *c2n,
                                           const int32_t *c2v, double lam, double mu,
                                           const op2hip_exec *ex, void *stream) {
  return p2::residual<float>(start, end, R, X, u, c2n, c2v, lam, mu, ex, stream_of(stream));
}

int op2hip_wrap_p2_elasticity_matrix(int32_t start, int32_t end, const op2hip_mat *A,
                                     const double *X, const int32_t *c2n, const int32_t *c2v,
                                     double lam, double mu, const op2hip_exec *ex, void *stream) {
  return p2::matrix<double, false>(start, end, A, nullptr, X, nullptr, c2n, c2v, lam, mu, ex,
                                   stream_of(stream));
}

int op2hip_wrap_p2_elasticity_matrix_f32(int32_t start, int32_t end, const op2hip_mat *A,
                                         const float *X, const int32_t *c2n, const int32_t *c2v,
                                         double lam, double mu, const op2hip_exec *ex,
                                         void *stream) {
  return p2::matrix<float, false>(start, end, A, nullptr, X, nullptr, c2n, c2v, lam, mu, ex,
                                  stream_of(stream));
}

int op2hip_wrap_p2_elasticity_fused(int32_t start, int32_t end, const op2hip_mat *A, double *R,
                                    const double *X, const double *u, const int32_t *c2n,
                                    const int32_t *c2v, double lam, double mu,
                                    const op2hip_exec *ex, void *stream) {
  return p2::matrix<double, true>(start, end, A, R, X, u, c2n, c2v, lam, mu, ex,
                                  stream_of(stream));
}

int op2hip_wrap_p2_elasticity_fused_f32(int32_t start, int32_t end, const op2hip_mat *A,
                                        float *R, const float *X, const float *u,
                                        const int32_t *c2n, const int32_t *c2v, double lam,
                                        double mu, const op2hip_exec *ex, void *stream) {
  return p2::matrix<float, true>(start, end, A, R, X, u, c2n, c2v, lam, mu, ex,
                                 stream_of(stream));
}

int op2hip_wrap_p2_elasticity_adjoint(int32_t start, int32_t end, double *ubar, double *Xbar,
                                      const double *Rbar, const double *X, const double *u,
                                      const int32_t *c2n, const int32_t *c2v, double lam,
                                      double mu, const op2hip_exec *ex, void *stream) {
  hipStream_t s = stream_of(stream);
  if (ex && ex->strategy == OP2HIP_STAGED)
    return launch_staged(start, end, p2::AdjointStaged<double>{ubar, Xbar, Rbar, X, u, lam, mu},
                         ex, s);
  const int bs = block_size_of(ex, 256);
  return for_each_launch(ex, start, end, [&](int32_t lo, int32_t hi, const int32_t *list,
                                             bool atomic) -> int {
    if (atomic)
      p2::adjoint_kernel<double, true><<<ceil_div(hi - lo, bs), bs, 0, s>>>(
          lo, hi, list, ubar, Xbar, Rbar, X, u, c2n, c2v, lam, mu);
    else
      p2::adjoint_kernel<double, false><<<ceil_div(hi - lo, bs), bs, 0, s>>>(
          lo, hi, list, ubar, Xbar, Rbar, X, u, c2n, c2v, lam, mu);
    OP2HIP_LAUNCHED();
    return OP2HIP_OK;
  });
}

}  // extern "C"
