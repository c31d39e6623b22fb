"""Element kernels for the P1 triangle and P2 tetrahedron workloads.

Each function returns an ``op2.Kernel`` whose C string is an ordinary PyOP2
local kernel (the form ``op2.Kernel(code, name)`` takes, buffer shapes as the
reference's wrapper packs them, ``pyop2/codegen/builder.py:340-347`` for Dats
and ``:519-529`` for Mats).  The same C string is what a CPU PyOP2 run would
compile.  On the device these kernels are not compiled from the string: their
code hash is registered with a hand-written HIP implementation (see
``pyop2_amd.registry``), so ``op2.par_loop`` launches the HIP kernel.

Argument conventions (INC buffers arrive zeroed, as in the reference):

* ``p1_lumped_mass``      ``(double M[3], const double X[3][2])``
* ``p1_laplace_residual`` ``(double R[3], const double X[3][2], const double u[3])``
* ``p1_laplace_matrix``   ``(double A[3][3], const double X[3][2])``
* ``p2_elasticity_residual`` ``(double R[10][3], const double X[4][3], const double u[10][3])``
* ``p2_elasticity_matrix``   ``(double A[30][30], const double X[4][3])``
* ``p2_elasticity_fused``    ``(double A[30][30], double R[10][3], const double X[4][3], const double u[10][3])``
* ``p2_elasticity_adjoint``  ``(double ub[10][3], double Xb[4][3], const double Rbar[10][3], const double X[4][3], const double u[10][3])``
* ``p2_elasticity_tlm``      ``(double dR[10][3], const double X[4][3], const double u[10][3], const double dX[4][3], const double du[10][3])``

The P2 kernels bake the Lame parameters in as literals (as form compilers
do).  The 4-point Gauss rule (a = 0.5854101966249685, b = 0.1381966011250105,
w = 1/24 on the reference tetrahedron) integrates the degree-2 integrands of
the P2 stiffness exactly.
"""
import functools
import hashlib
import re

__all__ = ["p1_lumped_mass", "p1_laplace_residual", "p1_laplace_matrix",
           "p2_elasticity_residual", "p2_elasticity_matrix", "p2_elasticity_fused",
           "p2_elasticity_adjoint", "p2_elasticity_tlm",
           "kernel_source", "code_digest", "QUAD_A", "QUAD_B", "QUAD_W"]

QUAD_A = 0.5854101966249685
QUAD_B = 0.1381966011250105
QUAD_W = 1.0 / 24.0

P1_LUMPED_MASS = r"""
static void p1_lumped_mass(double M[3], const double X[3][2])
{
  const double det = (X[1][0] - X[0][0]) * (X[2][1] - X[0][1])
                   - (X[2][0] - X[0][0]) * (X[1][1] - X[0][1]);
  const double m = fabs(det) / 6.0;
  M[0] += m;
  M[1] += m;
  M[2] += m;
}
"""

_P1_GEOMETRY = r"""
  double b[3], c[3];
  b[0] = X[1][1] - X[2][1];
  b[1] = X[2][1] - X[0][1];
  b[2] = X[0][1] - X[1][1];
  c[0] = X[2][0] - X[1][0];
  c[1] = X[0][0] - X[2][0];
  c[2] = X[1][0] - X[0][0];
  const double det = (X[1][0] - X[0][0]) * (X[2][1] - X[0][1])
                   - (X[2][0] - X[0][0]) * (X[1][1] - X[0][1]);
  const double s = 1.0 / (2.0 * fabs(det));
"""

P1_LAPLACE_RESIDUAL = r"""
static void p1_laplace_residual(double R[3], const double X[3][2], const double u[3])
{""" + _P1_GEOMETRY + r"""
  for (int i = 0; i < 3; ++i) {
    double acc = 0.0;
    for (int j = 0; j < 3; ++j)
      acc += (b[i] * b[j] + c[i] * c[j]) * s * u[j];
    R[i] += acc;
  }
}
"""

P1_LAPLACE_MATRIX = r"""
static void p1_laplace_matrix(double A[3][3], const double X[3][2])
{""" + _P1_GEOMETRY + r"""
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      A[i][j] += (b[i] * b[j] + c[i] * c[j]) * s;
}
"""

# Affine tetrahedron geometry: J[r][c] = X[c+1][r] - X[0][r], K = J^{-1}
# (K[m][k] = d xi_m / d x_k), barycentric gradients gl[0..3], wdet = w |det J|.
_P2_GEOMETRY = r"""
  const double qa = %(qa)r, qb = %(qb)r, qw = %(qw)r;
  double J[3][3], K[3][3], gl[4][3];
  for (int r = 0; r < 3; ++r)
    for (int cc = 0; cc < 3; ++cc)
      J[r][cc] = X[cc + 1][r] - X[0][r];
  const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
  const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
  const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
  const double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
  const double idet = 1.0 / det;
  K[0][0] = c00 * idet;
  K[1][0] = c01 * idet;
  K[2][0] = c02 * idet;
  K[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * idet;
  K[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * idet;
  K[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * idet;
  K[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * idet;
  K[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * idet;
  K[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * idet;
  for (int k = 0; k < 3; ++k) {
    gl[0][k] = -(K[0][k] + K[1][k] + K[2][k]);
    gl[1][k] = K[0][k];
    gl[2][k] = K[1][k];
    gl[3][k] = K[2][k];
  }
  const double wdet = qw * fabs(det);
  static const int ea[6] = {0, 0, 0, 1, 1, 2};
  static const int eb[6] = {1, 2, 3, 2, 3, 3};
"""

# Physical gradients of the 10 P2 basis functions at quadrature point q.
_P2_BASIS_GRADS = r"""
    double L[4] = {qb, qb, qb, qb};
    L[q] = qa;
    double G[10][3];
    for (int i = 0; i < 4; ++i)
      for (int k = 0; k < 3; ++k)
        G[i][k] = (4.0 * L[i] - 1.0) * gl[i][k];
    for (int e = 0; e < 6; ++e)
      for (int k = 0; k < 3; ++k)
        G[4 + e][k] = 4.0 * (L[ea[e]] * gl[eb[e]][k] + L[eb[e]] * gl[ea[e]][k]);
"""

P2_ELASTICITY_RESIDUAL = r"""
static void p2_elasticity_residual(double R[10][3], const double X[4][3], const double u[10][3])
{
  const double lmbda = %(lam)r, mu = %(mu)r;""" + _P2_GEOMETRY + r"""
  for (int q = 0; q < 4; ++q) {""" + _P2_BASIS_GRADS + r"""
    double gu[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
    for (int j = 0; j < 10; ++j)
      for (int bb = 0; bb < 3; ++bb)
        for (int k = 0; k < 3; ++k)
          gu[bb][k] += u[j][bb] * G[j][k];
    const double dv = gu[0][0] + gu[1][1] + gu[2][2];
    double s[3][3];
    for (int a = 0; a < 3; ++a)
      for (int k = 0; k < 3; ++k)
        s[a][k] = mu * (gu[a][k] + gu[k][a]) + (a == k ? lmbda * dv : 0.0);
    for (int i = 0; i < 10; ++i)
      for (int a = 0; a < 3; ++a) {
        double acc = 0.0;
        for (int k = 0; k < 3; ++k)
          acc += s[a][k] * G[i][k];
        R[i][a] += wdet * acc;
      }
  }
}
"""

P2_ELASTICITY_MATRIX = r"""
static void p2_elasticity_matrix(double A[30][30], const double X[4][3])
{
  const double lmbda = %(lam)r, mu = %(mu)r;""" + _P2_GEOMETRY + r"""
  for (int q = 0; q < 4; ++q) {""" + _P2_BASIS_GRADS + r"""
    for (int i = 0; i < 10; ++i)
      for (int j = 0; j < 10; ++j) {
        const double dot = G[i][0] * G[j][0] + G[i][1] * G[j][1] + G[i][2] * G[j][2];
        for (int a = 0; a < 3; ++a)
          for (int bb = 0; bb < 3; ++bb)
            A[3 * i + a][3 * j + bb] += wdet * (lmbda * G[i][a] * G[j][bb]
                + mu * (G[i][bb] * G[j][a] + (a == bb ? dot : 0.0)));
      }
  }
}
"""


P2_ELASTICITY_FUSED = r"""
static void p2_elasticity_fused(double A[30][30], double R[10][3], const double X[4][3], const double u[10][3])
{
  const double lmbda = %(lam)r, mu = %(mu)r;""" + _P2_GEOMETRY + r"""
  for (int q = 0; q < 4; ++q) {""" + _P2_BASIS_GRADS + r"""
    for (int i = 0; i < 10; ++i)
      for (int j = 0; j < 10; ++j) {
        const double dot = G[i][0] * G[j][0] + G[i][1] * G[j][1] + G[i][2] * G[j][2];
        for (int a = 0; a < 3; ++a)
          for (int bb = 0; bb < 3; ++bb)
            A[3 * i + a][3 * j + bb] += wdet * (lmbda * G[i][a] * G[j][bb]
                + mu * (G[i][bb] * G[j][a] + (a == bb ? dot : 0.0)));
      }
  }
  for (int i = 0; i < 10; ++i)
    for (int a = 0; a < 3; ++a) {
      double acc = 0.0;
      for (int j = 0; j < 10; ++j)
        for (int bb = 0; bb < 3; ++bb)
          acc += A[3 * i + a][3 * j + bb] * u[j][bb];
      R[i][a] += acc;
    }
}
"""


# Adjoint of the residual for J-type functionals: with E = Rbar_e,
#   ub += K_e^T E                    (= the residual operator applied to E),
#   Xb += d(E . R_e(u, X))/dX,  d/dJ = wdet sum_q [phi I - gu^T sE - gE^T su] K^T,
# phi = su : gE, J[r][c] = X[c+1][r] - X[0][r] (X[m+1] gets column m, X[0] minus the sum).
P2_ELASTICITY_ADJOINT = r"""
static void p2_elasticity_adjoint(double ub[10][3], double Xb[4][3], const double E[10][3],
                                  const double X[4][3], const double u[10][3])
{
  const double lmbda = %(lam)r, mu = %(mu)r;""" + _P2_GEOMETRY + r"""
  double P[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
  for (int q = 0; q < 4; ++q) {""" + _P2_BASIS_GRADS + r"""
    double gu[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
    double ge[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
    for (int j = 0; j < 10; ++j)
      for (int bb = 0; bb < 3; ++bb)
        for (int k = 0; k < 3; ++k) {
          gu[bb][k] += u[j][bb] * G[j][k];
          ge[bb][k] += E[j][bb] * G[j][k];
        }
    const double dvu = gu[0][0] + gu[1][1] + gu[2][2];
    const double dve = ge[0][0] + ge[1][1] + ge[2][2];
    double su[3][3], se[3][3];
    for (int a = 0; a < 3; ++a)
      for (int k = 0; k < 3; ++k) {
        su[a][k] = mu * (gu[a][k] + gu[k][a]) + (a == k ? lmbda * dvu : 0.0);
        se[a][k] = mu * (ge[a][k] + ge[k][a]) + (a == k ? lmbda * dve : 0.0);
      }
    for (int i = 0; i < 10; ++i)
      for (int a = 0; a < 3; ++a) {
        double acc = 0.0;
        for (int k = 0; k < 3; ++k)
          acc += se[a][k] * G[i][k];
        ub[i][a] += wdet * acc;
      }
    double phi = 0.0;
    for (int a = 0; a < 3; ++a)
      for (int k = 0; k < 3; ++k)
        phi += su[a][k] * ge[a][k];
    for (int l = 0; l < 3; ++l)
      for (int k = 0; k < 3; ++k) {
        double acc = (l == k) ? phi : 0.0;
        for (int bb = 0; bb < 3; ++bb)
          acc -= gu[bb][l] * se[bb][k] + ge[bb][l] * su[bb][k];
        P[l][k] += acc;
      }
  }
  for (int l = 0; l < 3; ++l) {
    double rowsum = 0.0;
    for (int m = 0; m < 3; ++m) {
      const double v = wdet * (P[l][0] * K[m][0] + P[l][1] * K[m][1] + P[l][2] * K[m][2]);
      Xb[m + 1][l] += v;
      rowsum += v;
    }
    Xb[0][l] -= rowsum;
  }
}
"""

# Tangent-linear residual: dR = dR/du du + dR/dX dX, with dK = -K dJ K,
# d|det J| = |det J| tr(K dJ) and grad phi linear in the barycentric gradients.
P2_ELASTICITY_TLM = r"""
static void p2_elasticity_tlm(double dR[10][3], const double X[4][3], const double u[10][3],
                              const double dX[4][3], const double du[10][3])
{
  const double lmbda = %(lam)r, mu = %(mu)r;""" + _P2_GEOMETRY + r"""
  double dJ[3][3], dK[3][3], dgl[4][3];
  for (int r = 0; r < 3; ++r)
    for (int cc = 0; cc < 3; ++cc)
      dJ[r][cc] = dX[cc + 1][r] - dX[0][r];
  double trKdJ = 0.0;
  for (int m = 0; m < 3; ++m)
    for (int r = 0; r < 3; ++r)
      trKdJ += K[m][r] * dJ[r][m];
  for (int m = 0; m < 3; ++m)
    for (int k = 0; k < 3; ++k) {
      double acc = 0.0;
      for (int r = 0; r < 3; ++r)
        for (int cc = 0; cc < 3; ++cc)
          acc += K[m][r] * dJ[r][cc] * K[cc][k];
      dK[m][k] = -acc;
    }
  for (int k = 0; k < 3; ++k) {
    dgl[0][k] = -(dK[0][k] + dK[1][k] + dK[2][k]);
    dgl[1][k] = dK[0][k];
    dgl[2][k] = dK[1][k];
    dgl[3][k] = dK[2][k];
  }
  const double dwdet = wdet * trKdJ;
  for (int q = 0; q < 4; ++q) {""" + _P2_BASIS_GRADS + r"""
    double dG[10][3];
    for (int i = 0; i < 4; ++i)
      for (int k = 0; k < 3; ++k)
        dG[i][k] = (4.0 * L[i] - 1.0) * dgl[i][k];
    for (int e = 0; e < 6; ++e)
      for (int k = 0; k < 3; ++k)
        dG[4 + e][k] = 4.0 * (L[ea[e]] * dgl[eb[e]][k] + L[eb[e]] * dgl[ea[e]][k]);
    double gu[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
    double dgu[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
    for (int j = 0; j < 10; ++j)
      for (int bb = 0; bb < 3; ++bb)
        for (int k = 0; k < 3; ++k) {
          gu[bb][k] += u[j][bb] * G[j][k];
          dgu[bb][k] += du[j][bb] * G[j][k] + u[j][bb] * dG[j][k];
        }
    const double dv = gu[0][0] + gu[1][1] + gu[2][2];
    const double ddv = dgu[0][0] + dgu[1][1] + dgu[2][2];
    double s[3][3], ds[3][3];
    for (int a = 0; a < 3; ++a)
      for (int k = 0; k < 3; ++k) {
        s[a][k] = mu * (gu[a][k] + gu[k][a]) + (a == k ? lmbda * dv : 0.0);
        ds[a][k] = mu * (dgu[a][k] + dgu[k][a]) + (a == k ? lmbda * ddv : 0.0);
      }
    for (int i = 0; i < 10; ++i)
      for (int a = 0; a < 3; ++a) {
        double acc0 = 0.0, acc1 = 0.0;
        for (int k = 0; k < 3; ++k) {
          acc0 += s[a][k] * G[i][k];
          acc1 += ds[a][k] * G[i][k] + s[a][k] * dG[i][k];
        }
        dR[i][a] += dwdet * acc0 + wdet * acc1;
      }
  }
}
"""


def kernel_source(family, lam=1.0, mu=0.5):
    """C source of element kernel ``family`` (P2 families take lam, mu)."""
    params = {"qa": QUAD_A, "qb": QUAD_B, "qw": QUAD_W,
              "lam": float(lam), "mu": float(mu)}
    table = {
        "p1_lumped_mass": P1_LUMPED_MASS,
        "p1_laplace_residual": P1_LAPLACE_RESIDUAL,
        "p1_laplace_matrix": P1_LAPLACE_MATRIX,
        "p2_elasticity_residual": P2_ELASTICITY_RESIDUAL % params,
        "p2_elasticity_matrix": P2_ELASTICITY_MATRIX % params,
        "p2_elasticity_fused": P2_ELASTICITY_FUSED % params,
        "p2_elasticity_adjoint": P2_ELASTICITY_ADJOINT % params,
        "p2_elasticity_tlm": P2_ELASTICITY_TLM % params,
    }
    try:
        return table[family]
    except KeyError:
        raise ValueError(f"unknown element kernel family {family!r}") from None


# C tokens: string and character literals, identifiers, numbers, operators
_TOKEN = re.compile(r'"(?:\\.|[^"\\])*"|\'(?:\\.|[^\'\\])*\'|[A-Za-z_]\w*'
                    r'|(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?[fFlLuU]*'
                    r'|->|\+\+|--|<<=|>>=|<<|>>|[-+*/%&|^!<>=]=|&&|\|\||[-+*/%&|^!~<>=?:;,.(){}\[\]#]')


def normalised_source(code, name):
    """The kernel's C tokens, comments and layout dropped and its function
    name replaced by a placeholder: two sources that differ only in
    formatting, comments or the kernel's name give the same text."""
    code = re.sub(r"/\*.*?\*/", " ", code, flags=re.S)
    code = re.sub(r"//[^\n]*", " ", code)
    toks = _TOKEN.findall(code)
    if "".join(toks) != re.sub(r"\s+", "", code):
        # a character no token covers: keep the text itself (nothing dropped)
        return name + "\0" + code
    return " ".join("@kernel" if t == name else t for t in toks)


@functools.lru_cache(maxsize=1024)
def code_digest(code, name):
    """Digest of a kernel's normalised source (``normalised_source``), the
    key of the HIP registry (kept per source text: a program that builds the
    same kernel for every loop call normalises it once)."""
    return hashlib.sha256(normalised_source(code, name).encode()).hexdigest()


def _make(family, **params):
    from pyop2_amd import registry
    from pyop2_amd.local_kernel import Kernel
    code = kernel_source(family, **params)
    registry.register_element_kernel(code, family, family, params)
    return Kernel(code, family)


def p1_lumped_mass():
    """Lumped P1 mass: |T|/3 to each vertex."""
    return _make("p1_lumped_mass")


def p1_laplace_residual():
    """P1 Laplace residual K_e u_e, K_e = (b b^T + c c^T) / (4 |T|)."""
    return _make("p1_laplace_residual")


def p1_laplace_matrix():
    """P1 Laplace element stiffness K_e (3x3)."""
    return _make("p1_laplace_matrix")


def p2_elasticity_residual(lam=1.0, mu=0.5):
    """P2 isotropic linear elasticity residual sigma(u):grad(phi_i e_a)."""
    return _make("p2_elasticity_residual", lam=float(lam), mu=float(mu))


def p2_elasticity_matrix(lam=1.0, mu=0.5):
    """P2 isotropic linear elasticity element matrix (30x30, dof 3*node+comp)."""
    return _make("p2_elasticity_matrix", lam=float(lam), mu=float(mu))


def p2_elasticity_fused(lam=1.0, mu=0.5):
    """Element matrix A_e and residual R_e = A_e u_e in one kernel."""
    return _make("p2_elasticity_fused", lam=float(lam), mu=float(mu))


def p2_elasticity_adjoint(lam=1.0, mu=0.5):
    """Adjoint of the residual loop: (ub INC, Xb INC, Rbar, X, u)."""
    return _make("p2_elasticity_adjoint", lam=float(lam), mu=float(mu))


def p2_elasticity_tlm(lam=1.0, mu=0.5):
    """Tangent-linear residual: (dR INC, X, u, dX, du) (generated device path)."""
    from pyop2_amd.local_kernel import Kernel
    return Kernel(kernel_source("p2_elasticity_tlm", lam=lam, mu=mu), "p2_elasticity_tlm")
