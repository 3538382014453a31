// TEST INFRASTRUCTURE ONLY -- CPU FP64 restatement of HOBAK's backward-Euler Newton step.
//
// This is the checker for the HIP path: the GPU library is compared against it in tests/ and in
// __graft_entry__.smoke(), and bench.py may time it as a CPU baseline.  The product never links
// it.  It is written from the reference's mathematics and follows the reference's order of
// operations where that order changes round-off (citations are relative to the reference root).
//
// Pinned by: the literal reference's outputs (tests/golden/*, made by tests/golden/make_golden.py
// from oracle/_ref), and by closed-form / invariant checks in tests/test_oracle_*.py.  The two
// parts the reference cannot express are pinned without it:
//   - block-Jacobi PCG: an independent textbook PCG in numpy that first reproduces the
//     reference's Jacobi iteration counts (tests/test_oracle_pcg.py);
//   - the penalty ground plane: the closed-form step of a mesh with zero elastic stiffness
//     (tests/test_oracle_math.py::test_penalty_ground_closed_form).
#include "hobak_oracle.h"

#include <omp.h>

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>
#include <map>
#include <array>

namespace {

typedef double M3[3][3];  // row-major [row][col]

inline void m3_from_colmajor(const double* a, M3 m) {
  for (int j = 0; j < 3; j++)
    for (int i = 0; i < 3; i++) m[i][j] = a[3 * j + i];
}
inline void m3_to_colmajor(const M3 m, double* a) {
  for (int j = 0; j < 3; j++)
    for (int i = 0; i < 3; i++) a[3 * j + i] = m[i][j];
}
inline void m3_copy(const M3 a, M3 b) { std::memcpy(b, a, sizeof(M3)); }
inline void m3_identity(M3 a) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) a[i][j] = (i == j) ? 1.0 : 0.0;
}
// C = A * B with ((a0 b0 + a1 b1) + a2 b2) summation.
inline void m3_mul(const M3 A, const M3 B, M3 C) {
  M3 t;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) t[i][j] = (A[i][0] * B[0][j] + A[i][1] * B[1][j]) + A[i][2] * B[2][j];
  m3_copy(t, C);
}
// C = A * B^T
inline void m3_mul_bt(const M3 A, const M3 B, M3 C) {
  M3 t;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) t[i][j] = (A[i][0] * B[j][0] + A[i][1] * B[j][1]) + A[i][2] * B[j][2];
  m3_copy(t, C);
}

// 3x3 determinant in the cofactor-expansion order of Eigen's determinant_impl<3>
// (ext/eigen-3.3.7/Eigen/src/LU/Determinant.h:18-23, 65-67).
inline double det3(const M3 m) {
  auto h = [&](int a, int b, int c) { return m[0][a] * (m[1][b] * m[2][c] - m[1][c] * m[2][b]); };
  return h(0, 1, 2) - h(1, 0, 2) + h(2, 0, 1);
}

// 3x3 inverse by cofactors (ext/eigen-3.3.7/Eigen/src/LU/InverseImpl.h:126-169).
inline void inv3(const M3 m, M3 r) {
  auto cof = [&](int i, int j) {
    const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
    return m[i1][j1] * m[i2][j2] - m[i1][j2] * m[i2][j1];
  };
  const double c0 = cof(0, 0), c1 = cof(1, 0), c2 = cof(2, 0);
  const double det = (c0 * m[0][0] + c1 * m[1][0]) + c2 * m[2][0];
  const double invdet = 1.0 / det;
  r[0][0] = c0 * invdet;
  r[0][1] = c1 * invdet;
  r[0][2] = c2 * invdet;
  for (int row = 1; row < 3; row++)
    for (int k = 0; k < 3; k++) r[row][k] = cof(k, row) * invdet;
}

// ---------------------------------------------------------------------------------------------
// Two-sided Jacobi SVD of a 3x3, following the steps of Eigen 3.3.7's JacobiSVD with
// NoQRPreconditioner (ext/eigen-3.3.7/Eigen/src/SVD/JacobiSVD.h:660-780), its 2x2 kernel
// (misc/RealSvd2x2.h:19-51) and Jacobi rotation construction (Jacobi/Jacobi.h:83-113).
// Rotations (c, s) act on a pair (x, y) as x' = c x + s y, y' = -s x + c y ("left") and
// x' = c x - s y, y' = s x + c y ("right", i.e. with the transposed rotation).
// ---------------------------------------------------------------------------------------------
inline void make_jacobi(double x, double y, double z, double& c, double& s) {
  const double deno = 2.0 * std::fabs(y);
  if (deno < DBL_MIN) {
    c = 1.0;
    s = 0.0;
    return;
  }
  const double tau = (x - z) / deno;
  const double w = std::sqrt(tau * tau + 1.0);
  const double t = (tau > 0.0) ? 1.0 / (tau + w) : 1.0 / (tau - w);
  const double sign_t = t > 0.0 ? 1.0 : -1.0;
  const double n = 1.0 / std::sqrt(t * t + 1.0);
  s = -sign_t * (y / std::fabs(y)) * std::fabs(t) * n;
  c = n;
}

void jacobi_svd3(const M3 F, M3 U, double S[3], M3 V) {
  double scale = 0.0;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) scale = std::max(scale, std::fabs(F[i][j]));
  if (scale == 0.0) scale = 1.0;
  M3 W;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) W[i][j] = F[i][j] / scale;
  m3_identity(U);
  m3_identity(V);
  const double precision = 2.0 * DBL_EPSILON;
  double maxDiag = std::max(std::max(std::fabs(W[0][0]), std::fabs(W[1][1])), std::fabs(W[2][2]));
  bool finished = false;
  int sweeps = 0;
  while (!finished && sweeps < 64) {
    finished = true;
    sweeps++;
    for (int p = 1; p < 3; p++) {
      for (int q = 0; q < p; q++) {
        const double thr = std::max(DBL_MIN, precision * maxDiag);
        if (std::fabs(W[p][q]) > thr || std::fabs(W[q][p]) > thr) {
          finished = false;
          // 2x2 real SVD of [[W_pp, W_pq], [W_qp, W_qq]]: first symmetrise with rot1 ...
          double m00 = W[p][p], m01 = W[p][q], m10 = W[q][p], m11 = W[q][q];
          double c1, s1;
          const double t = m00 + m11;
          const double d = m10 - m01;
          if (std::fabs(d) < DBL_MIN) {
            s1 = 0.0;
            c1 = 1.0;
          } else {
            const double u = t / d;
            const double tmp = std::sqrt(1.0 + u * u);
            s1 = 1.0 / tmp;
            c1 = u / tmp;
          }
          {
            const double x0 = m00, y0 = m10, x1 = m01, y1 = m11;
            m00 = c1 * x0 + s1 * y0;
            m10 = -s1 * x0 + c1 * y0;
            m01 = c1 * x1 + s1 * y1;
            m11 = -s1 * x1 + c1 * y1;
          }
          // ... then diagonalise the symmetric result with a Jacobi rotation jr.
          double cr, sr;
          make_jacobi(m00, m01, m11, cr, sr);
          // jl = rot1 * jr^T
          const double cl = c1 * cr - s1 * (-sr);
          const double sl = c1 * (-sr) + s1 * cr;
          // W <- jl applied to rows p, q
          for (int i = 0; i < 3; i++) {
            const double x = W[p][i], y = W[q][i];
            W[p][i] = cl * x + sl * y;
            W[q][i] = -sl * x + cl * y;
          }
          // U <- U * jl^T (columns p, q)
          for (int i = 0; i < 3; i++) {
            const double x = U[i][p], y = U[i][q];
            U[i][p] = cl * x + sl * y;
            U[i][q] = -sl * x + cl * y;
          }
          // W <- W * jr, V <- V * jr (columns p, q)
          for (int i = 0; i < 3; i++) {
            const double x = W[i][p], y = W[i][q];
            W[i][p] = cr * x - sr * y;
            W[i][q] = sr * x + cr * y;
          }
          for (int i = 0; i < 3; i++) {
            const double x = V[i][p], y = V[i][q];
            V[i][p] = cr * x - sr * y;
            V[i][q] = sr * x + cr * y;
          }
          maxDiag = std::max(maxDiag, std::max(std::fabs(W[p][p]), std::fabs(W[q][q])));
        }
      }
    }
  }
  for (int i = 0; i < 3; i++) {
    const double a = W[i][i];
    S[i] = std::fabs(a);
    if (a < 0.0)
      for (int r = 0; r < 3; r++) U[r][i] = -U[r][i];
  }
  for (int i = 0; i < 3; i++) S[i] *= scale;
  // descending selection sort, first maximum wins ties
  for (int i = 0; i < 3; i++) {
    int pos = i;
    for (int k = i + 1; k < 3; k++)
      if (S[k] > S[pos]) pos = k;
    if (S[pos] == 0.0) break;
    if (pos != i) {
      std::swap(S[i], S[pos]);
      for (int r = 0; r < 3; r++) {
        std::swap(U[r][i], U[r][pos]);
        std::swap(V[r][i], V[r][pos]);
      }
    }
  }
}

// "Rotation variant" SVD: reflections are folded into Sigma[2] (src/util/MATRIX_UTIL.cpp:76-95).
void svd_rv(const M3 F, M3 U, double S[3], M3 V) {
  jacobi_svd3(F, U, S, V);
  M3 UVt;
  m3_mul_bt(U, V, UVt);
  const double L22 = det3(UVt);
  const double detU = det3(U);
  const double detV = det3(V);
  if (detU < 0.0 && detV > 0)
    for (int r = 0; r < 3; r++) U[r][2] *= L22;
  if (detU > 0.0 && detV < 0.0)
    for (int r = 0; r < 3; r++) V[r][2] *= L22;
  S[2] = S[2] * L22;
}

// Cyclic Jacobi eigensolver for a symmetric 3x3 (eigenvalues ascending, eigenvectors in columns).
// The clamped Hessian only needs the spectral projector, so any accurate symmetric eigensolver
// gives the same result as the reference's SelfAdjointEigenSolver (src/Hyperelastic/Volume/SNH.cpp:130)
// to round-off.
void sym_eig3(const M3 Ain, double w[3], M3 Q) {
  M3 A;
  m3_copy(Ain, A);
  m3_identity(Q);
  for (int sweep = 0; sweep < 50; sweep++) {
    const double off = A[0][1] * A[0][1] + A[0][2] * A[0][2] + A[1][2] * A[1][2];
    const double dia = A[0][0] * A[0][0] + A[1][1] * A[1][1] + A[2][2] * A[2][2];
    if (off <= 1e-40 * dia || off == 0.0) break;
    for (int p = 0; p < 2; p++)
      for (int q = p + 1; q < 3; q++) {
        if (A[p][q] == 0.0) continue;
        const double theta = (A[q][q] - A[p][p]) / (2.0 * A[p][q]);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0);
        const double s = t * c;
        for (int k = 0; k < 3; k++) {
          const double akp = A[k][p], akq = A[k][q];
          A[k][p] = c * akp - s * akq;
          A[k][q] = s * akp + c * akq;
        }
        for (int k = 0; k < 3; k++) {
          const double apk = A[p][k], aqk = A[q][k];
          A[p][k] = c * apk - s * aqk;
          A[q][k] = s * apk + c * aqk;
        }
        for (int k = 0; k < 3; k++) {
          const double qkp = Q[k][p], qkq = Q[k][q];
          Q[k][p] = c * qkp - s * qkq;
          Q[k][q] = s * qkp + c * qkq;
        }
      }
  }
  for (int i = 0; i < 3; i++) w[i] = A[i][i];
  for (int i = 0; i < 3; i++)
    for (int k = i + 1; k < 3; k++)
      if (w[k] < w[i]) {
        std::swap(w[i], w[k]);
        for (int r = 0; r < 3; r++) std::swap(Q[r][i], Q[r][k]);
      }
}

// column-major flatten (src/util/MATRIX_UTIL.cpp:33-43): v[3 j + i] = A(i, j)
inline void flatten(const M3 A, double v[9]) {
  for (int j = 0; j < 3; j++)
    for (int i = 0; i < 3; i++) v[3 * j + i] = A[i][j];
}

// ∂J/∂F columns: f1 x f2, f2 x f0, f0 x f1 (src/util/MATRIX_UTIL.cpp:561-568)
void pJpF(const M3 F, M3 P) {
  double f[3][3];
  for (int c = 0; c < 3; c++)
    for (int r = 0; r < 3; r++) f[c][r] = F[r][c];
  auto cross = [](const double* a, const double* b, double* o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
  };
  double c0[3], c1[3], c2[3];
  cross(f[1], f[2], c0);
  cross(f[2], f[0], c1);
  cross(f[0], f[1], c2);
  for (int r = 0; r < 3; r++) {
    P[r][0] = c0[r];
    P[r][1] = c1[r];
    P[r][2] = c2[r];
  }
}

struct Material {
  int kind;  // 0 SNH, 1 ARAP
  double mu, lambda;  // for SNH: reparameterised lambda' = lambda + mu
  double alpha;       // SNH: 1 + mu / lambda'
  Material(int k, double m, double l) : kind(k), mu(m) {
    if (k == 0) {
      lambda = l + m;  // src/Hyperelastic/Volume/SNH.cpp:23-32
      alpha = 1.0 + mu / lambda;
    } else {
      lambda = l;  // ARAP keeps but never uses lambda (src/Hyperelastic/Volume/ARAP.cpp:26-29)
      alpha = 0.0;
    }
  }
};

double psi(const Material& m, const M3 F) {
  if (m.kind == 0) {  // SNH.cpp:34-43
    double Ic = 0;
    for (int j = 0; j < 3; j++)
      for (int i = 0; i < 3; i++) Ic += F[i][j] * F[i][j];
    const double Jm1 = det3(F) - m.alpha;
    return 0.5 * (m.mu * (Ic - 3.0) + m.lambda * Jm1 * Jm1);
  }
  M3 U, V;
  double S[3];
  svd_rv(F, U, S, V);  // ARAP.cpp:39-43: mu ||F - R||^2
  M3 R;
  m3_mul_bt(U, V, R);
  double s = 0;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) s += (F[i][j] - R[i][j]) * (F[i][j] - R[i][j]);
  return m.mu * s;
}

void pk1(const Material& m, const M3 F, M3 P) {
  if (m.kind == 0) {  // SNH.cpp:45-50: mu F + lambda' (J - alpha) dJ/dF
    M3 pj;
    pJpF(F, pj);
    const double Jm = det3(F) - m.alpha;
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) P[i][j] = m.mu * F[i][j] + m.lambda * Jm * pj[i][j];
    return;
  }
  // ARAP through HYPERELASTIC::PK1(F) (HYPERELASTIC.cpp:61-67) -> ARAP::PK1 (ARAP.cpp:48-55):
  // P = R * 2 mu (S - I), R = U V^T, S = V Sigma V^T.
  M3 U, V;
  double S[3];
  svd_rv(F, U, S, V);
  M3 R, VS, Sm;
  m3_mul_bt(U, V, R);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) VS[i][j] = V[i][j] * S[j];
  m3_mul_bt(VS, V, Sm);
  M3 T;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) T[i][j] = 2.0 * m.mu * (Sm[i][j] - (i == j ? 1.0 : 0.0));
  m3_mul(R, T, P);
}

// Unclamped dP/dF (SNH.cpp:52-76 and ARAP.cpp:60-105); used by finite-difference tests only.
void hessian(const Material& m, const M3 F, double H[9][9]) {
  if (m.kind == 0) {
    M3 pj;
    pJpF(F, pj);
    double g[9];
    flatten(pj, g);
    const double scale = m.lambda * (det3(F) - m.alpha);
    for (int a = 0; a < 9; a++)
      for (int b = 0; b < 9; b++) H[a][b] = (a == b ? m.mu : 0.0) + m.lambda * g[a] * g[b];
    // fractal cross product: blocks (i, j+3) = -f2hat, etc.
    auto hat = [&](int c, M3 o) {
      const double x = F[0][c] * scale, y = F[1][c] * scale, z = F[2][c] * scale;
      o[0][0] = 0; o[0][1] = -z; o[0][2] = y;
      o[1][0] = z; o[1][1] = 0; o[1][2] = -x;
      o[2][0] = -y; o[2][1] = x; o[2][2] = 0;
    };
    M3 f0, f1, f2;
    hat(0, f0);
    hat(1, f1);
    hat(2, f2);
    for (int j = 0; j < 3; j++)
      for (int i = 0; i < 3; i++) {
        H[i][j + 3] += -f2[i][j];
        H[i + 3][j] += f2[i][j];
        H[i][j + 6] += f1[i][j];
        H[i + 6][j] += -f1[i][j];
        H[i + 3][j + 6] += -f0[i][j];
        H[i + 6][j + 3] += f0[i][j];
      }
    return;
  }
  M3 U, V;
  double S[3];
  svd_rv(F, U, S, V);
  const double T[3][3][3] = {{{0, 0, 0}, {0, 0, -1}, {0, 1, 0}},
                             {{0, 0, 1}, {0, 0, 0}, {-1, 0, 0}},
                             {{0, 1, 0}, {-1, 0, 0}, {0, 0, 0}}};
  double q[3][9];
  for (int k = 0; k < 3; k++) {
    M3 t, UT, Q;
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) t[i][j] = T[k][i][j];
    m3_mul(U, t, UT);
    m3_mul_bt(UT, V, Q);
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) Q[i][j] *= 1.0 / std::sqrt(2.0);
    flatten(Q, q[k]);
  }
  const double lam[3] = {m.mu * 2.0 / (S[1] + S[2]), m.mu * 2.0 / (S[0] + S[2]),
                         m.mu * 2.0 / (S[0] + S[1])};
  for (int a = 0; a < 9; a++)
    for (int b = 0; b < 9; b++) {
      double v = (a == b) ? m.mu : 0.0;
      for (int k = 0; k < 3; k++) v -= lam[k] * q[k][a] * q[k][b];
      H[a][b] = 2.0 * v;
    }
}

// Twist (0-2) and flip (3-5) eigenvectors, src/util/MATRIX_UTIL.cpp:420-460.
void twist_flip_vectors(const M3 U, const M3 V, double Q[9][9]) {
  const double T[6][3][3] = {
      {{0, 0, 0}, {0, 0, -1}, {0, 1, 0}}, {{0, 0, 1}, {0, 0, 0}, {-1, 0, 0}},
      {{0, 1, 0}, {-1, 0, 0}, {0, 0, 0}}, {{0, 0, 0}, {0, 0, 1}, {0, 1, 0}},
      {{0, 0, 1}, {0, 0, 0}, {1, 0, 0}},  {{0, 1, 0}, {1, 0, 0}, {0, 0, 0}}};
  const double r = 1.0 / std::sqrt(2.0);
  for (int k = 0; k < 6; k++) {
    M3 t, UT, Qm;
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) t[i][j] = T[k][i][j];
    m3_mul(U, t, UT);
    m3_mul_bt(UT, V, Qm);
    double v[9];
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) Qm[i][j] = r * Qm[i][j];
    flatten(Qm, v);
    for (int a = 0; a < 9; a++) Q[a][k] = v[a];
  }
}

// Scaling eigenvectors U diag(q) V^T (src/util/MATRIX_UTIL.cpp:465-497).
void scaling_vectors(const M3 U, const M3 Qs, const M3 V, double Q[9][9]) {
  for (int k = 0; k < 3; k++) {
    M3 Ud, Qm;
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) Ud[i][j] = U[i][j] * Qs[j][k];
    m3_mul_bt(Ud, V, Qm);
    double v[9];
    flatten(Qm, v);
    for (int a = 0; a < 9; a++) Q[a][6 + k] = v[a];
  }
}

// Analytic eigen-clamped dP/dF: SNH.cpp:94-144, ARAP.cpp:113-135.
void clamped_hessian(const Material& m, const M3 U, const double S[3], const M3 V, double H[9][9]) {
  double lam[9];
  double Q[9][9];
  twist_flip_vectors(U, V, Q);
  if (m.kind == 0) {
    const double J = S[0] * S[1] * S[2];
    const double front = m.lambda * (J - 1.0) - m.mu;
    for (int i = 0; i < 3; i++) {
      lam[i] = front * S[i] + m.mu;
      lam[3 + i] = -front * S[i] + m.mu;
    }
    M3 A;
    const double s0s0 = S[0] * S[0], s1s1 = S[1] * S[1], s2s2 = S[2] * S[2];
    A[0][0] = m.mu + m.lambda * s1s1 * s2s2;
    A[1][1] = m.mu + m.lambda * s0s0 * s2s2;
    A[2][2] = m.mu + m.lambda * s0s0 * s1s1;
    const double fo = m.lambda * (2.0 * J - 1.0) - m.mu;
    A[0][1] = fo * S[2];
    A[0][2] = fo * S[1];
    A[1][2] = fo * S[0];
    A[1][0] = A[0][1];
    A[2][0] = A[0][2];
    A[2][1] = A[1][2];
    double w[3];
    M3 Qs;
    sym_eig3(A, w, Qs);
    for (int k = 0; k < 3; k++) lam[6 + k] = w[k];
    scaling_vectors(U, Qs, V, Q);
  } else {
    M3 I;
    m3_identity(I);
    scaling_vectors(U, I, V, Q);
    for (int k = 0; k < 9; k++) lam[k] = 2.0 * m.mu;
    lam[0] = m.mu * (2.0 - 4.0 / (S[1] + S[2]));
    lam[1] = m.mu * (2.0 - 4.0 / (S[0] + S[2]));
    lam[2] = m.mu * (2.0 - 4.0 / (S[0] + S[1]));
  }
  for (int k = 0; k < 9; k++)
    if (lam[k] < 0.0) lam[k] = 0.0;
  for (int a = 0; a < 9; a++)
    for (int b = 0; b < 9; b++) {
      double v = 0.0;
      for (int k = 0; k < 9; k++) v += Q[a][k] * lam[k] * Q[b][k];
      H[a][b] = v;
    }
}

// ∂F/∂x (9x12) from DmInv (src/Geometry/TET_MESH.cpp:197-263).
void pFpx(const M3 Dm, double P[9][12]) {
  const double m = Dm[0][0], n = Dm[0][1], o = Dm[0][2];
  const double p = Dm[1][0], q = Dm[1][1], r = Dm[1][2];
  const double s = Dm[2][0], t = Dm[2][1], u = Dm[2][2];
  const double t1 = -m - p - s, t2 = -n - q - t, t3 = -o - r - u;
  std::memset(P, 0, sizeof(double) * 108);
  const double c0[4] = {t1, m, p, s}, c1[4] = {t2, n, q, t}, c2[4] = {t3, o, r, u};
  for (int v = 0; v < 4; v++)
    for (int d = 0; d < 3; d++) {
      P[0 + d][3 * v + d] = c0[v];
      P[3 + d][3 * v + d] = c1[v];
      P[6 + d][3 * v + d] = c2[v];
    }
}

// Element force (TET_MESH.cpp:744-749) and element clamped stiffness (TET_MESH_FASTER.cpp:95-102)
// Pextra (may be NULL) is added to the elastic PK1 before the force, as computeInternalForces
// does with the damping PK1 (TET_MESH.cpp:831-834).
void element(const Material& mat, const M3 F, const M3 U, const double S[3], const M3 V,
             const double P9x12[9][12], double vol, double f[12], double H[12][12],
             const double (*Pextra)[3] = nullptr) {
  M3 P;
  pk1(mat, F, P);
  if (Pextra)
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) P[i][j] = P[i][j] + Pextra[i][j];
  double vp[9];
  flatten(P, vp);
  for (int k = 0; k < 12; k++) {
    double s = 0;
    for (int r = 0; r < 9; r++) s += P9x12[r][k] * vp[r];
    f[k] = -vol * s;
  }
  if (!H) return;
  double H9[9][9];
  clamped_hessian(mat, U, S, V, H9);
  for (int a = 0; a < 9; a++)
    for (int b = 0; b < 9; b++) H9[a][b] = -vol * H9[a][b];
  double T[12][9];
  for (int i = 0; i < 12; i++)
    for (int b = 0; b < 9; b++) {
      double s = 0;
      for (int a = 0; a < 9; a++) s += P9x12[a][i] * H9[a][b];
      T[i][b] = s;
    }
  for (int i = 0; i < 12; i++)
    for (int j = 0; j < 12; j++) {
      double s = 0;
      for (int b = 0; b < 9; b++) s += T[i][b] * P9x12[b][j];
      H[i][j] = s;
    }
}

// ---------------------------------------------------------------------------------------------
// Energy-based damping: GREEN_DAMPING (src/Damping/Volume/GREEN_DAMPING.cpp:38-97) on the rate of
// the Green strain, Edot = (Fdot^T F + F^T Fdot) / 2.
//   psi = mu |Edot|_F^2
//   P   = mu F (F^T Fdot + Fdot^T F)
//   dP/dFdot, column (i, j) (flattened column-major, index 3 j + i) = mu (F F^T E_ij + F E_ij^T F)
//   clampedHessian: the same matrix with its eigenvalues clamped at zero (MATRIX_UTIL.cpp:100-111)
// ---------------------------------------------------------------------------------------------
// Cyclic Jacobi eigensolver for a symmetric 9x9 (eigenvectors in columns).  As for sym_eig3, the
// clamp only needs the spectral projector, so any accurate symmetric eigensolver reproduces the
// reference's SelfAdjointEigenSolver result to round-off.
void sym_eig9(const double Ain[9][9], double w[9], double Q[9][9]) {
  double A[9][9];
  std::memcpy(A, Ain, sizeof(A));
  for (int i = 0; i < 9; i++)
    for (int j = 0; j < 9; j++) Q[i][j] = i == j ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 100; sweep++) {
    double off = 0.0, dia = 0.0;
    for (int i = 0; i < 9; i++) {
      dia += A[i][i] * A[i][i];
      for (int j = i + 1; j < 9; j++) off += A[i][j] * A[i][j];
    }
    if (off <= 1e-40 * dia || off == 0.0) break;
    for (int p = 0; p < 8; p++)
      for (int q = p + 1; q < 9; q++) {
        if (A[p][q] == 0.0) continue;
        const double theta = (A[q][q] - A[p][p]) / (2.0 * A[p][q]);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0);
        const double s = t * c;
        for (int k = 0; k < 9; k++) {
          const double akp = A[k][p], akq = A[k][q];
          A[k][p] = c * akp - s * akq;
          A[k][q] = s * akp + c * akq;
        }
        for (int k = 0; k < 9; k++) {
          const double apk = A[p][k], aqk = A[q][k];
          A[p][k] = c * apk - s * aqk;
          A[q][k] = s * apk + c * aqk;
        }
        for (int k = 0; k < 9; k++) {
          const double qkp = Q[k][p], qkq = Q[k][q];
          Q[k][p] = c * qkp - s * qkq;
          Q[k][q] = s * qkp + c * qkq;
        }
      }
  }
  for (int i = 0; i < 9; i++) w[i] = A[i][i];
}

inline void m3_tmul(const M3 A, const M3 B, M3 C) {  // C = A^T B
  M3 At;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) At[i][j] = A[j][i];
  m3_mul(At, B, C);
}

double green_psi(double mu, const M3 F, const M3 Fd) {
  M3 FdF, E;
  m3_tmul(Fd, F, FdF);  // Fdot^T F
  double s = 0.0;
  for (int j = 0; j < 3; j++)  // squaredNorm, column by column
    for (int i = 0; i < 3; i++) {
      E[i][j] = 0.5 * (FdF[i][j] + FdF[j][i]);
      s += E[i][j] * E[i][j];
    }
  return mu * s;
}

void green_pk1(double mu, const M3 F, const M3 Fd, M3 P) {
  M3 A, B, Sm, muF;
  m3_tmul(F, Fd, A);   // F^T Fdot
  m3_tmul(Fd, F, B);   // Fdot^T F
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      Sm[i][j] = A[i][j] + B[i][j];
      muF[i][j] = mu * F[i][j];
    }
  m3_mul(muF, Sm, P);  // (mu F) (F^T Fdot + Fdot^T F)
}

void green_hessian(double mu, const M3 F, double H[9][9]) {
  M3 FFt;
  m3_mul_bt(F, F, FFt);
  for (int j = 0; j < 3; j++)
    for (int i = 0; i < 3; i++) {
      // column 3 j + i: F F^T E_ij + F E_ij^T F;  (F F^T E_ij)(a, b) = FFt(a, i) [b == j],
      // (F E_ij^T F)(a, b) = F(a, j) F(i, b)
      const int col = 3 * j + i;
      for (int b = 0; b < 3; b++)
        for (int a = 0; a < 3; a++) {
          const double v = (b == j ? FFt[a][i] : 0.0) + F[a][j] * F[i][b];
          H[3 * b + a][col] = mu * v;
        }
    }
}

void green_clamped_hessian(double mu, const M3 F, double H[9][9]) {
  double H0[9][9], w[9], Q[9][9];
  green_hessian(mu, F, H0);
  sym_eig9(H0, w, Q);
  for (int k = 0; k < 9; k++)
    if (!(w[k] > 0.0)) w[k] = 0.0;
  for (int a = 0; a < 9; a++)
    for (int b = 0; b < 9; b++) {
      double v = 0.0;
      for (int k = 0; k < 9; k++) v += Q[a][k] * w[k] * Q[b][k];
      H[a][b] = v;
    }
}

// TET_MESH::computeFdots (TET_MESH.cpp:660-684): Fdot = [v1 - v0, v2 - v0, v3 - v0] DmInv.
void tet_fdot(const int* T, const double* vel, const M3 Di, M3 Fd) {
  M3 Vm;
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++) Vm[r][c] = vel[3 * T[c + 1] + r] - vel[3 * T[0] + r];
  m3_mul(Vm, Di, Fd);
}

// Element damping stiffness -V pFpx^T clampedHessian pFpx (TET_MESH_FASTER.cpp:144-150).
void element_damping(double mu, const M3 F, const double P9x12[9][12], double vol, double C[12][12]) {
  double H9[9][9];
  green_clamped_hessian(mu, F, H9);
  for (int a = 0; a < 9; a++)
    for (int b = 0; b < 9; b++) H9[a][b] = -vol * H9[a][b];
  double T[12][9];
  for (int i = 0; i < 12; i++)
    for (int b = 0; b < 9; b++) {
      double s = 0;
      for (int a = 0; a < 9; a++) s += P9x12[a][i] * H9[a][b];
      T[i][b] = s;
    }
  for (int i = 0; i < 12; i++)
    for (int j = 0; j < 12; j++) {
      double s = 0;
      for (int b = 0; b < 9; b++) s += T[i][b] * P9x12[b][j];
      C[i][j] = s;
    }
}

// Deterministic parallel sum: a fixed number of chunks, reduced in chunk order.
template <class Fn>
double det_sum(int n, Fn fn) {
  const int chunks = 64;
  double part[chunks];
#pragma omp parallel for schedule(static)
  for (int c = 0; c < chunks; c++) {
    const long lo = (long)n * c / chunks, hi = (long)n * (c + 1) / chunks;
    double s = 0;
    for (long i = lo; i < hi; i++) s += fn((int)i);
    part[c] = s;
  }
  double s = 0;
  for (int c = 0; c < chunks; c++) s += part[c];
  return s;
}

// A kinematic collision CUBE (src/Geometry/CUBE.cpp:23-193, KINEMATIC_SHAPE.h:55-73): world =
// R S local + t with S = scale I; the box is [-0.5, 0.5]^3 in local coordinates.  The products
// follow Eigen's left-to-right evaluation: worldVertexToLocal = (S^-1 R^T)(w - t) with S^-1 from
// Eigen's 3x3 cofactor inverse, localVertexToWorld = (R S) l + t.
struct Cube {
  double R[3][3];
  double s, sinv;
  double t[3];
  Cube(const double* center, double scale) {
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) R[i][j] = i == j ? 1.0 : 0.0;
    s = scale;
    sinv = (scale * scale) * (1.0 / ((scale * scale) * scale));
    for (int i = 0; i < 3; i++) t[i] = center[i];
  }
  void toLocal(const double* w, double* l) const {
    const double d[3] = {w[0] - t[0], w[1] - t[1], w[2] - t[2]};
    for (int i = 0; i < 3; i++) {
      const double m0 = sinv * R[0][i], m1 = sinv * R[1][i], m2 = sinv * R[2][i];
      l[i] = (m0 * d[0] + m1 * d[1]) + m2 * d[2];
    }
  }
  void toWorld(const double* l, double* w) const {
    for (int i = 0; i < 3; i++)
      w[i] = ((R[i][0] * s) * l[0] + (R[i][1] * s) * l[1]) + (R[i][2] * s) * l[2] + t[i];
  }
  void normalToWorld(const double* n, double* w) const {
    for (int i = 0; i < 3; i++) w[i] = (R[i][0] * n[0] + R[i][1] * n[1]) + R[i][2] * n[2];
  }
  static bool insideLocal(const double* l) {
    return l[0] <= 0.5 && l[0] >= -0.5 && l[1] <= 0.5 && l[1] >= -0.5 && l[2] <= 0.5 && l[2] >= -0.5;
  }
  bool inside(const double* w) const {  // CUBE.cpp:37-48
    double l[3];
    toLocal(w, l);
    return insideLocal(l);
  }
  double signedDistance(const double* w) const {  // CUBE.cpp:90-119
    double l[3];
    toLocal(w, l);
    if (insideLocal(l)) {
      const double xm = std::min(0.5 - l[0], l[0] - (-0.5));
      const double ym = std::min(0.5 - l[1], l[1] - (-0.5));
      const double zm = std::min(0.5 - l[2], l[2] - (-0.5));
      return -std::fabs(std::min(std::min(xm, ym), zm)) * s;
    }
    double d[3] = {0, 0, 0};
    for (int k = 0; k < 3; k++) {
      if (l[k] > 0.5)
        d[k] = l[k] - 0.5;
      else if (l[k] < -0.5)
        d[k] = -0.5 - l[k];
    }
    return std::sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]) * s;
  }
  // CUBE.cpp:124-191: the nearest face (first minimum of the six distances) in local coordinates
  void closest(const double* w, double* cl, double* nl) const {
    double l[3];
    toLocal(w, l);
    const double diffs[6] = {0.5 + l[0], 0.5 - l[0], 0.5 + l[1], 0.5 - l[1], 0.5 + l[2], 0.5 - l[2]};
    int mi = 0;
    double mf = diffs[0];
    for (int x = 1; x < 6; x++)
      if (diffs[x] < mf) {
        mf = diffs[x];
        mi = x;
      }
    const int axis = mi / 2;
    const double sign = (mi % 2) ? 1.0 : -1.0;
    for (int k = 0; k < 3; k++) {
      cl[k] = l[k];
      nl[k] = 0.0;
    }
    cl[axis] = 0.5 * sign;
    nl[axis] = sign;
  }
};

// A Baraff-Witkin plane constraint (src/Geometry/CONSTRAINTS.h PLANE_CONSTRAINT)
struct Plane {
  int obj, vertex;
  double cl[3], nl[3];  // closest point and normal, in the object's local frame
  bool sep;
};

struct Sim {
  int nv = 0, nt = 0, dofs = 0;
  std::vector<double> rest;  // nv*3
  std::vector<int> tets;     // nt*4
  Material mat{0, 1, 1};
  double dt = 1.0 / 60.0, ra = 0.01, rb = 0.01;
  double gravity[3] = {0, 0, 0};
  std::vector<double> vol, oneRing, DmInv;  // DmInv nt*9 row-major
  std::vector<double> M;                    // mass diagonal per DOF
  std::vector<char> pinnedDof;              // per DOF
  std::vector<int> pinned;
  std::vector<double> pinTargets;           // npinned*3
  double tol = 1e-8;
  int maxIters = 0;
  int precond = 0;
  int newtonIters = 1;
  double newtonTol = 1e-7;
  bool ground = false;
  double groundY = 0, groundK = 0;
  // energy-based damping (GREEN_DAMPING) in place of Rayleigh damping, and the mesh's held
  // velocity gradients (nt*9 row-major; zero until computeFdots is called)
  bool green = false;
  double greenMu = 0.0;
  std::vector<double> Fdot;
  // BACKWARD_EULER_VELOCITY in place of BACKWARD_EULER_POSITION (single Newton iteration)
  bool velocity = false;

  // BSR pattern, block rows by vertex, columns ascending
  std::vector<int> rowPtr, colIdx;
  // contributions per block: (tet, x, y) in ascending tet order
  std::vector<int> contribPtr;
  std::vector<int> contrib;  // tet*16 + x*4 + y
  std::vector<int> diagBlock;  // per vertex, index of its diagonal block
  // per-vertex force gather list: (tet*4 + local) ascending tet
  std::vector<int> vtxPtr, vtxList;

  std::vector<double> Krest;  // nnzb*9
  std::vector<double> u, v;   // state

  // kinematic collision objects and their plane constraints (TIMESTEPPER.h:178-187)
  std::vector<Cube> cubes;
  std::vector<int> surface;   // surface vertices, ascending (TET_MESH.cpp:567-593)
  std::vector<char> inColl;   // per vertex
  std::vector<Plane> planes;  // in creation order
  void computeSurface();
  // per-vertex constraint filter S: 0 free (I), 1 plane (I - n n^T), 2 pinned (0)
  void buildFilter(std::vector<char>& mode, std::vector<double>& nrm) const;

  // last system (for tests)
  std::vector<double> lastRhs, lastLhs;

  void setup();
  // damped: add the damping PK1 to the forces; CT (may be NULL) gets the element damping
  // stiffnesses.
  void evalElements(const std::vector<double>& uu, std::vector<double>& fT,
                    std::vector<double>* HT, std::vector<double>* CT = nullptr,
                    bool damped = false) const;
  void computeFdots(const double* vel);
  void assembleK(const std::vector<double>& HT, std::vector<double>& K) const;
  void gatherForces(const std::vector<double>& fT, std::vector<double>& R) const;
  void spmv(const std::vector<double>& vals, const double* x, double* y) const;
  int pcg(const std::vector<double>& lhs, const std::vector<double>& rhs, std::vector<double>& x,
          double& err) const;
  int step(double* err, int* itersPerNewton, int* newtonDone);
};

// Surface vertices: vertices of tet faces that no other tet shares (TET_MESH.cpp:294-354,
// 567-593; faces (0,1,3), (0,2,1), (0,3,2), (1,2,3) of each tet).
void Sim::computeSurface() {
  std::map<std::array<int, 3>, int> count;
  const int F[4][3] = {{0, 1, 3}, {0, 2, 1}, {0, 3, 2}, {1, 2, 3}};
  for (int t = 0; t < nt; t++)
    for (int f = 0; f < 4; f++) {
      std::array<int, 3> k = {tets[4 * t + F[f][0]], tets[4 * t + F[f][1]], tets[4 * t + F[f][2]]};
      std::sort(k.begin(), k.end());
      count[k]++;
    }
  std::vector<char> on(nv, 0);
  for (auto& kv : count)
    if (kv.second < 2)
      for (int c = 0; c < 3; c++) on[kv.first[c]] = 1;
  surface.clear();
  for (int i = 0; i < nv; i++)
    if (on[i]) surface.push_back(i);
  inColl.assign(nv, 0);
}

// TIMESTEPPER::buildConstraintMatrix (TIMESTEPPER.cpp:128-165): I - n n^T per live plane
// constraint (n normalized), then the kinematic (pinned) vertices set to zero, last.
void Sim::buildFilter(std::vector<char>& mode, std::vector<double>& nrm) const {
  mode.assign(nv, 0);
  nrm.assign(3 * (size_t)nv, 0.0);
  for (const Plane& p : planes) {
    if (p.sep) continue;
    double n[3];
    cubes[p.obj].normalToWorld(p.nl, n);
    const double len = std::sqrt((n[0] * n[0] + n[1] * n[1]) + n[2] * n[2]);
    for (int c = 0; c < 3; c++) nrm[3 * (size_t)p.vertex + c] = n[c] / len;
    mode[p.vertex] = 1;
  }
  for (int k : pinned) mode[k] = 2;
}

void Sim::setup() {
  dofs = 3 * nv;
  vol.assign(nt, 0.0);
  DmInv.assign(9 * (size_t)nt, 0.0);
  for (int t = 0; t < nt; t++) {  // TET_MESH.cpp:121-151, 178-191
    const int* T = &tets[4 * t];
    double d[3][3];
    for (int k = 0; k < 3; k++)
      for (int c = 0; c < 3; c++) d[k][c] = rest[3 * T[k + 1] + c] - rest[3 * T[0] + c];
    const double cx = d[0][1] * d[1][2] - d[0][2] * d[1][1];
    const double cy = d[0][2] * d[1][0] - d[0][0] * d[1][2];
    const double cz = d[0][0] * d[1][1] - d[0][1] * d[1][0];
    vol[t] = (d[2][0] * cx + d[2][1] * cy + d[2][2] * cz) / 6.0;
    M3 Dm, Di;
    for (int r = 0; r < 3; r++)
      for (int c = 0; c < 3; c++) Dm[r][c] = d[c][r];
    inv3(Dm, Di);
    for (int r = 0; r < 3; r++)
      for (int c = 0; c < 3; c++) DmInv[9 * (size_t)t + 3 * r + c] = Di[r][c];
  }
  oneRing.assign(nv, 0.0);  // TET_MESH.cpp:156-173
  for (int t = 0; t < nt; t++) {
    const double q = 0.25 * vol[t];
    for (int k = 0; k < 4; k++) oneRing[tets[4 * t + k]] += q;
  }
  M.assign(dofs, 0.0);
  // BSR pattern
  std::vector<std::vector<int>> nbr(nv);
  for (int t = 0; t < nt; t++)
    for (int a = 0; a < 4; a++)
      for (int b = 0; b < 4; b++) nbr[tets[4 * t + a]].push_back(tets[4 * t + b]);
  rowPtr.assign(nv + 1, 0);
  colIdx.clear();
  for (int i = 0; i < nv; i++) {
    auto& n = nbr[i];
    std::sort(n.begin(), n.end());
    n.erase(std::unique(n.begin(), n.end()), n.end());
    rowPtr[i + 1] = rowPtr[i] + (int)n.size();
    colIdx.insert(colIdx.end(), n.begin(), n.end());
  }
  const int nnzb = rowPtr[nv];
  diagBlock.assign(nv, -1);
  for (int i = 0; i < nv; i++)
    for (int k = rowPtr[i]; k < rowPtr[i + 1]; k++)
      if (colIdx[k] == i) diagBlock[i] = k;
  auto findBlock = [&](int i, int j) {
    const int* b = colIdx.data() + rowPtr[i];
    const int* e = colIdx.data() + rowPtr[i + 1];
    return (int)(std::lower_bound(b, e, j) - colIdx.data());
  };
  std::vector<int> cnt(nnzb + 1, 0);
  for (int t = 0; t < nt; t++)
    for (int x = 0; x < 4; x++)
      for (int y = 0; y < 4; y++) cnt[findBlock(tets[4 * t + x], tets[4 * t + y]) + 1]++;
  contribPtr.assign(nnzb + 1, 0);
  for (int k = 0; k < nnzb; k++) contribPtr[k + 1] = contribPtr[k] + cnt[k + 1];
  contrib.assign(contribPtr[nnzb], 0);
  std::vector<int> fill(contribPtr.begin(), contribPtr.end() - 1);
  for (int t = 0; t < nt; t++)  // ascending tet order, as TET_MESH_FASTER.cpp:213-236 pushes
    for (int y = 0; y < 4; y++)
      for (int x = 0; x < 4; x++) {
        const int k = findBlock(tets[4 * t + x], tets[4 * t + y]);
        contrib[fill[k]++] = t * 16 + x * 4 + y;
      }
  std::vector<int> vc(nv + 1, 0);
  for (int t = 0; t < nt; t++)
    for (int a = 0; a < 4; a++) vc[tets[4 * t + a] + 1]++;
  vtxPtr.assign(nv + 1, 0);
  for (int i = 0; i < nv; i++) vtxPtr[i + 1] = vtxPtr[i] + vc[i + 1];
  vtxList.assign(vtxPtr[nv], 0);
  std::vector<int> vf(vtxPtr.begin(), vtxPtr.end() - 1);
  for (int t = 0; t < nt; t++)
    for (int a = 0; a < 4; a++) vtxList[vf[tets[4 * t + a]]++] = 4 * t + a;

  u.assign(dofs, 0.0);
  v.assign(dofs, 0.0);
  Fdot.assign(9 * (size_t)nt, 0.0);
  // Rayleigh stiffness at the rest state (TIMESTEPPER.cpp:213-236), constant across steps.
  std::vector<double> fT, HT;
  std::vector<double> zero(dofs, 0.0);
  evalElements(zero, fT, &HT);
  assembleK(HT, Krest);
}

void Sim::evalElements(const std::vector<double>& uu, std::vector<double>& fT,
                       std::vector<double>* HT, std::vector<double>* CT, bool damped) const {
  fT.assign(12 * (size_t)nt, 0.0);
  if (HT) HT->assign(144 * (size_t)nt, 0.0);
  if (CT) CT->assign(144 * (size_t)nt, 0.0);
#pragma omp parallel for schedule(static)
  for (int t = 0; t < nt; t++) {
    const int* T = &tets[4 * t];
    double x[4][3];
    for (int k = 0; k < 4; k++)
      for (int c = 0; c < 3; c++) x[k][c] = rest[3 * T[k] + c] + uu[3 * T[k] + c];  // setDisplacement
    M3 Ds, Di, F;
    for (int r = 0; r < 3; r++)
      for (int c = 0; c < 3; c++) Ds[r][c] = x[c + 1][r] - x[0][r];
    for (int r = 0; r < 3; r++)
      for (int c = 0; c < 3; c++) Di[r][c] = DmInv[9 * (size_t)t + 3 * r + c];
    m3_mul(Ds, Di, F);  // TET_MESH.cpp:706-714
    M3 U, V;
    double S[3];
    svd_rv(F, U, S, V);  // TET_MESH.cpp:689-701
    double P9x12[9][12];
    pFpx(Di, P9x12);
    double H[12][12];
    M3 Pd, Fd;
    if (damped) {  // computeInternalForces(hyperelastic, damping): P = elastic + damping PK1
      for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) Fd[r][c] = Fdot[9 * (size_t)t + 3 * r + c];
      green_pk1(greenMu, F, Fd, Pd);
    }
    element(mat, F, U, S, V, P9x12, vol[t], &fT[12 * (size_t)t], HT ? H : nullptr,
            damped ? Pd : nullptr);
    if (HT)
      for (int i = 0; i < 12; i++)
        for (int j = 0; j < 12; j++) (*HT)[144 * (size_t)t + 12 * i + j] = H[i][j];
    if (CT) {
      element_damping(greenMu, F, P9x12, vol[t], H);
      for (int i = 0; i < 12; i++)
        for (int j = 0; j < 12; j++) (*CT)[144 * (size_t)t + 12 * i + j] = H[i][j];
    }
  }
}

void Sim::computeFdots(const double* vel) {
  Fdot.assign(9 * (size_t)nt, 0.0);
  for (int t = 0; t < nt; t++) {
    M3 Di, Fd;
    for (int r = 0; r < 3; r++)
      for (int c = 0; c < 3; c++) Di[r][c] = DmInv[9 * (size_t)t + 3 * r + c];
    tet_fdot(&tets[4 * t], vel, Di, Fd);
    for (int r = 0; r < 3; r++)
      for (int c = 0; c < 3; c++) Fdot[9 * (size_t)t + 3 * r + c] = Fd[r][c];
  }
}

// Gather assembly: each block sums its contributions in ascending tet order
// (TET_MESH_FASTER.cpp:113-131).  Block values are row-major 3x3: entry (a, b) at 3a + b.
void Sim::assembleK(const std::vector<double>& HT, std::vector<double>& K) const {
  const int nnzb = rowPtr[nv];
  K.assign(9 * (size_t)nnzb, 0.0);
#pragma omp parallel for schedule(static)
  for (int k = 0; k < nnzb; k++) {
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) {
        double s = 0;
        for (int c = contribPtr[k]; c < contribPtr[k + 1]; c++) {
          const int code = contrib[c];
          const int t = code >> 4, x = (code >> 2) & 3, y = code & 3;
          s += HT[144 * (size_t)t + 12 * (3 * x + a) + (3 * y + b)];
        }
        K[9 * (size_t)k + 3 * a + b] = s;
      }
  }
}

// Force scatter in ascending tet order (TET_MESH.cpp:756-767), done as a per-vertex gather.
void Sim::gatherForces(const std::vector<double>& fT, std::vector<double>& R) const {
  R.assign(dofs, 0.0);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nv; i++)
    for (int c = 0; c < 3; c++) {
      double s = 0;
      for (int k = vtxPtr[i]; k < vtxPtr[i + 1]; k++) {
        const int code = vtxList[k];
        s += fT[12 * (size_t)(code >> 2) + 3 * (code & 3) + c];
      }
      R[3 * i + c] = s;
    }
}

void Sim::spmv(const std::vector<double>& vals, const double* x, double* y) const {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nv; i++) {
    double s0 = 0, s1 = 0, s2 = 0;
    for (int k = rowPtr[i]; k < rowPtr[i + 1]; k++) {
      const double* B = &vals[9 * (size_t)k];
      const double* xj = x + 3 * colIdx[k];
      s0 += B[0] * xj[0] + B[1] * xj[1] + B[2] * xj[2];
      s1 += B[3] * xj[0] + B[4] * xj[1] + B[5] * xj[2];
      s2 += B[6] * xj[0] + B[7] * xj[1] + B[8] * xj[2];
    }
    y[3 * i] = s0;
    y[3 * i + 1] = s1;
    y[3 * i + 2] = s2;
  }
}

// Conjugate gradients with Eigen 3.3.7's stopping rule and iteration count
// (ext/eigen-3.3.7/Eigen/src/IterativeLinearSolvers/ConjugateGradient.h:28-91):
// x0 = 0; threshold = max(tol^2 |b|^2, DBL_MIN) on the recursively updated residual; the
// converging pass is not counted.  Preconditioner: Jacobi as DiagonalPreconditioner
// (BasicPreconditioners.h:65-79: 1/a_ii, or 1 where the diagonal is zero) or 3x3 block Jacobi.
int Sim::pcg(const std::vector<double>& lhs, const std::vector<double>& b, std::vector<double>& x,
             double& err) const {
  const int n = dofs;
  x.assign(n, 0.0);
  const int maxIt = maxIters > 0 ? maxIters : 2 * n;
  std::vector<double> Minv(precond == 0 ? n : 9 * (size_t)nv);
  if (precond == 0) {
    for (int i = 0; i < nv; i++)
      for (int c = 0; c < 3; c++) {
        const double d = diagBlock[i] >= 0 ? lhs[9 * (size_t)diagBlock[i] + 4 * c] : 0.0;
        Minv[3 * i + c] = (d != 0.0) ? 1.0 / d : 1.0;
      }
  } else {
    for (int i = 0; i < nv; i++) {
      M3 B, Bi;
      for (int a = 0; a < 3; a++)
        for (int c = 0; c < 3; c++) B[a][c] = lhs[9 * (size_t)diagBlock[i] + 3 * a + c];
      inv3(B, Bi);
      for (int a = 0; a < 3; a++)
        for (int c = 0; c < 3; c++) Minv[9 * (size_t)i + 3 * a + c] = Bi[a][c];
    }
  }
  auto apply = [&](const std::vector<double>& r, std::vector<double>& z) {
    if (precond == 0) {
#pragma omp parallel for schedule(static)
      for (int i = 0; i < n; i++) z[i] = Minv[i] * r[i];
    } else {
#pragma omp parallel for schedule(static)
      for (int i = 0; i < nv; i++) {
        const double* B = &Minv[9 * (size_t)i];
        const double r0 = r[3 * i], r1 = r[3 * i + 1], r2 = r[3 * i + 2];
        z[3 * i] = B[0] * r0 + B[1] * r1 + B[2] * r2;
        z[3 * i + 1] = B[3] * r0 + B[4] * r1 + B[5] * r2;
        z[3 * i + 2] = B[6] * r0 + B[7] * r1 + B[8] * r2;
      }
    }
  };
  std::vector<double> r(b), p(n), z(n), t(n);
  const double rhsNorm2 = det_sum(n, [&](int i) { return b[i] * b[i]; });
  if (rhsNorm2 == 0) {
    err = 0;
    return 0;
  }
  const double threshold = std::max(tol * tol * rhsNorm2, DBL_MIN);
  double resNorm2 = rhsNorm2;
  if (resNorm2 < threshold) {
    err = std::sqrt(resNorm2 / rhsNorm2);
    return 0;
  }
  apply(r, p);
  double absNew = det_sum(n, [&](int i) { return r[i] * p[i]; });
  int it = 0;
  while (it < maxIt) {
    spmv(lhs, p.data(), t.data());
    const double alpha = absNew / det_sum(n, [&](int i) { return p[i] * t[i]; });
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
      x[i] += alpha * p[i];
      r[i] -= alpha * t[i];
    }
    resNorm2 = det_sum(n, [&](int i) { return r[i] * r[i]; });
    if (resNorm2 < threshold) break;
    apply(r, z);
    const double absOld = absNew;
    absNew = det_sum(n, [&](int i) { return r[i] * z[i]; });
    const double beta = absNew / absOld;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
    it++;
  }
  err = std::sqrt(resNorm2 / rhsNorm2);
  return it;
}

// The 3x3 block of the constraint filter S at one vertex: I, I - n n^T (TIMESTEPPER.cpp:149) or 0.
static void filterBlock(int mode, const double* n, double S[3][3]) {
  for (int a = 0; a < 3; a++)
    for (int c = 0; c < 3; c++)
      S[a][c] = mode == 2 ? 0.0 : mode == 1 ? (a == c ? 1.0 : 0.0) - n[a] * n[c] : (a == c ? 1.0 : 0.0);
}

// One stepSimulation (SIMULATION_SCENE.h:120-126): f_ext = one-ring volume * gravity, then
//   newtonIters == 1: BACKWARD_EULER_POSITION::solveRayleighDamped (BACKWARD_EULER_POSITION.cpp:138-327),
//                     or solveEnergyDamped (:332-473) with GREEN_DAMPING: R gains the damping
//                     force at the held Fdot (the step never updates it), b = (M/dt) v + R +
//                     f_ext (:384) and A = M/dt^2 - K.  The step does assemble the damping
//                     stiffness C (:378), but computeCollisionResponse(R, K, C) (:381) takes C as
//                     its collisionC argument and empties it (TIMESTEPPER.cpp:657) before
//                     A = M/dt^2 - C/dt - K (:389), so with self-collisions off C is zero there.
//   newtonIters  > 1: BDF_1::solveRayleighDamped (BDF_1.cpp:78-239) with residual test maxR > tol
//   velocity:         BACKWARD_EULER_VELOCITY::solveRayleighDamped (BACKWARD_EULER_VELOCITY.cpp:
//                     107-235), which solves for the velocity change dv:
//                       b = dt (R + dt K v + f_ext) (:157), A = M - dt C - dt^2 K (:161), where
//                       C = alpha M + beta K_rest (:123) is emptied by computeCollisionResponse
//                       (:153; TIMESTEPPER.cpp:657), so A = M - dt^2 K;
//                       plane-constraint targets are velocity changes (cl - x)/dt - v (:51-78);
//                       dv = y + z, v += dv, u += dt v, v = S v (:190-201), then the same
//                       separation test and constraint bookkeeping as the position form.
int Sim::step(double* errOut, int* itersPerNewton, int* newtonDone) {
  const int nnzb = rowPtr[nv];
  const double invDt = 1.0 / dt;
  const double invDt2 = invDt * invDt;
  std::vector<double> fext(dofs);
  for (int i = 0; i < nv; i++)
    for (int c = 0; c < 3; c++) fext[3 * i + c] = oneRing[i] * gravity[c];
  // C = alpha M + beta K_rest (TIMESTEPPER.cpp:232-233); D = M/dt^2 - C/dt; E = M/dt - C.
  // Energy damping has no Rayleigh part: D = M/dt^2, E = M/dt (BACKWARD_EULER_POSITION.cpp:384, 389).
  // The velocity form's C is emptied before A is formed: A = M - dt^2 K (D = M here).
  const double ra_ = (green || velocity) ? 0.0 : ra, rb_ = (green || velocity) ? 0.0 : rb;
  std::vector<double> Cv(9 * (size_t)nnzb), D(9 * (size_t)nnzb), E(9 * (size_t)nnzb);
  for (int i = 0; i < nv; i++)
    for (int k = rowPtr[i]; k < rowPtr[i + 1]; k++)
      for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
          const size_t e = 9 * (size_t)k + 3 * a + b;
          const bool dg = (colIdx[k] == i) && (a == b);
          const double m = dg ? M[3 * i + a] : 0.0;
          const double c = dg ? ra_ * m + rb_ * Krest[e] : rb_ * Krest[e];
          Cv[e] = c;
          D[e] = velocity ? m : dg ? m * invDt2 - c * invDt : -(c * invDt);
          E[e] = dg ? invDt * m - c : -c;
        }
  const std::vector<double> uOld(u), vOld(v);
  for (size_t k = 0; k < pinned.size(); k++)  // applyKinematicConstraints (TIMESTEPPER.cpp:105-122)
    for (int c = 0; c < 3; c++) u[3 * pinned[k] + c] = pinTargets[3 * k + c];

  // Plane constraints against kinematic collision objects (BACKWARD_EULER_POSITION.cpp:163-181).
  // xk: the positions the constraint logic sees during this step (the mesh's vertices).
  const bool contact = !cubes.empty() && newtonIters == 1;
  std::vector<char> mode;
  std::vector<double> nrm, zc, xk;
  if (contact) {
    xk.resize(dofs);
    for (int i = 0; i < dofs; i++) xk[i] = rest[i] + u[i];
    // findNewSurfaceConstraints (TIMESTEPPER.cpp:446-526)
    for (int o = 0; o < (int)cubes.size(); o++)
      for (int sv : surface) {
        if (inColl[sv]) continue;
        const double* x = &xk[3 * (size_t)sv];
        if (!cubes[o].inside(x)) continue;
        Plane p;
        p.obj = o;
        p.vertex = sv;
        p.sep = false;
        cubes[o].closest(x, p.cl, p.nl);
        double n[3];
        cubes[o].normalToWorld(p.nl, n);
        const double* vv = &v[3 * (size_t)sv];
        if ((vv[0] * n[0] + vv[1] * n[1]) + vv[2] * n[2] >= -FLT_EPSILON) continue;
        planes.push_back(p);
        inColl[sv] = 1;
      }
    buildFilter(mode, nrm);
    // updateConstraintTargets (BACKWARD_EULER_POSITION.cpp:85-110; the velocity form's
    // BACKWARD_EULER_VELOCITY.cpp:51-78 targets (cl - x)/dt - v), then z = (I - S) targets
    std::vector<double> tg(dofs, 0.0);
    for (const Plane& p : planes) {
      if (p.sep) continue;
      double cw[3];
      cubes[p.obj].toWorld(p.cl, cw);
      for (int c = 0; c < 3; c++) {
        const size_t i = 3 * (size_t)p.vertex + c;
        tg[i] = velocity ? (cw[c] - xk[i]) / dt - v[i] : cw[c] - xk[i];
      }
    }
    zc.assign(dofs, 0.0);
    for (int i = 0; i < nv; i++) {
      double S[3][3], ImS[3][3];
      filterBlock(mode[i], &nrm[3 * (size_t)i], S);
      for (int a = 0; a < 3; a++)
        for (int c = 0; c < 3; c++) ImS[a][c] = (a == c ? 1.0 : 0.0) - S[a][c];
      for (int a = 0; a < 3; a++)
        zc[3 * (size_t)i + a] = (ImS[a][0] * tg[3 * (size_t)i] + ImS[a][1] * tg[3 * (size_t)i + 1]) +
                                ImS[a][2] * tg[3 * (size_t)i + 2];
    }
  }

  int total = 0, done = 0;
  double err = 0;
  double maxR = 10 * newtonTol;
  std::vector<double> fT, HT, K, R, b(dofs), rhs(dofs), lhs(9 * (size_t)nnzb), y, tmp(dofs);
  while (done < newtonIters && (newtonIters == 1 || maxR > newtonTol)) {
    evalElements(u, fT, &HT, nullptr, green);
    assembleK(HT, K);
    const std::vector<double>& Dq = D;
    gatherForces(fT, R);
    if (ground) {  // penalty floor: psi = k/2 (y_g - y)^2 below the plane (no reference counterpart)
      for (int i = 0; i < nv; i++) {
        const double y = rest[3 * i + 1] + u[3 * i + 1];
        if (y < groundY) {
          R[3 * i + 1] += groundK * (groundY - y);
          K[9 * (size_t)diagBlock[i] + 4] -= groundK;
        }
      }
    }
    if (velocity) {
      // b = dt (R + dt K v + f_ext)   (BACKWARD_EULER_VELOCITY.cpp:157)
      spmv(K, v.data(), tmp.data());
      for (int i = 0; i < dofs; i++) b[i] = dt * ((R[i] + dt * tmp[i]) + fext[i]);
    } else if (newtonIters == 1) {
      // b = (M/dt - C) v + R + f_ext   (BACKWARD_EULER_POSITION.cpp:198)
      spmv(E, v.data(), tmp.data());
      for (int i = 0; i < dofs; i++) b[i] = (tmp[i] + R[i]) + fext[i];
    } else {
      // BDF_1.cpp:143-150: a = -(u-u0)/dt^2 + v0/dt; vb = (u-u0)/dt - v0; b = C vb + (M a + R + f)
      std::vector<double> vb(dofs), Ma(dofs);
      for (int i = 0; i < dofs; i++) {
        const double a = -invDt2 * (u[i] - uOld[i]) + invDt * vOld[i];
        Ma[i] = M[i] * a;
        vb[i] = invDt * (u[i] - uOld[i]) - vOld[i];
      }
      spmv(Cv, vb.data(), tmp.data());
      for (int i = 0; i < dofs; i++) b[i] = tmp[i] + (((Ma[i] + R[i]) + fext[i]));
    }
    // A = M/dt^2 - C/dt - K (velocity form: M - dt^2 K); rhs = S b; LHS = S A S + (I - S)
    // (:208, :242-243)
    const double kdt = velocity ? dt * dt : 1.0;
    if (contact) {
      // c = b - A z (:212); rhs = S c; LHS = S A S + (I - S), each product summed in Eigen's
      // column order
      std::vector<double> Av(9 * (size_t)nnzb), Az(dofs);
      for (size_t q = 0; q < Av.size(); q++) Av[q] = velocity ? Dq[q] - kdt * K[q] : Dq[q] - K[q];
      spmv(Av, zc.data(), Az.data());
      std::vector<double> Sb(9 * (size_t)nv);
      for (int i = 0; i < nv; i++) {
        double S[3][3];
        filterBlock(mode[i], &nrm[3 * (size_t)i], S);
        for (int a = 0; a < 3; a++)
          for (int c = 0; c < 3; c++) Sb[9 * (size_t)i + 3 * a + c] = S[a][c];
      }
      for (int i = 0; i < nv; i++) {
        const double* Si = &Sb[9 * (size_t)i];
        double cv[3];
        for (int a = 0; a < 3; a++) cv[a] = b[3 * i + a] - Az[3 * i + a];
        for (int a = 0; a < 3; a++) rhs[3 * i + a] = (Si[3 * a] * cv[0] + Si[3 * a + 1] * cv[1]) + Si[3 * a + 2] * cv[2];
        for (int k = rowPtr[i]; k < rowPtr[i + 1]; k++) {
          const int j = colIdx[k];
          const double* Sj = &Sb[9 * (size_t)j];
          const double* A = &Av[9 * (size_t)k];
          double SA[3][3];
          for (int a = 0; a < 3; a++)
            for (int d = 0; d < 3; d++) SA[a][d] = (Si[3 * a] * A[d] + Si[3 * a + 1] * A[3 + d]) + Si[3 * a + 2] * A[6 + d];
          for (int a = 0; a < 3; a++)
            for (int c = 0; c < 3; c++) {
              double x = (SA[a][0] * Sj[c] + SA[a][1] * Sj[3 + c]) + SA[a][2] * Sj[6 + c];
              if (j == i) x += (a == c ? 1.0 : 0.0) - Si[3 * a + c];
              lhs[9 * (size_t)k + 3 * a + c] = x;
            }
        }
      }
    } else
    for (int i = 0; i < nv; i++) {
      const bool pi = pinnedDof[3 * i];
      for (int k = rowPtr[i]; k < rowPtr[i + 1]; k++) {
        const int j = colIdx[k];
        const bool pj = pinnedDof[3 * j];
        for (int e = 0; e < 9; e++) {
          const size_t q = 9 * (size_t)k + e;
          double a = velocity ? Dq[q] - kdt * K[q] : Dq[q] - K[q];
          if (pi || pj) a = (j == i && (e % 4 == 0)) ? 1.0 : 0.0;
          lhs[q] = a;
        }
      }
      for (int c = 0; c < 3; c++) rhs[3 * i + c] = pi ? 0.0 : b[3 * i + c];
    }
    maxR = 0;
    for (int i = 0; i < dofs; i++) maxR += rhs[i] * rhs[i];
    const int it = pcg(lhs, rhs, y, err);
    if (itersPerNewton) itersPerNewton[done] = it;
    total += it;
    if (velocity) {
      // vDelta = y + z; v += vDelta; u += dt v; v = S v (BACKWARD_EULER_VELOCITY.cpp:190-201)
      for (int i = 0; i < dofs; i++) {
        v[i] = v[i] + (contact ? y[i] + zc[i] : y[i]);
        u[i] = u[i] + dt * v[i];
      }
      if (contact) {
        for (int i = 0; i < nv; i++) {
          double S[3][3], w[3];
          filterBlock(mode[i], &nrm[3 * (size_t)i], S);
          for (int c = 0; c < 3; c++) w[c] = v[3 * i + c];
          for (int a = 0; a < 3; a++) v[3 * i + a] = (S[a][0] * w[0] + S[a][1] * w[1]) + S[a][2] * w[2];
        }
      } else {
        for (int i = 0; i < dofs; i++)
          if (pinnedDof[i]) v[i] = 0.0;
      }
    } else if (contact) {  // xDelta = y + z (:252)
      for (int i = 0; i < dofs; i++) u[i] += y[i] + zc[i];
    } else {
      for (int i = 0; i < dofs; i++) u[i] += y[i];
    }
    lastRhs = rhs;
    lastLhs = lhs;
    done++;
  }
  if (contact) {
    // findSeparatingSurfaceConstraints(b) (TIMESTEPPER.cpp:252-329) at this step's positions
    bool changed = false;
    for (Plane& p : planes) {
      const double* x = &xk[3 * (size_t)p.vertex];
      if (cubes[p.obj].signedDistance(x) > 1e-6) {
        p.sep = true;
        changed = true;
        continue;
      }
      double xd[3] = {b[3 * p.vertex], b[3 * p.vertex + 1], b[3 * p.vertex + 2]};
      const double len = std::sqrt((xd[0] * xd[0] + xd[1] * xd[1]) + xd[2] * xd[2]);
      if (len > 1.0)
        for (int c = 0; c < 3; c++) xd[c] /= len;
      double n[3];
      cubes[p.obj].normalToWorld(p.nl, n);
      if ((xd[0] * n[0] + xd[1] * n[1]) + xd[2] * n[2] > 1e-6) {
        p.sep = true;
        changed = true;
      }
    }
    if (changed) {  // deleteSurfaceConstraints (TIMESTEPPER.cpp:561-586)
      std::vector<Plane> keep;
      for (const Plane& p : planes) {
        if (p.sep)
          inColl[p.vertex] = 0;
        else
          keep.push_back(p);
      }
      planes.swap(keep);
    }
    // updateSurfaceConstraints (TIMESTEPPER.cpp:531-553); S is rebuilt only when a constraint
    // went away, so otherwise the velocity filter keeps this step's normals (:283-296)
    for (Plane& p : planes) cubes[p.obj].closest(&xk[3 * (size_t)p.vertex], p.cl, p.nl);
    if (changed) buildFilter(mode, nrm);
    // the velocity form filtered v with this step's S before the test (above)
    for (int i = 0; i < nv && !velocity; i++) {
      double S[3][3], w[3];
      filterBlock(mode[i], &nrm[3 * (size_t)i], S);
      for (int c = 0; c < 3; c++) w[c] = invDt * (u[3 * i + c] - uOld[3 * i + c]);
      for (int a = 0; a < 3; a++) v[3 * i + a] = (S[a][0] * w[0] + S[a][1] * w[1]) + S[a][2] * w[2];
    }
  } else if (!velocity) {
    // v = (u - u_old)/dt, then filtered by S (BACKWARD_EULER_POSITION.cpp:300-314)
    for (int i = 0; i < dofs; i++) v[i] = pinnedDof[i] ? 0.0 : invDt * (u[i] - uOld[i]);
  }
  if (errOut) *errOut = err;
  if (newtonDone) *newtonDone = done;
  return total;
}

}  // namespace

extern "C" {

void orc_set_threads(int n) { omp_set_num_threads(n); }

void orc_svd_rv(const double* F, double* U, double* S, double* V) {
  M3 f, u, v;
  m3_from_colmajor(F, f);
  svd_rv(f, u, S, v);
  m3_to_colmajor(u, U);
  m3_to_colmajor(v, V);
}

double orc_psi(int material, double mu, double lambda, const double* F) {
  M3 f;
  m3_from_colmajor(F, f);
  return psi(Material(material, mu, lambda), f);
}

void orc_pk1(int material, double mu, double lambda, const double* F, double* P) {
  M3 f, p;
  m3_from_colmajor(F, f);
  pk1(Material(material, mu, lambda), f, p);
  m3_to_colmajor(p, P);
}

void orc_hessian(int material, double mu, double lambda, const double* F, double* H9) {
  M3 f;
  m3_from_colmajor(F, f);
  double H[9][9];
  hessian(Material(material, mu, lambda), f, H);
  for (int j = 0; j < 9; j++)
    for (int i = 0; i < 9; i++) H9[9 * j + i] = H[i][j];
}

void orc_clamped_hessian_usv(int material, double mu, double lambda, const double* Uc,
                             const double* S, const double* Vc, double* H9) {
  M3 u, v;
  m3_from_colmajor(Uc, u);
  m3_from_colmajor(Vc, v);
  double H[9][9];
  clamped_hessian(Material(material, mu, lambda), u, S, v, H);
  for (int j = 0; j < 9; j++)
    for (int i = 0; i < 9; i++) H9[9 * j + i] = H[i][j];
}

void orc_clamped_hessian(int material, double mu, double lambda, const double* F, double* H9) {
  M3 f, u, v;
  double s[3];
  m3_from_colmajor(F, f);
  svd_rv(f, u, s, v);
  double H[9][9];
  clamped_hessian(Material(material, mu, lambda), u, s, v, H);
  for (int j = 0; j < 9; j++)
    for (int i = 0; i < 9; i++) H9[9 * j + i] = H[i][j];
}

double orc_green_psi(double mu, const double* F, const double* Fdot) {
  M3 f, fd;
  m3_from_colmajor(F, f);
  m3_from_colmajor(Fdot, fd);
  return green_psi(mu, f, fd);
}

void orc_green_pk1(double mu, const double* F, const double* Fdot, double* P) {
  M3 f, fd, p;
  m3_from_colmajor(F, f);
  m3_from_colmajor(Fdot, fd);
  green_pk1(mu, f, fd, p);
  m3_to_colmajor(p, P);
}

void orc_green_hessian(double mu, const double* F, int clamped, double* H9) {
  M3 f;
  m3_from_colmajor(F, f);
  double H[9][9];
  if (clamped)
    green_clamped_hessian(mu, f, H);
  else
    green_hessian(mu, f, H);
  for (int j = 0; j < 9; j++)
    for (int i = 0; i < 9; i++) H9[9 * j + i] = H[i][j];
}

void orc_tet_force_hessian(int material, double mu, double lambda, const double* x,
                           const double* DmInvC, double restVolume, double* f12, double* H144) {
  Material m(material, mu, lambda);
  M3 Ds, Di, F, U, V;
  double S[3];
  m3_from_colmajor(DmInvC, Di);
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++) Ds[r][c] = x[3 * (c + 1) + r] - x[r];
  m3_mul(Ds, Di, F);
  svd_rv(F, U, S, V);
  double P9x12[9][12], H[12][12];
  pFpx(Di, P9x12);
  element(m, F, U, S, V, P9x12, restVolume, f12, H);
  for (int j = 0; j < 12; j++)
    for (int i = 0; i < 12; i++) H144[12 * j + i] = H[i][j];
}

void* orc_create(const double* rest_xyz, int nv, const int* tets, int nt, int material,
                 double mu, double lambda, double dt, double rayleigh_alpha, double rayleigh_beta,
                 double mass_scale, const double* gravity, const int* pinned, int npinned,
                 const double* pin_targets, double pcg_tol, int pcg_max_iters, int precond,
                 int newton_iters, double newton_tol) {
  Sim* s = new Sim;
  s->nv = nv;
  s->nt = nt;
  s->rest.assign(rest_xyz, rest_xyz + 3 * (size_t)nv);
  s->tets.assign(tets, tets + 4 * (size_t)nt);
  s->mat = Material(material, mu, lambda);
  s->dt = dt;
  s->ra = rayleigh_alpha;
  s->rb = rayleigh_beta;
  for (int c = 0; c < 3; c++) s->gravity[c] = gravity[c];
  s->tol = pcg_tol;
  s->maxIters = pcg_max_iters;
  s->precond = precond;
  s->newtonIters = newton_iters < 1 ? 1 : newton_iters;
  s->newtonTol = newton_tol;
  s->setup();
  for (int i = 0; i < nv; i++)  // buildMassMatrix then scaleMass (TIMESTEPPER.cpp:176-208)
    for (int c = 0; c < 3; c++) s->M[3 * i + c] = s->oneRing[i] * mass_scale;
  s->pinnedDof.assign(s->dofs, 0);
  s->pinned.assign(pinned, pinned + npinned);
  s->pinTargets.assign(pin_targets, pin_targets + 3 * (size_t)npinned);
  for (int k = 0; k < npinned; k++)
    for (int c = 0; c < 3; c++) s->pinnedDof[3 * pinned[k] + c] = 1;
  return s;
}

void orc_destroy(void* h) { delete (Sim*)h; }

// TIMESTEPPER::addKinematicCollisionObject with a CUBE(center, scale) (TIMESTEPPER.cpp:433-441)
void orc_add_collision_cube(void* h, const double* center, double scale) {
  Sim* s = (Sim*)h;
  if (s->surface.empty()) s->computeSurface();
  s->cubes.emplace_back(center, scale);
}

// Live plane constraints: count, and their vertex ids in creation order.
int orc_plane_constraints(void* h, int* vertex_ids) {
  Sim* s = (Sim*)h;
  if (vertex_ids)
    for (size_t i = 0; i < s->planes.size(); i++) vertex_ids[i] = s->planes[i].vertex;
  return (int)s->planes.size();
}

// Surface vertices (ascending): count, and the ids.
int orc_surface_vertices(void* h, int* ids) {
  Sim* s = (Sim*)h;
  if (s->surface.empty()) s->computeSurface();
  if (ids) std::copy(s->surface.begin(), s->surface.end(), ids);
  return (int)s->surface.size();
}

void orc_set_ground(void* h, double y, double k) {
  Sim* s = (Sim*)h;
  s->ground = true;
  s->groundY = y;
  s->groundK = k;
}

void orc_set_state(void* h, const double* u, const double* v) {
  Sim* s = (Sim*)h;
  if (u) s->u.assign(u, u + s->dofs);
  if (v) s->v.assign(v, v + s->dofs);
}

void orc_get_state(void* h, double* u, double* v) {
  Sim* s = (Sim*)h;
  if (u) std::copy(s->u.begin(), s->u.end(), u);
  if (v) std::copy(s->v.begin(), s->v.end(), v);
}

int orc_step(void* h, double* pcg_error, double* seconds, int* iters_per_newton, int* newton_done) {
  Sim* s = (Sim*)h;
  auto t0 = std::chrono::steady_clock::now();
  int it = s->step(pcg_error, iters_per_newton, newton_done);
  auto t1 = std::chrono::steady_clock::now();
  if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
  return it;
}

int orc_nnzb(void* h) { return ((Sim*)h)->rowPtr.back(); }

void orc_bsr_pattern(void* h, int* row_ptr, int* col_idx) {
  Sim* s = (Sim*)h;
  std::copy(s->rowPtr.begin(), s->rowPtr.end(), row_ptr);
  std::copy(s->colIdx.begin(), s->colIdx.end(), col_idx);
}

void orc_forces_hessian(void* h, const double* u, double* R, double* Kvals) {
  Sim* s = (Sim*)h;
  std::vector<double> uu(u, u + s->dofs), fT, HT, K, Rv;
  s->evalElements(uu, fT, Kvals ? &HT : nullptr);
  s->gatherForces(fT, Rv);
  std::copy(Rv.begin(), Rv.end(), R);
  if (Kvals) {
    s->assembleK(HT, K);
    std::copy(K.begin(), K.end(), Kvals);
  }
}

// Energy damping with GREEN_DAMPING(mu) in place of Rayleigh damping (solveEnergyDamped,
// BACKWARD_EULER_POSITION.cpp:332-473).  -1 for a multi-iteration handle: the reference's BDF_1
// has no energy-damped solve (BDF_1.cpp:244-255).
int orc_set_green_damping(void* h, double mu) {
  Sim* s = (Sim*)h;
  if (s->newtonIters > 1) return -1;
  s->green = true;
  s->greenMu = mu;
  return 0;
}

void orc_compute_fdots(void* h, const double* vel) { ((Sim*)h)->computeFdots(vel); }

// BACKWARD_EULER_VELOCITY in place of BACKWARD_EULER_POSITION.  Single-iteration Rayleigh handles
// without the penalty ground only (returns -1 otherwise).
int orc_set_velocity_form(void* h) {
  Sim* s = (Sim*)h;
  if (s->newtonIters > 1 || s->green || s->ground) return -1;
  s->velocity = true;
  return 0;
}

// computeInternalForces(hyperelastic, damping) and the assembled damping stiffness C
// (TET_MESH_FASTER::computeDampingHessian) at displacement u with the held Fdot.
void orc_damped_forces_hessian(void* h, const double* u, double* R, double* Cvals) {
  Sim* s = (Sim*)h;
  std::vector<double> uu(u, u + s->dofs), fT, CT, C, Rv;
  s->evalElements(uu, fT, nullptr, Cvals ? &CT : nullptr, true);
  s->gatherForces(fT, Rv);
  std::copy(Rv.begin(), Rv.end(), R);
  if (Cvals) {
    s->assembleK(CT, C);
    std::copy(C.begin(), C.end(), Cvals);
  }
}

void orc_last_system(void* h, double* rhs, double* lhs_vals) {
  Sim* s = (Sim*)h;
  if (rhs && !s->lastRhs.empty()) std::copy(s->lastRhs.begin(), s->lastRhs.end(), rhs);
  if (lhs_vals && !s->lastLhs.empty()) std::copy(s->lastLhs.begin(), s->lastLhs.end(), lhs_vals);
}

void orc_rest_volumes(void* h, double* vol, double* one_ring) {
  Sim* s = (Sim*)h;
  if (vol) std::copy(s->vol.begin(), s->vol.end(), vol);
  if (one_ring) std::copy(s->oneRing.begin(), s->oneRing.end(), one_ring);
}

}  // extern "C"
