// Per-element math for the fused element kernel (gfx950, one lane per tet).
//
// Everything here is written for registers: 3x3 matrices are row-major [row][col] arrays indexed
// only by compile-time constants, so the compiler keeps them in VGPRs (no scratch).
//
// Reference behaviour reproduced (paths relative to the reference root):
//   svd_rv                      src/util/MATRIX_UTIL.cpp:76-95 on Eigen's two-sided Jacobi SVD
//                               (ext/eigen-3.3.7/Eigen/src/SVD/JacobiSVD.h:660-780)
//   SNH::PK1 / clampedHessian   src/Hyperelastic/Volume/SNH.cpp:45-50, 94-144
//   ARAP::PK1 / clampedHessian  src/Hyperelastic/Volume/ARAP.cpp:48-55, 113-135
#pragma once
#include <hip/hip_runtime.h>

#include <cfloat>

namespace hobak {

#define HD __device__ __forceinline__

template <typename T>
HD T hmax(T a, T b) { return a > b ? a : b; }

HD double det3(const double m[3][3]) {
  // cofactor expansion along the first row, in the order of Eigen's determinant_impl<3>
  return m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) -
         m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
         m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
}

template <typename T>
HD T det3t(const T m[3][3]) {
  return m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) -
         m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
         m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
}

// The SVD below rounds every product and sum on its own (no fused multiply-add), in the order of
// Eigen's JacobiSVD.  Its result at a matrix with repeated singular values (an improper rotation
// has sigma = (1, 1, 1) up to sign) is one of infinitely many valid factorizations, chosen by
// round-off; ARAP's PK1 and stiffness depend on that choice, so only identical rounding reproduces
// the reference's values there.

// Jacobi rotation that diagonalises the symmetric 2x2 [[x, y], [y, z]] (Jacobi/Jacobi.h:83-113).
HD void make_jacobi(double x, double y, double z, double& c, double& s) {
#pragma clang fp contract(off)
  const double deno = 2.0 * fabs(y);
  if (deno < DBL_MIN) {
    c = 1.0;
    s = 0.0;
    return;
  }
  const double tau = (x - z) / deno;
  const double w = sqrt(tau * tau + 1.0);
  const double t = (tau > 0.0) ? 1.0 / (tau + w) : 1.0 / (tau - w);
  const double sign_t = t > 0.0 ? 1.0 : -1.0;
  const double n = 1.0 / sqrt(t * t + 1.0);
  s = -sign_t * (y > 0.0 ? 1.0 : -1.0) * fabs(t) * n;
  c = n;
}

// One 2x2 step of the two-sided Jacobi sweep on the (P, Q) plane (JacobiSVD.h:704-731 with
// misc/RealSvd2x2.h:19-51).  Returns whether the plane was still off-diagonal.
template <int P, int Q>
HD bool jacobi_pair(double W[3][3], double U[3][3], double V[3][3], double& maxDiag) {
#pragma clang fp contract(off)
  const double thr = hmax(DBL_MIN, 2.0 * DBL_EPSILON * maxDiag);
  if (!(fabs(W[P][Q]) > thr || fabs(W[Q][P]) > thr)) return false;
  double m00 = W[P][P], m01 = W[P][Q], m10 = W[Q][P], m11 = W[Q][Q];
  double c1, s1;
  const double t = m00 + m11;
  const double d = m10 - m01;
  if (fabs(d) < DBL_MIN) {
    s1 = 0.0;
    c1 = 1.0;
  } else {
    const double u = t / d;
    const double tmp = sqrt(1.0 + u * u);
    s1 = 1.0 / tmp;
    c1 = u / tmp;
  }
  {
    const double x0 = m00, y0 = m10, x1 = m01, y1 = m11;
    m00 = c1 * x0 + s1 * y0;
    m01 = c1 * x1 + s1 * y1;
    m11 = -s1 * x1 + c1 * y1;
  }
  double cr, sr;
  make_jacobi(m00, m01, m11, cr, sr);
  const double cl = c1 * cr + s1 * sr;
  const double sl = -c1 * sr + s1 * cr;
#pragma unroll
  for (int i = 0; i < 3; i++) {
    const double x = W[P][i], y = W[Q][i];
    W[P][i] = cl * x + sl * y;
    W[Q][i] = -sl * x + cl * y;
  }
#pragma unroll
  for (int i = 0; i < 3; i++) {
    const double x = U[i][P], y = U[i][Q];
    U[i][P] = cl * x + sl * y;
    U[i][Q] = -sl * x + cl * y;
  }
#pragma unroll
  for (int i = 0; i < 3; i++) {
    const double x = W[i][P], y = W[i][Q];
    W[i][P] = cr * x - sr * y;
    W[i][Q] = sr * x + cr * y;
  }
#pragma unroll
  for (int i = 0; i < 3; i++) {
    const double x = V[i][P], y = V[i][Q];
    V[i][P] = cr * x - sr * y;
    V[i][Q] = sr * x + cr * y;
  }
  maxDiag = hmax(maxDiag, hmax(fabs(W[P][P]), fabs(W[Q][Q])));
  return true;
}

template <int I, int J>
HD void swap_cols(double S[3], double U[3][3], double V[3][3]) {
  double t = S[I];
  S[I] = S[J];
  S[J] = t;
#pragma unroll
  for (int r = 0; r < 3; r++) {
    t = U[r][I];
    U[r][I] = U[r][J];
    U[r][J] = t;
    t = V[r][I];
    V[r][I] = V[r][J];
    V[r][J] = t;
  }
}

// F = U diag(S) V^T with S descending and U, V orthogonal (Eigen JacobiSVD semantics), then the
// svd_rv fix-up that folds a reflection into S[2].
HD void svd_rv(const double F[3][3], double U[3][3], double S[3], double V[3][3]) {
#pragma clang fp contract(off)
  double scale = 0.0;
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) scale = hmax(scale, fabs(F[i][j]));
  if (scale == 0.0) scale = 1.0;
  const double inv = 1.0 / scale;
  double W[3][3];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      W[i][j] = F[i][j] / scale;
      U[i][j] = (i == j) ? 1.0 : 0.0;
      V[i][j] = (i == j) ? 1.0 : 0.0;
    }
  (void)inv;
  double maxDiag = hmax(hmax(fabs(W[0][0]), fabs(W[1][1])), fabs(W[2][2]));
  // Sweeps until no plane needs a rotation; the cap is far above the handful a 3x3 needs and
  // guarantees every lane leaves the loop.
  for (int sweep = 0; sweep < 24; sweep++) {
    bool any = jacobi_pair<1, 0>(W, U, V, maxDiag);
    any |= jacobi_pair<2, 0>(W, U, V, maxDiag);
    any |= jacobi_pair<2, 1>(W, U, V, maxDiag);
    if (!any) break;
  }
#pragma unroll
  for (int i = 0; i < 3; i++) {
    const double a = W[i][i];
    S[i] = fabs(a) * scale;
    if (a < 0.0) {
#pragma unroll
      for (int r = 0; r < 3; r++) U[r][i] = -U[r][i];
    }
  }
  // descending selection sort, first maximum wins (Eigen's maxCoeff order)
  {
    int pos = 0;
    if (S[1] > S[pos]) pos = 1;
    if (S[2] > S[pos]) pos = 2;
    if (S[pos] != 0.0) {
      if (pos == 1) swap_cols<0, 1>(S, U, V);
      if (pos == 2) swap_cols<0, 2>(S, U, V);
      if (S[2] > S[1] && S[2] != 0.0) swap_cols<1, 2>(S, U, V);
    }
  }
  double UVt[3][3];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) UVt[i][j] = (U[i][0] * V[j][0] + U[i][1] * V[j][1]) + U[i][2] * V[j][2];
  const double L22 = det3(UVt);
  const double detU = det3(U);
  const double detV = det3(V);
  if (detU < 0.0 && detV > 0.0) {
#pragma unroll
    for (int r = 0; r < 3; r++) U[r][2] *= L22;
  }
  if (detU > 0.0 && detV < 0.0) {
#pragma unroll
    for (int r = 0; r < 3; r++) V[r][2] *= L22;
  }
  S[2] = S[2] * L22;
}

// Cyclic Jacobi eigen-decomposition of a symmetric 3x3 (eigenvectors in the columns of Q).
template <typename T, int P, int Q>
HD void eig_rot(T A[3][3], T Qm[3][3]) {
  if (A[P][Q] == T(0)) return;
  const T theta = (A[Q][Q] - A[P][P]) / (T(2) * A[P][Q]);
  const T t = (theta >= T(0) ? T(1) : T(-1)) / (fabs(theta) + sqrt(theta * theta + T(1)));
  const T c = T(1) / sqrt(t * t + T(1));
  const T s = t * c;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const T akp = A[k][P], akq = A[k][Q];
    A[k][P] = c * akp - s * akq;
    A[k][Q] = s * akp + c * akq;
  }
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const T apk = A[P][k], aqk = A[Q][k];
    A[P][k] = c * apk - s * aqk;
    A[Q][k] = s * apk + c * aqk;
  }
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const T qkp = Qm[k][P], qkq = Qm[k][Q];
    Qm[k][P] = c * qkp - s * qkq;
    Qm[k][Q] = s * qkp + c * qkq;
  }
}

// Projection of a symmetric 3x3 onto the PSD cone: the matrix itself when it is positive
// definite (Sylvester's criterion), otherwise Q max(Lambda, 0) Q^T.
template <typename T>
HD void psd_project3(T A[3][3]) {
  const T m1 = A[0][0];
  const T m2 = A[0][0] * A[1][1] - A[0][1] * A[1][0];
  const T m3 = det3t(A);
  if (m1 > T(0) && m2 > T(0) && m3 > T(0)) return;
  T W[3][3], Qm[3][3];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      W[i][j] = A[i][j];
      Qm[i][j] = (i == j) ? T(1) : T(0);
    }
  for (int sweep = 0; sweep < 24; sweep++) {
    const T off = W[0][1] * W[0][1] + W[0][2] * W[0][2] + W[1][2] * W[1][2];
    const T dia = W[0][0] * W[0][0] + W[1][1] * W[1][1] + W[2][2] * W[2][2];
    if (!(off > T(1e-40) * dia)) break;
    eig_rot<T, 0, 1>(W, Qm);
    eig_rot<T, 0, 2>(W, Qm);
    eig_rot<T, 1, 2>(W, Qm);
  }
  T lam[3];
#pragma unroll
  for (int k = 0; k < 3; k++) lam[k] = W[k][k] > T(0) ? W[k][k] : T(0);
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      T s = T(0);
#pragma unroll
      for (int k = 0; k < 3; k++) s += Qm[i][k] * lam[k] * Qm[j][k];
      A[i][j] = s;
    }
}

// The clamped Hessian of an isotropic energy in the singular-value frame X = U^T F V:
//   Hs        3x3 block acting on (X00, X11, X22)   (scaling modes)
//   d[i], o[i] the 2x2 block [[d, o], [o, d]] on (X_jk, X_kj), {i, j, k} = {0, 1, 2}
//               (twist eigenvalue lt with vector (1,-1)/sqrt2, flip lf with (1,1)/sqrt2:
//                d = (lt + lf)/2, o = (lf - lt)/2)
// All eigenvalues clamped at zero as in SNH.cpp:137-140 and ARAP.cpp:129-131.
template <typename T>
struct SigmaHessian {
  T Hs[3][3];
  T d[3];
  T o[3];
};

template <typename T>
HD void snh_sigma_hessian(T mu, T lam, const T S[3], SigmaHessian<T>& h) {
  const T J = S[0] * S[1] * S[2];
  const T front = lam * (J - T(1)) - mu;
#pragma unroll
  for (int i = 0; i < 3; i++) {
    T lt = front * S[i] + mu;
    T lf = -front * S[i] + mu;
    lt = lt > T(0) ? lt : T(0);
    lf = lf > T(0) ? lf : T(0);
    h.d[i] = T(0.5) * (lt + lf);
    h.o[i] = T(0.5) * (lf - lt);
  }
  const T s00 = S[0] * S[0], s11 = S[1] * S[1], s22 = S[2] * S[2];
  h.Hs[0][0] = mu + lam * s11 * s22;
  h.Hs[1][1] = mu + lam * s00 * s22;
  h.Hs[2][2] = mu + lam * s00 * s11;
  const T fo = lam * (T(2) * J - T(1)) - mu;
  h.Hs[0][1] = h.Hs[1][0] = fo * S[2];
  h.Hs[0][2] = h.Hs[2][0] = fo * S[1];
  h.Hs[1][2] = h.Hs[2][1] = fo * S[0];
  psd_project3<T>(h.Hs);
}

template <typename T>
HD void arap_sigma_hessian(T mu, const T S[3], SigmaHessian<T>& h) {
  const T two_mu = T(2) * mu;
  const T lt[3] = {mu * (T(2) - T(4) / (S[1] + S[2])), mu * (T(2) - T(4) / (S[0] + S[2])),
                   mu * (T(2) - T(4) / (S[0] + S[1]))};
#pragma unroll
  for (int i = 0; i < 3; i++) {
    const T t = lt[i] >= T(0) ? lt[i] : T(0);
    h.d[i] = T(0.5) * (t + two_mu);
    h.o[i] = T(0.5) * (two_mu - t);
  }
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) h.Hs[i][j] = (i == j) ? two_mu : T(0);
}

// PK1 of SNH from F: mu F + lambda' (J - alpha) dJ/dF (SNH.cpp:45-50, MATRIX_UTIL.cpp:561-568).
template <typename T>
HD void snh_pk1(T mu, T lam, T alpha, const T F[3][3], T P[3][3]) {
  // columns of dJ/dF: f1 x f2, f2 x f0, f0 x f1 (f_c = column c of F)
  T G[3][3];
  G[0][0] = F[1][1] * F[2][2] - F[2][1] * F[1][2];
  G[1][0] = F[2][1] * F[0][2] - F[0][1] * F[2][2];
  G[2][0] = F[0][1] * F[1][2] - F[1][1] * F[0][2];
  G[0][1] = F[1][2] * F[2][0] - F[2][2] * F[1][0];
  G[1][1] = F[2][2] * F[0][0] - F[0][2] * F[2][0];
  G[2][1] = F[0][2] * F[1][0] - F[1][2] * F[0][0];
  G[0][2] = F[1][0] * F[2][1] - F[2][0] * F[1][1];
  G[1][2] = F[2][0] * F[0][1] - F[0][0] * F[2][1];
  G[2][2] = F[0][0] * F[1][1] - F[1][0] * F[0][1];
  const T Jm = det3t(F) - alpha;
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) P[i][j] = mu * F[i][j] + lam * Jm * G[i][j];
}

// PK1 of ARAP from its SVD: R 2 mu (S - I), R = U V^T, S = V Sigma V^T (ARAP.cpp:48-55).
template <typename T>
HD void arap_pk1(T mu, const T U[3][3], const T Sg[3], const T V[3][3], T P[3][3]) {
  T R[3][3], Sm[3][3];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      R[i][j] = U[i][0] * V[j][0] + U[i][1] * V[j][1] + U[i][2] * V[j][2];
      Sm[i][j] = V[i][0] * Sg[0] * V[j][0] + V[i][1] * Sg[1] * V[j][1] + V[i][2] * Sg[2] * V[j][2];
    }
  const T tm = T(2) * mu;
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      T s = T(0);
#pragma unroll
      for (int k = 0; k < 3; k++) s += R[i][k] * (tm * (Sm[k][j] - (k == j ? T(1) : T(0))));
      P[i][j] = s;
    }
}

// Reduced-precision PK1s with the near-rest cancellation taken out in FP64.  At F = I both forms
// above are differences of O(mu) terms that cancel to O(mu * strain), which in FP32 leaves a
// relative force error of ~1e-7 / strain.  Rewritten with strain-like factors that vanish at rest,
// formed in FP64 from the FP64 F and SVD, the T arithmetic that follows carries only T's relative
// error of the force itself:
//   SNH   mu F + lam (J - alpha) cof F = mu (F - cof F) + lam (J - 1) cof F   (alpha = 1 + mu / lam)
//   ARAP  R 2 mu (S - I) = 2 mu U (Sigma - I) V^T
template <typename T>
HD void snh_pk1_split(T mu, T lam, const double F[3][3], T P[3][3]) {
  double G[3][3];
  G[0][0] = F[1][1] * F[2][2] - F[2][1] * F[1][2];
  G[1][0] = F[2][1] * F[0][2] - F[0][1] * F[2][2];
  G[2][0] = F[0][1] * F[1][2] - F[1][1] * F[0][2];
  G[0][1] = F[1][2] * F[2][0] - F[2][2] * F[1][0];
  G[1][1] = F[2][2] * F[0][0] - F[0][2] * F[2][0];
  G[2][1] = F[0][2] * F[1][0] - F[1][2] * F[0][0];
  G[0][2] = F[1][0] * F[2][1] - F[2][0] * F[1][1];
  G[1][2] = F[2][0] * F[0][1] - F[0][0] * F[2][1];
  G[2][2] = F[0][0] * F[1][1] - F[1][0] * F[0][1];
  const T j1 = (T)(det3t(F) - 1.0);
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) P[i][j] = mu * (T)(F[i][j] - G[i][j]) + lam * j1 * (T)G[i][j];
}

template <typename T>
HD void arap_pk1_split(T mu, const T U[3][3], const double Sd[3], const T V[3][3], T P[3][3]) {
  T d[3];
#pragma unroll
  for (int k = 0; k < 3; k++) d[k] = (T)(Sd[k] - 1.0);
  const T tm = T(2) * mu;
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++)
      P[i][j] = tm * ((U[i][0] * d[0] * V[j][0] + U[i][1] * d[1] * V[j][1]) + U[i][2] * d[2] * V[j][2]);
}

}  // namespace hobak
