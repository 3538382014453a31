// Fused per-tet element kernel: deformation gradient -> FP64 SVD -> PK1 force -> eigen-clamped
// stiffness, one lane per tet, outputs staged as coalesced SoA planes.
//
// Reference path replaced (relative to the reference root):
//   TET_MESH::computeF / computeFs        src/Geometry/TET_MESH.cpp:643-655, 706-714
//   TET_MESH::computeSVDs -> svd_rv       src/Geometry/TET_MESH.cpp:689-701
//   TET_MESH::computeHyperelasticForces   src/Geometry/TET_MESH.cpp:736-770 (per-tet part)
//   TET_MESH_FASTER::computeHyperelasticClampedHessian element loop
//                                         src/Geometry/TET_MESH_FASTER.cpp:93-103
//   TET_MESH::computeInternalForces       src/Geometry/TET_MESH.cpp:816-858 (per-tet part: the
//                                         GREEN_DAMPING force, damping_force_kernel)
//   TET_MESH::computeFdots                src/Geometry/TET_MESH.cpp:660-684 (fdot_kernel)
//   TET_MESH_FASTER::computeDampingHessian element loop, src/Geometry/TET_MESH_FASTER.cpp:144-150
//                                         (damping_hessian_kernel)
//
// The reference forms the 12x12 stiffness as dF/dx^T (-V H9) dF/dx with dense 9x12 products.
// Here the clamped H9 is kept in the singular-value frame, where it is block diagonal (a 3x3
// scaling block and three 2x2 twist/flip blocks), and each 3x3 vertex-pair block of the stiffness
// is formed directly as  V_e * U * M_vw * U^T  with
//   M_vw[a][a']  = sum_{b,b'} w_v[b] Hhat[(a,b),(a',b')] w_w[b'],   w = (G DmInv) V,
// G the 4x3 map from corner positions to edge vectors.  That is the same matrix up to round-off,
// at about a third of the flops, and it needs only ~40 live values per lane.  Blocks and forces
// are staged tet-major (block pair(x, y) of tet t at 9 * (pair * nt + t), the force of corner x at
// 3 * (x * nt + t)); a wave's 64 consecutive tets pass each block through an LDS tile, so every
// store instruction writes 64 consecutive values.  The assembly and the force gather read them
// through per-slot and per-vertex lists in ascending tet order.  The corner positions are staged
// in LDS: a workgroup's 128 consecutive tets touch a few dozen to a few hundred distinct vertices
// (a list built with the mesh, ascending), whose x = X + u the workgroup loads once, coalesced,
// into three coordinate planes; each tet then reads its corners there through 16-bit local
// indices (eloc) instead of gathering 4 x 2 x 24 bytes from HBM.
#include "device_math.h"
#include "kernels.h"

namespace hobak {

// Inputs of one tet's evaluation, read from the instance record.
struct ElementIn {
  int nt;
  double mu, lam, alpha;
  const double (*xs)[kElementVerts];           // LDS: the workgroup's vertex positions, by plane
  const HOBAK_GLOBAL unsigned short* eloc;     // 4 per tet: its corners' indices into xs
  const HOBAK_GLOBAL double* dminv;
  const HOBAK_GLOBAL double* vol;
  HOBAK_GLOBAL void* stage_h;
  HOBAK_GLOBAL void* stage_f;
  int with_hessian;
};

// Orders a wave's LDS accesses: the tile written by every lane is then read by every lane.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The store of N values per lane for a wave's 64 consecutive tets (lane l: the wave's l-th tet) to
// dst[N * l + e], of which the first nvalid tets are written.  Through the wave's LDS tile, so
// that each store instruction writes 64 consecutive values (whole lines), where a direct store
// would write 64 pieces N values apart.
template <typename T, int N>
__device__ __forceinline__ void wave_store(HOBAK_GLOBAL T* dst, const T (&v)[N], int nvalid, T* tile) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int e = 0; e < N; e++) tile[N * lane + e] = v[e];
  wave_sync();
  const int n = N * nvalid;
#pragma unroll
  for (int k = 0; k < N; k++) {
    const int i = 64 * k + lane;
    if (i < n) dst[i] = tile[i];
  }
  wave_sync();  // the tile is written again by the next store
}

// GREEN_DAMPING stress mu F (F^T Fdot + Fdot^T F) (src/Damping/Volume/GREEN_DAMPING.cpp:49-52),
// with Fdot read from its 9 planes.
__device__ __forceinline__ void green_pk1(double mu, const double (&F)[3][3], const HOBAK_GLOBAL double* fdot,
                                          size_t nt, int t, double (&P)[3][3]) {
  double Fd[3][3];
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int c = 0; c < 3; c++) Fd[r][c] = fdot[(3 * r + c) * nt + t];
  double X[3][3];  // F^T Fdot
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) X[i][j] = (F[0][i] * Fd[0][j] + F[1][i] * Fd[1][j]) + F[2][i] * Fd[2][j];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++)
      P[i][j] = ((mu * F[i][0]) * (X[0][j] + X[j][0]) + (mu * F[i][1]) * (X[1][j] + X[j][1])) +
                (mu * F[i][2]) * (X[2][j] + X[j][2]);
}

// Tet t of a wave whose tets start at t0, the first nvalid of them real (the other lanes repeat
// the last tet, and their results are not stored).
template <typename T, int MAT>
__device__ __forceinline__ void element_tet(const ElementIn& a, int t, int t0, int nvalid, T* tile) {
  const size_t nt = (size_t)a.nt;
  // corner positions x = X + u (TET_MESH::setDisplacement), staged in LDS, and edge matrix Ds
  const unsigned long long lc = ((const HOBAK_GLOBAL unsigned long long*)a.eloc)[t];  // 4 x 16 bits
  const int loc[4] = {(int)(lc & 0xffff), (int)((lc >> 16) & 0xffff), (int)((lc >> 32) & 0xffff),
                      (int)(lc >> 48)};
  double x[4][3];
#pragma unroll
  for (int k = 0; k < 4; k++)
#pragma unroll
    for (int c = 0; c < 3; c++) x[k][c] = a.xs[c][loc[k]];
  double Di[3][3];
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int c = 0; c < 3; c++) Di[r][c] = a.dminv[(3 * r + c) * nt + t];
  double F[3][3];
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int c = 0; c < 3; c++)
      F[r][c] = ((x[1][r] - x[0][r]) * Di[0][c] + (x[2][r] - x[0][r]) * Di[1][c]) +
                (x[3][r] - x[0][r]) * Di[2][c];

  double Ud[3][3], Sd[3], Vd[3][3];
  svd_rv(F, Ud, Sd, Vd);

  const T vol = (T)a.vol[t];
  // Wd: row v = d x_v-coefficient of F's rows (the non-zero pattern of dF/dx, TET_MESH.cpp:197-245)
  T Wd[4][3];
#pragma unroll
  for (int j = 0; j < 3; j++) {
    Wd[0][j] = (T)((-Di[0][j] - Di[1][j]) - Di[2][j]);
    Wd[1][j] = (T)Di[0][j];
    Wd[2][j] = (T)Di[1][j];
    Wd[3][j] = (T)Di[2][j];
  }
  T U[3][3], V[3][3], S[3];
#pragma unroll
  for (int i = 0; i < 3; i++) {
    S[i] = (T)Sd[i];
#pragma unroll
    for (int j = 0; j < 3; j++) {
      U[i][j] = (T)Ud[i][j];
      V[i][j] = (T)Vd[i][j];
    }
  }

  // PK1 and the element force f_v = -V_e P Wd_v^T
  T P[3][3];
  if constexpr (sizeof(T) < sizeof(double)) {
    // reduced precision: strain-like factors in FP64 (see snh_pk1_split)
    if (MAT == 0)
      snh_pk1_split<T>((T)a.mu, (T)a.lam, F, P);
    else
      arap_pk1_split<T>((T)a.mu, U, Sd, V, P);
  } else if (MAT == 0) {
    T Ft[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int j = 0; j < 3; j++) Ft[i][j] = (T)F[i][j];
    snh_pk1<T>((T)a.mu, (T)a.lam, (T)a.alpha, Ft, P);
  } else {
    arap_pk1<T>((T)a.mu, U, S, V, P);
  }
  auto sf = (HOBAK_GLOBAL T*)a.stage_f;
#pragma unroll
  for (int v = 0; v < 4; v++) {
    T f[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const T s = (P[c][0] * Wd[v][0] + P[c][1] * Wd[v][1]) + P[c][2] * Wd[v][2];
      f[c] = -vol * s;
    }
    wave_store<T, 3>(sf + 3 * ((size_t)v * nt + t0), f, nvalid, tile);
  }
  if (!a.with_hessian) return;

  SigmaHessian<T> h;
  if (MAT == 0)
    snh_sigma_hessian<T>((T)a.mu, (T)a.lam, S, h);
  else
    arap_sigma_hessian<T>((T)a.mu, S, h);

  T w[4][3];
#pragma unroll
  for (int v = 0; v < 4; v++)
#pragma unroll
    for (int b = 0; b < 3; b++) w[v][b] = Wd[v][0] * V[0][b] + Wd[v][1] * V[1][b] + Wd[v][2] * V[2][b];

  auto sh = (HOBAK_GLOBAL T*)a.stage_h;
#pragma unroll
  for (int v = 0; v < 4; v++) {
#pragma unroll
    for (int v2 = v; v2 < 4; v2++) {
      T M[3][3];
#pragma unroll
      for (int p = 0; p < 3; p++)
#pragma unroll
        for (int q = 0; q < 3; q++) {
          if (p == q) {
            T s = h.Hs[p][p] * w[v][p] * w[v2][p];
#pragma unroll
            for (int b = 0; b < 3; b++)
              if (b != p) s += h.d[3 - p - b] * w[v][b] * w[v2][b];
            M[p][q] = s;
          } else {
            M[p][q] = h.Hs[p][q] * w[v][p] * w[v2][q] + h.o[3 - p - q] * w[v][q] * w[v2][p];
          }
        }
      T UM[3][3];
#pragma unroll
      for (int c = 0; c < 3; c++)
#pragma unroll
        for (int q = 0; q < 3; q++) UM[c][q] = U[c][0] * M[0][q] + U[c][1] * M[1][q] + U[c][2] * M[2][q];
      T B[9];  // block (v, v2): rows of corner v
#pragma unroll
      for (int c = 0; c < 3; c++)
#pragma unroll
        for (int c2 = 0; c2 < 3; c2++)
          B[3 * c + c2] = vol * (UM[c][0] * U[c2][0] + UM[c][1] * U[c2][1] + UM[c][2] * U[c2][2]);
      wave_store<T, 9>(sh + 9 * ((size_t)staged_pair(v, v2) * nt + t0), B, nvalid, tile);
    }
  }
}

// Segmented over instances (one lane per tet, kElementBlock tets per virtual workgroup); the
// material is a per-instance, workgroup-uniform branch.
template <typename T>
__global__ __launch_bounds__(kElementBlock) void element_kernel(Batch b, int src, int with_hessian,
                                                                const double* u_ext, void* sf_ext) {
  __shared__ T tiles[kElementBlock / 64][9 * 64];  // one store tile per wave
  __shared__ double xs[3][kElementVerts];          // the workgroup's vertices x = X + u
  for (int ws = blockIdx.x; ws < b.total; ws += gridDim.x) {
    const int vb = xcd_vb(ws, b.total);
    int lb, lg;
    const HOBAK_GLOBAL InstDev* __restrict__ I = seg_find(b, vb, lb, lg);
    ElementIn a;
    a.nt = I->nt;
    a.mu = I->mu;
    a.lam = I->lam;
    a.alpha = I->alpha;
    a.xs = xs;
    a.eloc = gbl(I->eloc);
    a.dminv = gbl(I->dminv);
    a.vol = gbl(I->vol);
    a.stage_h = gbl(I->stage_h);
    a.stage_f = gbl(sf_ext ? sf_ext : I->stage_f);
    a.with_hessian = with_hessian;
    const int material = I->material;
    // stage the workgroup's vertex positions (the same sum X + u each corner gather made)
    {
      const auto rest = gbl(I->rest);
      const auto u = gbl(src == ELEM_EXT ? u_ext : src == ELEM_TMP ? I->tmp64 : I->u);
      const auto evert = gbl(I->evert);
      const int v0 = gbl(I->evptr)[lb], nvx = gbl(I->evptr)[lb + 1] - v0;
      for (int i = threadIdx.x; i < nvx; i += kElementBlock) {
        const size_t v = 3 * (size_t)evert[v0 + i];
#pragma unroll
        for (int c = 0; c < 3; c++) xs[c][i] = rest[v + c] + u[v + c];
      }
    }
    __syncthreads();
    const int t = lb * kElementBlock + threadIdx.x;
    const int t0 = t - (threadIdx.x & 63);  // the wave's first tet
    if (t0 < a.nt) {                        // (uniform over the wave)
      const int nvalid = min(64, a.nt - t0);
      const int tc = min(t, a.nt - 1);
      T* tile = tiles[threadIdx.x >> 6];
      if (material == 0)
        element_tet<T, 0>(a, tc, t0, nvalid, tile);
      else
        element_tet<T, 1>(a, tc, t0, nvalid, tile);
    }
    __syncthreads();  // xs is staged again for the next workgroup of this launch slot
  }
}

template <typename T>
hipError_t launch_element(const Batch& b, int src, int with_hessian, int cap, hipStream_t s,
                          const double* u_ext, void* stage_f_ext) {
  if (b.total <= 0) return hipSuccess;
  const int grid = cap > 0 && cap < b.total ? cap : b.total;
  hipLaunchKernelGGL((element_kernel<T>), dim3(grid), dim3(kElementBlock), 0, s, b, src, with_hessian,
                     u_ext, stage_f_ext);
  return hipGetLastError();
}

template hipError_t launch_element<double>(const Batch&, int, int, int, hipStream_t, const double*, void*);
template hipError_t launch_element<float>(const Batch&, int, int, int, hipStream_t, const double*, void*);

// The GREEN_DAMPING part of computeInternalForces, after the element kernel: for each tet of an
// instance with energy damping, f_v += -V P_damping d_v^T into the staged forces (one lane per
// tet; instances without damping leave at once).  The reference sums the two stresses before the
// dF/dx^T product (TET_MESH.cpp:831-834); adding the two products instead differs by round-off
// only, and keeps the element kernel, and so every undamped instance, bit for bit unchanged.
template <typename T>
__global__ __launch_bounds__(kElementBlock) void damping_force_kernel(Batch b, int src) {
  for (int ws = blockIdx.x; ws < b.total; ws += gridDim.x) {
    const int vb = xcd_vb(ws, b.total);
    int lb, lg;
    const HOBAK_GLOBAL InstDev* __restrict__ I = seg_find(b, vb, lb, lg);
    if (!I->damping) continue;
    const int t = lb * kElementBlock + threadIdx.x;
    const int nt = I->nt;
    if (t >= nt) continue;
    const size_t n = (size_t)nt;
    auto tets = gbl((const int*)I->tets);
    auto rest = gbl(I->rest);
    auto u = gbl(src == ELEM_TMP ? I->tmp64 : I->u);
    auto dminv = gbl(I->dminv);
    const int vid[4] = {tets[4 * (size_t)t], tets[4 * (size_t)t + 1], tets[4 * (size_t)t + 2],
                        tets[4 * (size_t)t + 3]};
    double x[4][3], Di[3][3], F[3][3];
#pragma unroll
    for (int k = 0; k < 4; k++)
#pragma unroll
      for (int c = 0; c < 3; c++) x[k][c] = rest[3 * (size_t)vid[k] + c] + u[3 * (size_t)vid[k] + c];
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
      for (int c = 0; c < 3; c++) Di[r][c] = dminv[(3 * r + c) * n + t];
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
      for (int c = 0; c < 3; c++)
        F[r][c] = ((x[1][r] - x[0][r]) * Di[0][c] + (x[2][r] - x[0][r]) * Di[1][c]) +
                  (x[3][r] - x[0][r]) * Di[2][c];
    double P[3][3];
    green_pk1(I->dmu, F, gbl(I->fdot), n, t, P);
    double d[4][3];
#pragma unroll
    for (int j = 0; j < 3; j++) {
      d[0][j] = (-Di[0][j] - Di[1][j]) - Di[2][j];
      d[1][j] = Di[0][j];
      d[2][j] = Di[1][j];
      d[3][j] = Di[2][j];
    }
    const double vol = gbl(I->vol)[t];
    auto sf = gbl((T*)I->stage_f);
#pragma unroll
    for (int v = 0; v < 4; v++) {
      auto dst = sf + 3 * ((size_t)v * n + t);
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const double s = (P[c][0] * d[v][0] + P[c][1] * d[v][1]) + P[c][2] * d[v][2];
        dst[c] = dst[c] + (T)(-vol * s);
      }
    }
  }
}

template <typename T>
hipError_t launch_damping_force(const Batch& b, int src, int cap, hipStream_t s) {
  if (b.total <= 0) return hipSuccess;
  const int grid = cap > 0 && cap < b.total ? cap : b.total;
  hipLaunchKernelGGL((damping_force_kernel<T>), dim3(grid), dim3(kElementBlock), 0, s, b, src);
  return hipGetLastError();
}

template hipError_t launch_damping_force<double>(const Batch&, int, int, hipStream_t);
template hipError_t launch_damping_force<float>(const Batch&, int, int, hipStream_t);

// TET_MESH::computeFdots: one lane per tet, Fdot = V DmInv with V's columns v_k - v_0
// (TET_MESH.cpp:666-682), the product in the same summation order as F = Ds DmInv.
__global__ __launch_bounds__(kBlock) void fdot_kernel(int nt, const int4* __restrict__ tets,
                                                      const double* __restrict__ dminv,
                                                      const double* __restrict__ vel,
                                                      double* __restrict__ fdot) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= nt) return;
  const size_t n = (size_t)nt;
  const int4 q = tets[t];
  const int vid[4] = {q.x, q.y, q.z, q.w};
  double w[4][3];
#pragma unroll
  for (int k = 0; k < 4; k++)
#pragma unroll
    for (int c = 0; c < 3; c++) w[k][c] = vel[3 * (size_t)vid[k] + c];
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int c = 0; c < 3; c++)
      fdot[(3 * r + c) * n + t] = ((w[1][r] - w[0][r]) * dminv[(3 * 0 + c) * n + t] +
                                   (w[2][r] - w[0][r]) * dminv[(3 * 1 + c) * n + t]) +
                                  (w[3][r] - w[0][r]) * dminv[(3 * 2 + c) * n + t];
}

hipError_t launch_fdots(const InstDev& h, const double* vel, double* fdot, hipStream_t s) {
  if (h.nt <= 0) return hipSuccess;
  hipLaunchKernelGGL(fdot_kernel, dim3((h.nt + kBlock - 1) / kBlock), dim3(kBlock), 0, s, h.nt,
                     (const int4*)h.tets, h.dminv, vel, fdot);
  return hipGetLastError();
}

// TET_MESH_FASTER::computeDampingHessian, per tet: the stiffness blocks of vol * dP/dFdot with
// dP/dFdot = mu (F F^T (x) I + F^T-coupling) (GREEN_DAMPING.cpp:58-72).  Block (v, w) of
// dF/dx^T H dF/dx is
//   mu vol ( (F F^T) (d_v . d_w) + (F d_w) (F d_v)^T ),
// d_v the row of dF/dx for corner v.  H is positive semi-definite (it is the Hessian of the
// convex quadratic mu |sym(F^T Fdot)|^2 in Fdot), so the reference's eigenvalue clamp
// (GREEN_DAMPING.cpp:79-88) changes it only by round-off and is not repeated here.  Staged with
// the positive sign, like the element stiffness; the assembled result is -C.
template <typename T>
__global__ __launch_bounds__(kElementBlock) void damping_hessian_kernel(Batch b, int src) {
  for (int ws = blockIdx.x; ws < b.total; ws += gridDim.x) {
    const int vb = xcd_vb(ws, b.total);
    int lb, lg;
    const HOBAK_GLOBAL InstDev* __restrict__ I = seg_find(b, vb, lb, lg);
    const int t = lb * kElementBlock + threadIdx.x;
    const int nt = I->nt;
    if (t >= nt) continue;
    const size_t n = (size_t)nt;
    auto tets = gbl((const int*)I->tets);
    auto rest = gbl(I->rest);
    auto u = gbl(src == ELEM_TMP ? I->tmp64 : I->u);
    auto dminv = gbl(I->dminv);
    const int vid[4] = {tets[4 * (size_t)t], tets[4 * (size_t)t + 1], tets[4 * (size_t)t + 2],
                        tets[4 * (size_t)t + 3]};
    double x[4][3], Di[3][3], F[3][3];
#pragma unroll
    for (int k = 0; k < 4; k++)
#pragma unroll
      for (int c = 0; c < 3; c++) x[k][c] = rest[3 * (size_t)vid[k] + c] + u[3 * (size_t)vid[k] + c];
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
      for (int c = 0; c < 3; c++) Di[r][c] = dminv[(3 * r + c) * n + t];
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
      for (int c = 0; c < 3; c++)
        F[r][c] = ((x[1][r] - x[0][r]) * Di[0][c] + (x[2][r] - x[0][r]) * Di[1][c]) +
                  (x[3][r] - x[0][r]) * Di[2][c];
    double d[4][3];
#pragma unroll
    for (int j = 0; j < 3; j++) {
      d[0][j] = (-Di[0][j] - Di[1][j]) - Di[2][j];
      d[1][j] = Di[0][j];
      d[2][j] = Di[1][j];
      d[3][j] = Di[2][j];
    }
    double G[3][3], g[4][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int j = 0; j < 3; j++) G[i][j] = (F[i][0] * F[j][0] + F[i][1] * F[j][1]) + F[i][2] * F[j][2];
#pragma unroll
    for (int v = 0; v < 4; v++)
#pragma unroll
      for (int i = 0; i < 3; i++) g[v][i] = (F[i][0] * d[v][0] + F[i][1] * d[v][1]) + F[i][2] * d[v][2];
    const double s = I->dmu * gbl(I->vol)[t];
    auto sh = gbl((T*)I->stage_h);
    // block (v, w) is s (G (d_v . d_w) + g_w g_v^T), G symmetric, so block (w, v) is its exact
    // transpose: the upper blocks are staged, rows of corner v
#pragma unroll
    for (int v = 0; v < 4; v++)
#pragma unroll
      for (int w = v; w < 4; w++) {
        const double dd = (d[v][0] * d[w][0] + d[v][1] * d[w][1]) + d[v][2] * d[w][2];
        auto dst = sh + 9 * ((size_t)staged_pair(v, w) * n + t);
#pragma unroll
        for (int c = 0; c < 3; c++)
#pragma unroll
          for (int c2 = 0; c2 < 3; c2++) dst[3 * c + c2] = (T)(s * (G[c][c2] * dd + g[w][c] * g[v][c2]));
      }
  }
}

template <typename T>
hipError_t launch_damping_hessian(const Batch& b, int src, int cap, hipStream_t s) {
  if (b.total <= 0) return hipSuccess;
  const int grid = cap > 0 && cap < b.total ? cap : b.total;
  hipLaunchKernelGGL((damping_hessian_kernel<T>), dim3(grid), dim3(kElementBlock), 0, s, b, src);
  return hipGetLastError();
}

template hipError_t launch_damping_hessian<double>(const Batch&, int, int, hipStream_t);
template hipError_t launch_damping_hessian<float>(const Batch&, int, int, hipStream_t);

}  // namespace hobak
