// Assembly, right-hand side, PCG and state-update kernels of the backward-Euler Newton step.
//
// Reference path replaced (relative to the reference root):
//   gather assembly            src/Geometry/TET_MESH_FASTER.cpp:113-131 (ascending-tet sums, no atomics)
//   force scatter              src/Geometry/TET_MESH.cpp:756-767 (ascending-tet order, as a gather)
//   system formation           src/Timestepper/BACKWARD_EULER_POSITION.cpp:194-244
//   BDF-1 right-hand side      src/Timestepper/BDF_1.cpp:143-150
//   Jacobi preconditioner      ext/eigen-3.3.7/Eigen/src/IterativeLinearSolvers/BasicPreconditioners.h:65-79
//   conjugate gradients        ext/eigen-3.3.7/Eigen/src/IterativeLinearSolvers/ConjugateGradient.h:28-91
//   state update               src/Timestepper/BACKWARD_EULER_POSITION.cpp:262-314
//
// Every reduction is deterministic: each workgroup writes its partial sum to a fixed slot, and
// the last workgroup to finish (found with an integer atomic counter) adds the partials in a
// fixed order.  There are no floating-point atomics anywhere.
#include "solver_kernels.h"

#include <cfloat>

namespace hobak {

// ---------------------------------------------------------------------------------------------
// reductions
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Block-reduce NV values, publish the partials, and return true in the last block to arrive,
// with the grid totals in `tot` (valid in every lane of that block).
template <int NV>
__device__ bool grid_reduce(double (&vals)[NV], const Reduce& R, double (&tot)[NV]) {
  __shared__ double sh[NV][kBlock / 64];
  __shared__ int is_last;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; k++) {
    const double s = wave_sum(vals[k]);
    if (lane == 0) sh[k][wave] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; k++) {
      double s = 0.0;
#pragma unroll
      for (int w = 0; w < kBlock / 64; w++) s += sh[k][w];
      R.partial[k * gridDim.x + blockIdx.x] = s;
    }
    __threadfence();
    const unsigned prev = atomicAdd(R.counter, 1u);
    is_last = (prev == gridDim.x - 1);
  }
  __syncthreads();
  if (!is_last) return false;
  __threadfence();
  double acc[NV];
#pragma unroll
  for (int k = 0; k < NV; k++) {
    double s = 0.0;
    for (int b = threadIdx.x; b < (int)gridDim.x; b += kBlock)
      s += __builtin_nontemporal_load(&R.partial[k * gridDim.x + b]);
    acc[k] = wave_sum(s);
  }
  __syncthreads();
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < NV; k++) sh[k][wave] = acc[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; k++) {
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; w++) s += sh[k][w];
    tot[k] = s;
  }
  if (threadIdx.x == 0) *R.counter = 0u;
  return true;
}

// First slot of row r; its k-th block sits at slot_base + k * kSlice.
__device__ __forceinline__ int slot_base(const int* slice_off, int r) {
  return slice_off[r >> 6] + (r & 63);
}

// ---------------------------------------------------------------------------------------------
// step bookkeeping
// ---------------------------------------------------------------------------------------------
// _positionOld = _position; _velocityOld = _velocity; applyKinematicConstraints
// (BACKWARD_EULER_POSITION.cpp:152-157, TIMESTEPPER.cpp:105-122)
__global__ void begin_step_kernel(int nv, double* u, const double* v, double* uOld, double* vOld,
                                  const unsigned char* pinned, const double* pin_target) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nv) return;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const size_t i = 3 * (size_t)r + c;
    uOld[i] = u[i];
    vOld[i] = v[i];
    if (pinned[r]) u[i] = pin_target[i];
  }
}

// _position += xDelta (BACKWARD_EULER_POSITION.cpp:262-266)
template <typename T>
__global__ void apply_dx_kernel(int n, double* u, const T* x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) u[i] += (double)x[i];
}

// _velocity = (u - u_old)/dt, then S-filtered (BACKWARD_EULER_POSITION.cpp:300-314)
__global__ void end_step_kernel(int nv, double invDt, const double* u, const double* uOld, double* v,
                                const unsigned char* pinned) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nv) return;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const size_t i = 3 * (size_t)r + c;
    v[i] = pinned[r] ? 0.0 : invDt * (u[i] - uOld[i]);
  }
}

// ---------------------------------------------------------------------------------------------
// assembly
// ---------------------------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void write_precond(const T (&a)[9], T* minv, int nrows_pad, int r,
                                              int precond) {
  if (precond == 0) {  // DiagonalPreconditioner: 1/a_ii, 1 where a_ii == 0
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const T d = a[4 * c];
      minv[c * (size_t)nrows_pad + r] = (d != T(0)) ? T(1) / d : T(1);
    }
    return;
  }
  // 3x3 inverse by cofactors
  auto cof = [&](int i, int j) {
    const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
    return a[3 * i1 + j1] * a[3 * i2 + j2] - a[3 * i1 + j2] * a[3 * i2 + j1];
  };
  const T c0 = cof(0, 0), c1 = cof(1, 0), c2 = cof(2, 0);
  const T det = (c0 * a[0] + c1 * a[3]) + c2 * a[6];
  const T invdet = T(1) / det;
  minv[0 * (size_t)nrows_pad + r] = c0 * invdet;
  minv[1 * (size_t)nrows_pad + r] = c1 * invdet;
  minv[2 * (size_t)nrows_pad + r] = c2 * invdet;
#pragma unroll
  for (int row = 1; row < 3; row++)
#pragma unroll
    for (int k = 0; k < 3; k++) minv[(3 * row + k) * (size_t)nrows_pad + r] = cof(k, row) * invdet;
}

// One lane per block row.  mode 0: out = sum of staged contributions (the stiffness, -K).
// mode 1: out = LHS = S (D - K) S + (I - S) with the preconditioner of its diagonal block.
template <typename T>
__global__ __launch_bounds__(kBlock) void assemble_kernel(AsmArgs a) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= a.nrows) return;
  const int width = a.row_len[r];
  const int base = slot_base(a.slice_off, r);
  const T* stage = (const T*)a.stage_h;
  const T* D = (const T*)a.D;
  T* out = (T*)a.out;
  const size_t nt = (size_t)a.nt;
  const size_t ns = (size_t)a.nslots;
  const bool pr = a.mode == 1 && a.pinned[r];
  for (int k = 0; k < width; k++) {
    const int slot = base + k * kSlice;
    const int c = a.col[slot];
    T acc[9];
#pragma unroll
    for (int e = 0; e < 9; e++) acc[e] = T(0);
    const int q1 = a.cptr[slot + 1];
    for (int q = a.cptr[slot]; q < q1; q++) {
      const int code = a.clist[q];
      const size_t tet = (size_t)(code >> 5);
      const int lb = code & 15;
      const T* src = stage + (size_t)lb * 9 * nt + tet;
      if (code & 16) {
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
          for (int j = 0; j < 3; j++) acc[3 * i + j] += src[(3 * j + i) * nt];
      } else {
#pragma unroll
        for (int e = 0; e < 9; e++) acc[e] += src[e * nt];
      }
    }
    if (a.mode == 0) {
#pragma unroll
      for (int e = 0; e < 9; e++) out[e * ns + slot] = acc[e];
      continue;
    }
    const bool pin = pr || a.pinned[c];
    if (a.ground_on && c == r && !pin) {
      const double y = a.rest[3 * (size_t)r + 1] + a.u[3 * (size_t)r + 1];
      if (y < a.ground_y) acc[4] += (T)a.ground_k;
    }
    T val[9];
#pragma unroll
    for (int e = 0; e < 9; e++) {
      if (pin)
        val[e] = (c == r && (e % 4) == 0) ? T(1) : T(0);
      else
        val[e] = D[e * ns + slot] + acc[e];
      out[e * ns + slot] = val[e];
    }
    if (c == r) write_precond<T>(val, (T*)a.minv, a.nrows_pad, r, a.precond);
  }
}

// D = M/dt^2 - C/dt, E = M/dt - C, C = alpha M + beta K_rest, from the assembled rest stiffness
// (-K_rest, mode-0 output).  Order of operations as BACKWARD_EULER_POSITION.cpp:198, 208 and
// TIMESTEPPER.cpp:232-233.
template <typename T>
__global__ void constant_parts_kernel(ConstArgs a) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= a.nrows) return;
  const int width = a.row_len[r];
  const int base = slot_base(a.slice_off, r);
  const size_t ns = (size_t)a.nslots;
  for (int k = 0; k < width; k++) {
    const int slot = base + k * kSlice;
    const int c = a.col[slot];
    const bool real = true;
#pragma unroll
    for (int e = 0; e < 9; e++) {
      const double krest = -a.kneg[e * ns + slot];
      const bool dg = real && (c == r) && (e % 4 == 0);
      const double m = dg ? a.mass[3 * (size_t)r + e / 4] : 0.0;
      const double cc = dg ? a.ra * m + a.rb * krest : a.rb * krest;
      double d = dg ? m * a.invDt2 - cc * a.invDt : -(cc * a.invDt);
      double ee = dg ? a.invDt * m - cc : -cc;
      if (!real) {
        d = 0.0;
        ee = 0.0;
      }
      ((T*)a.D)[e * ns + slot] = (T)d;
      ((T*)a.E)[e * ns + slot] = (T)ee;
      ((T*)a.C)[e * ns + slot] = (T)(real ? cc : 0.0);
    }
  }
}

// ---------------------------------------------------------------------------------------------
// right-hand side + PCG initialisation
// ---------------------------------------------------------------------------------------------
// Per block row: R = ascending-tet gather of staged forces; b = (E v + R) + f_ext (BEP) or
// b = C vb + ((M a + R) + f_ext) (BDF_1); rhs = S b.  Then x = 0, r = rhs, z = M^-1 r, p = z and
// the grid sums |rhs|^2 and r.z seed Eigen's CG (ConjugateGradient.h:43-66).
template <typename T>
__global__ __launch_bounds__(kBlock) void rhs_init_kernel(RhsArgs a) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  double part[2] = {0.0, 0.0};
  if (r < a.nrows) {
    const T* sf = (const T*)a.stage_f;
    const size_t nt = (size_t)a.nt;
    double R[3] = {0.0, 0.0, 0.0};
    for (int q = a.vptr[r]; q < a.vptr[r + 1]; q++) {
      const int code = a.vlist[q];
      const size_t tet = (size_t)(code >> 2);
      const int loc = code & 3;
#pragma unroll
      for (int c = 0; c < 3; c++) R[c] += (double)sf[(3 * loc + c) * nt + tet];
    }
    if (a.ground_on) {
      const double y = a.rest[3 * (size_t)r + 1] + a.u[3 * (size_t)r + 1];
      if (y < a.ground_y) R[1] += a.ground_k * (a.ground_y - y);
    }
    const int width = a.row_len[r];
    const int base = slot_base(a.slice_off, r);
    const size_t ns = (size_t)a.nslots;
    const T* Mx = (const T*)a.Mvec;  // E (BEP) or C (BDF_1)
    double s[3] = {0.0, 0.0, 0.0};
    for (int k = 0; k < width; k++) {
      const int slot = base + k * kSlice;
      const int c = a.col[slot];
      double xv[3];
#pragma unroll
      for (int j = 0; j < 3; j++) {
        const size_t i = 3 * (size_t)c + j;
        xv[j] = a.bdf ? a.invDt * (a.u[i] - a.uOld[i]) - a.vOld[i] : a.vel[i];
      }
#pragma unroll
      for (int i = 0; i < 3; i++) {
        s[i] += (double)Mx[(3 * i + 0) * ns + slot] * xv[0];
        s[i] += (double)Mx[(3 * i + 1) * ns + slot] * xv[1];
        s[i] += (double)Mx[(3 * i + 2) * ns + slot] * xv[2];
      }
    }
    const bool pin = a.pinned[r];
    double b[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const size_t i = 3 * (size_t)r + c;
      const double fext = a.one_ring[r] * a.g[c];
      if (a.bdf) {
        const double acc = -a.invDt2 * (a.u[i] - a.uOld[i]) + a.invDt * a.vOld[i];
        b[c] = s[c] + (((a.mass[i] * acc + R[c]) + fext));
      } else {
        b[c] = (s[c] + R[c]) + fext;
      }
      if (pin) b[c] = 0.0;
      a.rhs64[i] = b[c];
      part[0] += b[c] * b[c];
    }
    if (a.init_cg) {
      T* rr = (T*)a.r;
      T* xx = (T*)a.x;
      T* zz = (T*)a.z;
      T* pp = (T*)a.p;
      const T* mi = (const T*)a.minv;
      const size_t np = (size_t)a.nrows_pad;
      T rv[3];
#pragma unroll
      for (int c = 0; c < 3; c++) rv[c] = (T)b[c];
      T zv[3];
      if (a.precond == 0) {
#pragma unroll
        for (int c = 0; c < 3; c++) zv[c] = mi[c * np + r] * rv[c];
      } else {
#pragma unroll
        for (int i = 0; i < 3; i++)
          zv[i] = mi[(3 * i) * np + r] * rv[0] + mi[(3 * i + 1) * np + r] * rv[1] +
                  mi[(3 * i + 2) * np + r] * rv[2];
      }
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const size_t i = 3 * (size_t)r + c;
        rr[i] = rv[c];
        xx[i] = T(0);
        zz[i] = zv[c];
        pp[i] = zv[c];
        part[1] += (double)rv[c] * (double)zv[c];
      }
    }
  }
  double tot[2];
  if (grid_reduce<2>(part, a.red, tot) && threadIdx.x == 0) {
    CgScalars* sc = a.sc;
    sc->maxR = tot[0];
    if (a.init_cg) {
      const double rhsNorm2 = tot[0];
      sc->rhsNorm2 = rhsNorm2;
      sc->iters = 0;
      sc->status = 0;
      sc->done = 0;
      sc->maxIters = a.max_iters;
      sc->beta = 0.0;
      sc->alpha = 0.0;
      sc->resNorm2 = rhsNorm2;
      if (rhsNorm2 == 0.0) {
        sc->done = 1;
      } else {
        const double thr = a.tol * a.tol * rhsNorm2;
        sc->threshold = thr > DBL_MIN ? thr : DBL_MIN;
        if (rhsNorm2 < sc->threshold) sc->done = 1;
        if (!(rhsNorm2 == rhsNorm2) || rhsNorm2 > DBL_MAX) {
          sc->done = 1;
          sc->status = 2;
        }
      }
      sc->absNew = tot[1];
    }
  }
}

// ---------------------------------------------------------------------------------------------
// PCG iteration: three kernels
//   spmv_dot:  t = A p,  pt = p.t,  alpha = absNew / pt
//   update:    x += alpha p, r -= alpha t, |r|^2 test, z = M^-1 r, absNew = r.z, beta
//   p_update:  p = z + beta p
// Each kernel returns at once when the solve is finished, so a fixed chunk of iterations can be
// queued (or replayed from a graph) and polled from the host between chunks.
// ---------------------------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kBlock) void cg_spmv_kernel(CgArgs a) {
  if (a.sc->done) return;
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  double part[1] = {0.0};
  const T* p = (const T*)a.p;
  if (r < a.nrows) {
    const int width = a.row_len[r];
    const int base = slot_base(a.slice_off, r);
    const size_t ns = (size_t)a.nslots;
    const T* A = (const T*)a.A;
    T s0 = T(0), s1 = T(0), s2 = T(0);
    for (int k = 0; k < width; k++) {
      const int slot = base + k * kSlice;
      const size_t c = 3 * (size_t)a.col[slot];
      const T p0 = p[c], p1 = p[c + 1], p2 = p[c + 2];
      s0 += A[0 * ns + slot] * p0 + A[1 * ns + slot] * p1 + A[2 * ns + slot] * p2;
      s1 += A[3 * ns + slot] * p0 + A[4 * ns + slot] * p1 + A[5 * ns + slot] * p2;
      s2 += A[6 * ns + slot] * p0 + A[7 * ns + slot] * p1 + A[8 * ns + slot] * p2;
    }
    T* t = (T*)a.t;
    const size_t i = 3 * (size_t)r;
    t[i] = s0;
    t[i + 1] = s1;
    t[i + 2] = s2;
    part[0] = (double)p[i] * (double)s0 + (double)p[i + 1] * (double)s1 + (double)p[i + 2] * (double)s2;
  }
  double tot[1];
  if (grid_reduce<1>(part, a.red, tot) && threadIdx.x == 0) {
    a.sc->alpha = a.sc->absNew / tot[0];
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void cg_update_kernel(CgArgs a) {
  if (a.sc->done) return;
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  double part[2] = {0.0, 0.0};
  if (r < a.nrows) {
    const T alpha = (T)a.sc->alpha;
    T* x = (T*)a.x;
    T* rr = (T*)a.r;
    T* z = (T*)a.z;
    const T* p = (const T*)a.p;
    const T* t = (const T*)a.t;
    const T* mi = (const T*)a.minv;
    const size_t np = (size_t)a.nrows_pad;
    T rv[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const size_t i = 3 * (size_t)r + c;
      x[i] += alpha * p[i];
      rv[c] = rr[i] - alpha * t[i];
      rr[i] = rv[c];
      part[0] += (double)rv[c] * (double)rv[c];
    }
    T zv[3];
    if (a.precond == 0) {
#pragma unroll
      for (int c = 0; c < 3; c++) zv[c] = mi[c * np + r] * rv[c];
    } else {
#pragma unroll
      for (int i = 0; i < 3; i++)
        zv[i] = mi[(3 * i) * np + r] * rv[0] + mi[(3 * i + 1) * np + r] * rv[1] +
                mi[(3 * i + 2) * np + r] * rv[2];
    }
#pragma unroll
    for (int c = 0; c < 3; c++) {
      z[3 * (size_t)r + c] = zv[c];
      part[1] += (double)rv[c] * (double)zv[c];
    }
  }
  double tot[2];
  if (grid_reduce<2>(part, a.red, tot) && threadIdx.x == 0) {
    CgScalars* sc = a.sc;
    sc->resNorm2 = tot[0];
    if (!(tot[0] == tot[0]) || tot[0] > DBL_MAX || !(tot[1] == tot[1])) {
      sc->done = 1;
      sc->status = 2;
    } else if (tot[0] < sc->threshold) {
      sc->done = 1;
    } else {
      const double absOld = sc->absNew;
      sc->absNew = tot[1];
      sc->beta = tot[1] / absOld;
      sc->iters += 1;
      if (sc->iters >= sc->maxIters) {
        sc->done = 1;
        sc->status = 1;
      }
    }
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void cg_p_kernel(CgArgs a) {
  if (a.sc->done) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 3 * a.nrows) return;
  const T beta = (T)a.sc->beta;
  T* p = (T*)a.p;
  p[i] = ((const T*)a.z)[i] + beta * p[i];
}

// ---------------------------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------------------------
static inline int grid_for(int n, int block = kBlock) { return (n + block - 1) / block; }

hipError_t launch_begin_step(int nv, double* u, const double* v, double* uOld, double* vOld,
                             const unsigned char* pinned, const double* target, hipStream_t s) {
  hipLaunchKernelGGL(begin_step_kernel, dim3(grid_for(nv)), dim3(kBlock), 0, s, nv, u, v, uOld, vOld,
                     pinned, target);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_apply_dx(int n, double* u, const T* x, hipStream_t s) {
  hipLaunchKernelGGL((apply_dx_kernel<T>), dim3(grid_for(n)), dim3(kBlock), 0, s, n, u, x);
  return hipGetLastError();
}

hipError_t launch_end_step(int nv, double invDt, const double* u, const double* uOld, double* v,
                           const unsigned char* pinned, hipStream_t s) {
  hipLaunchKernelGGL(end_step_kernel, dim3(grid_for(nv)), dim3(kBlock), 0, s, nv, invDt, u, uOld, v,
                     pinned);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_assemble(const AsmArgs& a, hipStream_t s) {
  hipLaunchKernelGGL((assemble_kernel<T>), dim3(grid_for(a.nrows)), dim3(kBlock), 0, s, a);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_constant_parts(const ConstArgs& a, hipStream_t s) {
  hipLaunchKernelGGL((constant_parts_kernel<T>), dim3(grid_for(a.nrows)), dim3(kBlock), 0, s, a);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_rhs_init(const RhsArgs& a, hipStream_t s) {
  hipLaunchKernelGGL((rhs_init_kernel<T>), dim3(grid_for(a.nrows)), dim3(kBlock), 0, s, a);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_cg_spmv(const CgArgs& a, hipStream_t s) {
  hipLaunchKernelGGL((cg_spmv_kernel<T>), dim3(grid_for(a.nrows)), dim3(kBlock), 0, s, a);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_cg_rest(const CgArgs& a, hipStream_t s) {
  hipLaunchKernelGGL((cg_update_kernel<T>), dim3(grid_for(a.nrows)), dim3(kBlock), 0, s, a);
  hipLaunchKernelGGL((cg_p_kernel<T>), dim3(grid_for(3 * a.nrows)), dim3(kBlock), 0, s, a);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_cg_iteration(const CgArgs& a, hipStream_t s) {
  hipError_t e = launch_cg_spmv<T>(a, s);
  if (e != hipSuccess) return e;
  return launch_cg_rest<T>(a, s);
}

// Plain streaming copy, used only to calibrate hardware byte counters against a known size.
__global__ void copy_kernel(const double4* __restrict__ src, double4* __restrict__ dst, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

hipError_t launch_copy(const void* src, void* dst, size_t bytes, hipStream_t s) {
  const size_t n = bytes / sizeof(double4);
  hipLaunchKernelGGL(copy_kernel, dim3(4096), dim3(kBlock), 0, s, (const double4*)src, (double4*)dst, n);
  return hipGetLastError();
}

template hipError_t launch_apply_dx<double>(int, double*, const double*, hipStream_t);
template hipError_t launch_apply_dx<float>(int, double*, const float*, hipStream_t);
template hipError_t launch_assemble<double>(const AsmArgs&, hipStream_t);
template hipError_t launch_assemble<float>(const AsmArgs&, hipStream_t);
template hipError_t launch_constant_parts<double>(const ConstArgs&, hipStream_t);
template hipError_t launch_constant_parts<float>(const ConstArgs&, hipStream_t);
template hipError_t launch_rhs_init<double>(const RhsArgs&, hipStream_t);
template hipError_t launch_rhs_init<float>(const RhsArgs&, hipStream_t);
template hipError_t launch_cg_iteration<double>(const CgArgs&, hipStream_t);
template hipError_t launch_cg_iteration<float>(const CgArgs&, hipStream_t);
template hipError_t launch_cg_spmv<double>(const CgArgs&, hipStream_t);
template hipError_t launch_cg_spmv<float>(const CgArgs&, hipStream_t);
template hipError_t launch_cg_rest<double>(const CgArgs&, hipStream_t);
template hipError_t launch_cg_rest<float>(const CgArgs&, hipStream_t);

}  // namespace hobak

namespace hobak {

// computeHyperelasticForces' scatter as a per-vertex gather in ascending tet order
// (TET_MESH.cpp:756-767); output in FP64.
template <typename T>
__global__ void force_gather_kernel(int nv, int nt, const int* vptr, const int* vlist,
                                    const T* stage_f, double* R) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nv) return;
  double s[3] = {0.0, 0.0, 0.0};
  for (int q = vptr[r]; q < vptr[r + 1]; q++) {
    const int code = vlist[q];
    const size_t tet = (size_t)(code >> 2);
    const int loc = code & 3;
#pragma unroll
    for (int c = 0; c < 3; c++) s[c] += (double)stage_f[(3 * loc + c) * (size_t)nt + tet];
  }
#pragma unroll
  for (int c = 0; c < 3; c++) R[3 * (size_t)r + c] = s[c];
}

template <typename T>
hipError_t launch_force_gather(int nv, int nt, const int* vptr, const int* vlist, const T* stage_f,
                               double* R, hipStream_t s) {
  hipLaunchKernelGGL((force_gather_kernel<T>), dim3((nv + kBlock - 1) / kBlock), dim3(kBlock), 0, s,
                     nv, nt, vptr, vlist, stage_f, R);
  return hipGetLastError();
}
template hipError_t launch_force_gather<double>(int, int, const int*, const int*, const double*,
                                                double*, hipStream_t);
template hipError_t launch_force_gather<float>(int, int, const int*, const int*, const float*,
                                               double*, hipStream_t);

}  // namespace hobak
