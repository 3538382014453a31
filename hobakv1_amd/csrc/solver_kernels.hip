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

#include <hip/hip_ext.h>

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

// Sum of one value over the workgroup, returned in every lane.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV]) {
  __shared__ double sh[NV][kBlock / 64];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; k++) {
    const double s = wave_sum(v[k]);
    if (lane == 0) sh[k][wave] = s;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; k++) {
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; w++) s += sh[k][w];
    v[k] = s;
  }
  __syncthreads();
}

// Publish this workgroup's sums of (a, b) to the Partials arrays.
__device__ __forceinline__ void publish(double a, double b, const Partials& P, bool two) {
  double v[2] = {a, b};
  block_sum<2>(v);
  if (threadIdx.x == 0) {
    P.a[blockIdx.x] = v[0];
    if (two) P.b[blockIdx.x] = v[1];
  }
}

// Every workgroup reduces the previous kernel's partials in the same fixed order, so all of
// them (and every run) obtain bit-identical totals.
__device__ __forceinline__ void gather_partials(const Partials& P, bool two, double& ta, double& tb) {
  double v[2] = {0.0, 0.0};
  for (int i = threadIdx.x; i < P.n; i += kBlock) {
    v[0] += P.a[i];
    if (two) v[1] += P.b[i];
  }
  block_sum<2>(v);
  ta = v[0];
  tb = v[1];
}

// Slot of the k-th block of row r (slice height 1 << sh).
__device__ __forceinline__ int slot_of(const int* slice_off, int sh, int r, int k) {
  return slice_off[r >> sh] + (k << sh) + (r & ((1 << sh) - 1));
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
__device__ __forceinline__ void write_precond(const T (&a)[9], T* minv, int r, int precond) {
  if (precond == 0) {  // DiagonalPreconditioner: 1/a_ii, 1 where a_ii == 0
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const T d = a[4 * c];
      minv[3 * (size_t)r + c] = (d != T(0)) ? T(1) / d : T(1);
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
  T* m = minv + (size_t)kMinvStride * r;
  m[0] = c0 * invdet;
  m[1] = c1 * invdet;
  m[2] = c2 * invdet;
#pragma unroll
  for (int row = 1; row < 3; row++)
#pragma unroll
    for (int k = 0; k < 3; k++) m[3 * row + k] = cof(k, row) * invdet;
}

// One lane per block slot.  mode 0: out = sum of staged contributions (the stiffness, -K).
// mode 1: out = LHS = S (D - K) S + (I - S); the diagonal slot also writes its preconditioner.
template <typename T>
__global__ __launch_bounds__(kBlock) void assemble_kernel(AsmArgs a) {
  const int slot = blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= a.nslots) return;
  const int r = a.slot_row[slot];
  if (r < 0) return;
  const int c = a.col[slot];
  const size_t ns = (size_t)a.nslots;
  T acc[9];
#pragma unroll
  for (int e = 0; e < 9; e++) acc[e] = T(0);
  // the slot's contributions are contiguous and in ascending tet order
  const T* src = (const T*)a.stage_h + 9 * (size_t)a.cptr[slot];
  const T* end = (const T*)a.stage_h + 9 * (size_t)a.cptr[slot + 1];
  for (; src < end; src += 9) {
#pragma unroll
    for (int e = 0; e < 9; e++) acc[e] += src[e];
  }
  T* out = (T*)a.out;
  if (a.mode == 0) {
#pragma unroll
    for (int e = 0; e < 9; e++) out[e * ns + slot] = acc[e];
    return;
  }
  const bool pin = a.pinned[r] || a.pinned[c];
  if (a.ground_on && c == r && !pin) {
    const double y = a.rest[3 * (size_t)r + 1] + a.u[3 * (size_t)r + 1];
    if (y < a.ground_y) acc[4] += (T)a.ground_k;
  }
  const T* D = (const T*)a.D;
  T val[9];
#pragma unroll
  for (int e = 0; e < 9; e++) {
    val[e] = pin ? ((c == r && (e % 4) == 0) ? T(1) : T(0)) : D[e * ns + slot] + acc[e];
    out[e * ns + slot] = val[e];
  }
  if (c == r) write_precond<T>(val, (T*)a.minv, r, a.precond);
}

// D = M/dt^2 - C/dt, E = M/dt - C, C = alpha M + beta K_rest, from the assembled rest stiffness
// (-K_rest, mode-0 output).  Order of operations as BACKWARD_EULER_POSITION.cpp:198, 208 and
// TIMESTEPPER.cpp:232-233.
template <typename T>
__global__ void constant_parts_kernel(ConstArgs a) {
  const int slot = blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= a.nslots) return;
  const int r = a.slot_row[slot];
  if (r < 0) return;
  const int c = a.col[slot];
  const size_t ns = (size_t)a.nslots;
#pragma unroll
  for (int e = 0; e < 9; e++) {
    const double krest = -(double)((const T*)a.kneg)[e * ns + slot];
    const bool dg = (c == r) && (e % 4 == 0);
    const double m = dg ? a.mass[3 * (size_t)r + e / 4] : 0.0;
    const double cc = dg ? a.ra * m + a.rb * krest : a.rb * krest;
    const double d = dg ? m * a.invDt2 - cc * a.invDt : -(cc * a.invDt);
    const double ee = dg ? a.invDt * m - cc : -cc;
    ((T*)a.D)[e * ns + slot] = (T)d;
    ((T*)a.E)[e * ns + slot] = (T)ee;
    ((T*)a.C)[e * ns + slot] = (T)cc;
  }
}

// ---------------------------------------------------------------------------------------------
// right-hand side + PCG initialisation
// ---------------------------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void apply_precond(const T* minv, int precond, int r, const T (&rv)[3],
                                              T (&zv)[3]) {
  if (precond == 0) {
#pragma unroll
    for (int c = 0; c < 3; c++) zv[c] = minv[3 * (size_t)r + c] * rv[c];
  } else {
    const T* m = minv + (size_t)kMinvStride * r;
#pragma unroll
    for (int i = 0; i < 3; i++) zv[i] = m[3 * i] * rv[0] + m[3 * i + 1] * rv[1] + m[3 * i + 2] * rv[2];
  }
}

// Per block row: R = ascending-tet gather of staged forces; b = (E v + R) + f_ext (BEP) or
// b = C vb + ((M a + R) + f_ext) (BDF_1); rhs = S b.  Then x = 0, r = rhs, z = M^-1 r and p = 0
// (the first SpMV forms p = z + 0 p), and the partial sums of |rhs|^2 and r.z seed Eigen's CG
// (ConjugateGradient.h:43-66).
template <typename T>
__global__ __launch_bounds__(kBlock) void rhs_init_kernel(RhsArgs a) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  double part[2] = {0.0, 0.0};
  if (r < a.nrows) {
    const T* sf = (const T*)a.stage_f;
    double R[3] = {0.0, 0.0, 0.0};
    for (int q = a.vptr[r]; q < a.vptr[r + 1]; q++) {
#pragma unroll
      for (int c = 0; c < 3; c++) R[c] += (double)sf[3 * (size_t)q + c];
    }
    if (a.ground_on) {
      const double y = a.rest[3 * (size_t)r + 1] + a.u[3 * (size_t)r + 1];
      if (y < a.ground_y) R[1] += a.ground_k * (a.ground_y - y);
    }
    const size_t ns = (size_t)a.nslots;
    const T* Mx = (const T*)a.Mvec;  // E (BEP) or C (BDF_1)
    double s[3] = {0.0, 0.0, 0.0};
    const int len = a.row_len[r];
    for (int k = 0; k < len; k++) {
      const int slot = slot_of(a.slice_off, a.slice_shift, r, k);
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
      T rv[3], zv[3];
#pragma unroll
      for (int c = 0; c < 3; c++) rv[c] = (T)b[c];
      apply_precond<T>((const T*)a.minv, a.precond, r, rv, zv);
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const size_t i = 3 * (size_t)r + c;
        ((T*)a.r)[i] = rv[c];
        ((T*)a.x)[i] = T(0);
        ((T*)a.z)[i] = zv[c];
        ((T*)a.p0)[i] = T(0);
        part[1] += (double)rv[c] * (double)zv[c];
      }
    }
  }
  publish(part[0], part[1], a.out, true);
}

// sc[0] from the right-hand side's partials: Eigen's early exits on a zero or already small
// right-hand side (ConjugateGradient.h:45-61) and the first r.z.
__global__ __launch_bounds__(kBlock) void cg_init_kernel(InitArgs a) {
  double bb, rz;
  gather_partials(a.in, true, bb, rz);
  if (threadIdx.x != 0) return;
  CgScalars* s = &a.sc[0];
  const double thr0 = a.tol * a.tol * bb;
  const double thr = thr0 > DBL_MIN ? thr0 : DBL_MIN;
  int done = (bb == 0.0 || bb < thr || a.max_iters <= 0) ? 1 : 0;
  int status = 0;
  if (!(bb == bb) || bb > DBL_MAX) {
    done = 1;
    status = 2;
  }
  s->maxR = bb;
  s->rhsNorm2 = bb;
  s->resNorm2 = bb;
  s->threshold = thr;
  s->absNew = rz;
  s->alpha = 0.0;
  s->beta = 0.0;
  s->iters = 0;
  s->maxIters = a.max_iters;
  s->done = done;
  s->status = status;
}

// ---------------------------------------------------------------------------------------------
// PCG iteration `it`: two kernels (Eigen's loop body, ConjugateGradient.h:69-88, regrouped)
//   spmv(it):   S = sc[it & 1].  For it > 0 it first finishes iteration it - 1: from the update
//               kernel's partials, |r|^2 < threshold -> done (the converging pass is not
//               counted); else beta = r.z / S.absNew, absNew = r.z, iters + 1 (done at
//               maxIters).  Workgroup 0 writes that as the next set sc[(it + 1) & 1] (a
//               finished set is carried forward).  Then p_new = z + beta p_old, formed on the fly
//               at every column (p double-buffered by parity), t = A p_new, partials of p.t.
//   update(it): S' = sc[(it + 1) & 1]; alpha = S'.absNew / (p.t); x += alpha p_new;
//               r -= alpha t; z = M^-1 r; partials of r.r and r.z.
// Kernels of a finished solve return at once, so chunks of iterations can be queued ahead and
// the host polls the flag between chunks.  Scalars are read and written field by field: a
// by-value copy of the struct becomes a per-lane LDS buffer that costs microseconds per launch.
// ---------------------------------------------------------------------------------------------
// One wavefront per H-row slice, L = 64 / H lanes per row: lane = H * kk + rr handles the row's
// blocks k = kk, kk + L, ...; a row's L partial sums are combined with cross-lane adds.
template <typename T, int H>
__global__ __launch_bounds__(kBlock) void cg_spmv_kernel(CgArgs a) {
  const CgScalars* S = &a.sc[a.it & 1];
  CgScalars* N = &a.sc[(a.it + 1) & 1];
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  const int done = S->done;
  const double absOld = S->absNew;
  const int iters = S->iters;
  if (lead) {
    N->rhsNorm2 = S->rhsNorm2;
    N->threshold = S->threshold;
    N->absNew = absOld;
    N->alpha = S->alpha;
    N->beta = S->beta;
    N->resNorm2 = S->resNorm2;
    N->maxR = S->maxR;
    N->iters = iters;
    N->maxIters = S->maxIters;
    N->done = done;
    N->status = S->status;
  }
  if (done) return;
  T beta = T(0);
  if (a.it > 0) {
    double rr, rz;
    gather_partials(a.rz, true, rr, rz);
    if (!(rr == rr) || rr > DBL_MAX || !(rz == rz)) {
      if (lead) {
        N->resNorm2 = rr;
        N->done = 1;
        N->status = 2;
      }
      return;
    }
    if (rr < S->threshold) {
      if (lead) {
        N->resNorm2 = rr;
        N->done = 1;
      }
      return;
    }
    const double b = rz / absOld;
    const int maxIters = S->maxIters;
    if (lead) {
      N->resNorm2 = rr;
      N->beta = b;
      N->absNew = rz;
      N->iters = iters + 1;
      if (iters + 1 >= maxIters) {
        N->done = 1;
        N->status = 1;
      }
    }
    if (iters + 1 >= maxIters) return;
    beta = (T)b;
  }
  constexpr int L = 64 / H;
  const int slice = (int)((blockIdx.x * (unsigned)blockDim.x + threadIdx.x) >> 6);
  const int lane = threadIdx.x & 63;
  const int kk = lane / H;
  const int rr = lane % H;
  const int r = slice * H + rr;
  const T* pold = (const T*)((a.it & 1) ? a.p1 : a.p0);
  T* pnew = (T*)((a.it & 1) ? a.p0 : a.p1);
  const T* z = (const T*)a.z;
  const T* A = (const T*)a.A;
  const size_t ns = (size_t)a.nslots;
  T s0 = T(0), s1 = T(0), s2 = T(0);
  const bool live = r < a.nrows;
  if (live) {
    const int len = a.row_len[r];
    const int base = a.slice_off[slice] + rr;
    for (int k = kk; k < len; k += L) {
      const int slot = base + k * H;
      const size_t c = 3 * (size_t)a.col[slot];
      const T p0 = z[c] + beta * pold[c];
      const T p1 = z[c + 1] + beta * pold[c + 1];
      const T p2 = z[c + 2] + beta * pold[c + 2];
      s0 += A[0 * ns + slot] * p0 + A[1 * ns + slot] * p1 + A[2 * ns + slot] * p2;
      s1 += A[3 * ns + slot] * p0 + A[4 * ns + slot] * p1 + A[5 * ns + slot] * p2;
      s2 += A[6 * ns + slot] * p0 + A[7 * ns + slot] * p1 + A[8 * ns + slot] * p2;
    }
  }
#pragma unroll
  for (int off = H; off < 64; off <<= 1) {
    s0 += __shfl_xor(s0, off, 64);
    s1 += __shfl_xor(s1, off, 64);
    s2 += __shfl_xor(s2, off, 64);
  }
  double pt = 0.0;
  if (live && kk == 0) {
    const size_t i = 3 * (size_t)r;
    T* t = (T*)a.t;
    t[i] = s0;
    t[i + 1] = s1;
    t[i + 2] = s2;
    const T q0 = z[i] + beta * pold[i];
    const T q1 = z[i + 1] + beta * pold[i + 1];
    const T q2 = z[i + 2] + beta * pold[i + 2];
    pnew[i] = q0;
    pnew[i + 1] = q1;
    pnew[i + 2] = q2;
    pt = (double)q0 * (double)s0 + (double)q1 * (double)s1 + (double)q2 * (double)s2;
  }
  publish(pt, 0.0, a.pt, false);
}

// One lane per DOF, 21 rows (63 lanes) per wavefront so that a row's three DOFs share a wave and
// the 3x3 preconditioner apply can read its siblings' residuals with cross-lane moves.  A fixed
// grid strides over the waves, which keeps the partial-sum count (and the work of every
// workgroup that later re-reduces them) small.
template <typename T>
__global__ __launch_bounds__(kBlock) void cg_update_kernel(CgArgs a) {
  const CgScalars* S = &a.sc[(a.it + 1) & 1];
  if (S->done) return;
  const double absNew = S->absNew;
  double pt, unused;
  gather_partials(a.pt, false, pt, unused);
  const T alpha = (T)(absNew / pt);
  const int lane = threadIdx.x & 63;
  const int c = lane - 3 * (lane / 3);
  const int b = lane - c;
  const int nwaves = (a.nrows + 20) / 21;
  const int stride = (int)(gridDim.x * (kBlock / 64));
  const T* p = (const T*)((a.it & 1) ? a.p0 : a.p1);  // the p_new this iteration's SpMV wrote
  const T* mi = (const T*)a.minv;
  T* x = (T*)a.x;
  T* rr = (T*)a.r;
  T* z = (T*)a.z;
  const T* t = (const T*)a.t;
  double rr2 = 0.0, rz = 0.0;
  for (int wave = (int)((blockIdx.x * (unsigned)blockDim.x + threadIdx.x) >> 6); wave < nwaves;
       wave += stride) {
    const int row = wave * 21 + lane / 3;
    const bool live = lane < 63 && row < a.nrows;
    const size_t i = 3 * (size_t)row + c;
    T rv = T(0);
    if (live) {
      x[i] += alpha * p[i];
      rv = rr[i] - alpha * t[i];
      rr[i] = rv;
    }
    const T r0 = __shfl(rv, b, 64);
    const T r1 = __shfl(rv, b + 1, 64);
    const T r2 = __shfl(rv, b + 2, 64);
    if (live) {
      T zv;
      if (a.precond == 0) {
        zv = mi[i] * rv;
      } else {
        const T* m = mi + (size_t)kMinvStride * row + 3 * c;
        zv = m[0] * r0 + m[1] * r1 + m[2] * r2;
      }
      z[i] = zv;
      rr2 += (double)rv * (double)rv;
      rz += (double)rv * (double)zv;
    }
  }
  publish(rr2, rz, a.rz, true);
}

// ---------------------------------------------------------------------------------------------
// FP64 iterative refinement of the FP32 solve
//   begin:     b2 = |S b|^2; x64 = 0; res64 = S b
//   init:      r = (T) res64, x = 0, z = M^-1 r, p = 0 -> the inner PCG (same kernels as above)
//   residual:  x64 += x; res64 = rhs64 - A x64 with A's T values promoted to FP64
//   check:     res2 = |res64|^2; done when res2 < tol^2 b2
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void refine_begin_kernel(RefineArgs a, Partials rhs_parts) {
  double b2, unused;
  gather_partials(rhs_parts, false, b2, unused);
  const int ndof = 3 * a.nrows;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < ndof; i += gridDim.x * blockDim.x) {
    a.x64[i] = 0.0;
    a.res64[i] = a.rhs64[i];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    a.state->b2 = b2;
    a.state->res2 = b2;
    a.state->sweeps = 0;
    a.state->status = (!(b2 == b2) || b2 > DBL_MAX) ? 2 : 0;
    a.state->done = (b2 == 0.0 || a.state->status) ? 1 : 0;
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void refine_init_kernel(RefineArgs a) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  double rr = 0.0, rz = 0.0;
  if (r < a.nrows) {
    T rv[3], zv[3];
#pragma unroll
    for (int c = 0; c < 3; c++) rv[c] = (T)a.res64[3 * (size_t)r + c];
    apply_precond<T>((const T*)a.minv, a.precond, r, rv, zv);
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const size_t i = 3 * (size_t)r + c;
      ((T*)a.r)[i] = rv[c];
      ((T*)a.x)[i] = T(0);
      ((T*)a.z)[i] = zv[c];
      ((T*)a.p0)[i] = T(0);
      rr += (double)rv[c] * (double)rv[c];
      rz += (double)rv[c] * (double)zv[c];
    }
  }
  publish(rr, rz, a.out, true);
}

// One wavefront per H-row slice as in the SpMV; x64 is advanced by the inner solution on the fly
// for every column, and the kernel's own rows write it back.
template <typename T, int H>
__global__ __launch_bounds__(kBlock) void refine_residual_kernel(RefineArgs a) {
  constexpr int L = 64 / H;
  const int slice = (int)((blockIdx.x * (unsigned)blockDim.x + threadIdx.x) >> 6);
  const int lane = threadIdx.x & 63;
  const int kk = lane / H;
  const int rr = lane % H;
  const int r = slice * H + rr;
  const T* A = (const T*)a.A;
  const T* xi = (const T*)a.x;
  const size_t ns = (size_t)a.nslots;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  const bool live = r < a.nrows;
  if (live) {
    const int len = a.row_len[r];
    const int base = a.slice_off[slice] + rr;
    for (int k = kk; k < len; k += L) {
      const int slot = base + k * H;
      const size_t c = 3 * (size_t)a.col[slot];
      const double x0 = a.x64[c] + (double)xi[c];
      const double x1 = a.x64[c + 1] + (double)xi[c + 1];
      const double x2 = a.x64[c + 2] + (double)xi[c + 2];
      s0 += (double)A[0 * ns + slot] * x0 + (double)A[1 * ns + slot] * x1 + (double)A[2 * ns + slot] * x2;
      s1 += (double)A[3 * ns + slot] * x0 + (double)A[4 * ns + slot] * x1 + (double)A[5 * ns + slot] * x2;
      s2 += (double)A[6 * ns + slot] * x0 + (double)A[7 * ns + slot] * x1 + (double)A[8 * ns + slot] * x2;
    }
  }
#pragma unroll
  for (int off = H; off < 64; off <<= 1) {
    s0 += __shfl_xor(s0, off, 64);
    s1 += __shfl_xor(s1, off, 64);
    s2 += __shfl_xor(s2, off, 64);
  }
  double res2 = 0.0;
  if (live && kk == 0) {
    const double s[3] = {s0, s1, s2};
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const size_t i = 3 * (size_t)r + c;
      const double res = a.rhs64[i] - s[c];
      a.res64[i] = res;
      res2 += res * res;
    }
  }
  publish(res2, 0.0, a.out, false);
}

// x64 += x for the rows themselves happens after every residual kernel has read the old x64:
// a separate pass keeps the on-the-fly sums above race-free.
template <typename T>
__global__ void refine_accumulate_kernel(int n, double* x64, const T* x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x64[i] += (double)x[i];
}

__global__ __launch_bounds__(kBlock) void refine_check_kernel(RefineArgs a) {
  double res2, unused;
  gather_partials(a.out, false, res2, unused);
  if (threadIdx.x != 0) return;
  RefineState* st = a.state;
  st->res2 = res2;
  st->sweeps += 1;
  if (!(res2 == res2) || res2 > DBL_MAX) {
    st->done = 1;
    st->status = 2;
  } else if (res2 < a.tol * a.tol * st->b2) {
    st->done = 1;
  }
}

// computeHyperelasticForces' scatter as a per-vertex gather in ascending tet order
// (TET_MESH.cpp:756-767); output in FP64.
template <typename T>
__global__ void force_gather_kernel(int nv, const int* vptr, const T* stage_f, double* R) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nv) return;
  double s[3] = {0.0, 0.0, 0.0};
  for (int q = vptr[r]; q < vptr[r + 1]; q++) {
#pragma unroll
    for (int c = 0; c < 3; c++) s[c] += (double)stage_f[3 * (size_t)q + c];
  }
#pragma unroll
  for (int c = 0; c < 3; c++) R[3 * (size_t)r + c] = s[c];
}

// Plain streaming copy, used only to calibrate hardware byte counters against a known size.
__global__ void copy_kernel(const double4* __restrict__ src, double4* __restrict__ dst, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// ---------------------------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------------------------
static inline int grid_for(long n, int block = kBlock) { return (int)((n + block - 1) / block); }

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
  hipLaunchKernelGGL((assemble_kernel<T>), dim3(grid_for(a.nslots)), dim3(kBlock), 0, s, a);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_constant_parts(const ConstArgs& a, hipStream_t s) {
  hipLaunchKernelGGL((constant_parts_kernel<T>), dim3(grid_for(a.nslots)), dim3(kBlock), 0, s, a);
  return hipGetLastError();
}

int rhs_grid(int nrows) { return grid_for(nrows); }
int spmv_grid(int nrows, int sh) { return grid_for((long)((nrows + (1 << sh) - 1) >> sh) * 64); }
// The update kernel strides over its work with at most kVectorBlocks workgroups.
constexpr int kVectorBlocks = 512;
int update_grid(int nrows) {
  const int g = grid_for((long)((nrows + 20) / 21) * 64);
  return g < kVectorBlocks ? g : kVectorBlocks;
}


template <typename T>
hipError_t launch_rhs_init(const RhsArgs& a, hipStream_t s) {
  hipLaunchKernelGGL((rhs_init_kernel<T>), dim3(rhs_grid(a.nrows)), dim3(kBlock), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_cg_init(const InitArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(cg_init_kernel, dim3(1), dim3(kBlock), 0, s, a);
  return hipGetLastError();
}

// With events, the launch carries them in its own dispatch (kernel start and end), so the
// measured interval is the kernel's execution, not the queue's.
template <typename T>
hipError_t launch_cg_spmv(const CgArgs& a, hipStream_t s, hipEvent_t start, hipEvent_t stop) {
  const dim3 g(spmv_grid(a.nrows, a.slice_shift)), b(kBlock);
#define SPMV_LAUNCH(H)                                                                \
  if (start)                                                                          \
    hipExtLaunchKernelGGL((cg_spmv_kernel<T, H>), g, b, 0, s, start, stop, 0, a);     \
  else                                                                                \
    hipLaunchKernelGGL((cg_spmv_kernel<T, H>), g, b, 0, s, a);
  switch (a.slice_shift) {
    case 3: SPMV_LAUNCH(8) break;
    case 4: SPMV_LAUNCH(16) break;
    case 5: SPMV_LAUNCH(32) break;
    case 6: SPMV_LAUNCH(64) break;
    default: return hipErrorInvalidValue;
  }
#undef SPMV_LAUNCH
  return hipGetLastError();
}

// The pt / rz partial arrays hold spmv_grid / update_grid entries.
template <typename T>
hipError_t launch_cg_rest(const CgArgs& a, hipStream_t s) {
  hipLaunchKernelGGL((cg_update_kernel<T>), dim3(update_grid(a.nrows)), dim3(kBlock), 0, s, a);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_cg_iteration(const CgArgs& a, hipStream_t s) {
  hipError_t e = launch_cg_spmv<T>(a, s);
  if (e != hipSuccess) return e;
  return launch_cg_rest<T>(a, s);
}

hipError_t launch_refine_begin(const RefineArgs& a, const Partials& rhs_parts, hipStream_t s) {
  const long g = grid_for(3L * a.nrows);
  hipLaunchKernelGGL(refine_begin_kernel, dim3(g < kVectorBlocks ? g : kVectorBlocks), dim3(kBlock), 0,
                     s, a, rhs_parts);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_refine_init(const RefineArgs& a, hipStream_t s) {
  hipLaunchKernelGGL((refine_init_kernel<T>), dim3(rhs_grid(a.nrows)), dim3(kBlock), 0, s, a);
  return hipGetLastError();
}

// residual: accumulate-free SpMV on x64 + x, then x64 += x; partials go to a.out (spmv grid)
template <typename T>
hipError_t launch_refine_residual(const RefineArgs& a, hipStream_t s) {
  const dim3 g(spmv_grid(a.nrows, a.slice_shift)), b(kBlock);
  switch (a.slice_shift) {
    case 3: hipLaunchKernelGGL((refine_residual_kernel<T, 8>), g, b, 0, s, a); break;
    case 4: hipLaunchKernelGGL((refine_residual_kernel<T, 16>), g, b, 0, s, a); break;
    case 5: hipLaunchKernelGGL((refine_residual_kernel<T, 32>), g, b, 0, s, a); break;
    case 6: hipLaunchKernelGGL((refine_residual_kernel<T, 64>), g, b, 0, s, a); break;
    default: return hipErrorInvalidValue;
  }
  hipLaunchKernelGGL((refine_accumulate_kernel<T>), dim3(grid_for(3L * a.nrows)), dim3(kBlock), 0, s,
                     3 * a.nrows, a.x64, (const T*)a.x);
  return hipGetLastError();
}

hipError_t launch_refine_check(const RefineArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(refine_check_kernel, dim3(1), dim3(kBlock), 0, s, a);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_force_gather(int nv, const int* vptr, const T* stage_f, double* R, hipStream_t s) {
  hipLaunchKernelGGL((force_gather_kernel<T>), dim3(grid_for(nv)), dim3(kBlock), 0, s, nv, vptr,
                     stage_f, R);
  return hipGetLastError();
}

hipError_t launch_copy(const void* src, void* dst, size_t bytes, hipStream_t s) {
  const size_t n = bytes / sizeof(double4);
  hipLaunchKernelGGL(copy_kernel, dim3(4096), dim3(kBlock), 0, s, (const double4*)src, (double4*)dst, n);
  return hipGetLastError();
}

#define INST(T)                                                                          \
  template hipError_t launch_apply_dx<T>(int, double*, const T*, hipStream_t);           \
  template hipError_t launch_assemble<T>(const AsmArgs&, hipStream_t);                   \
  template hipError_t launch_constant_parts<T>(const ConstArgs&, hipStream_t);           \
  template hipError_t launch_rhs_init<T>(const RhsArgs&, hipStream_t);                   \
  template hipError_t launch_cg_iteration<T>(const CgArgs&, hipStream_t);                \
  template hipError_t launch_cg_spmv<T>(const CgArgs&, hipStream_t, hipEvent_t, hipEvent_t); \
  template hipError_t launch_cg_rest<T>(const CgArgs&, hipStream_t);                     \
  template hipError_t launch_force_gather<T>(int, const int*, const T*, double*, hipStream_t); \
  template hipError_t launch_refine_init<T>(const RefineArgs&, hipStream_t);                 \
  template hipError_t launch_refine_residual<T>(const RefineArgs&, hipStream_t);
INST(double)
INST(float)
#undef INST

}  // namespace hobak
