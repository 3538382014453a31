// Assembly, right-hand side, PCG, refinement and state-update kernels of the backward-Euler
// Newton step, all segmented over a batch of instances (see kernels.h).
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
// Every reduction is deterministic and fence-free: each workgroup of a producing kernel writes
// its partial sum to a fixed slot of its instance's partial array, and every workgroup of the
// consuming (next) kernel re-reduces that array in the same fixed order.  There are no
// floating-point atomics anywhere.  The PCG kernels of an instance whose solve is done (or not
// active) return at once, which is the per-instance convergence mask of a batched solve.
#include <hip/hip_ext.h>

#include <cfloat>

#include "kernels.h"

namespace hobak {

// ---------------------------------------------------------------------------------------------
// reductions
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Sum of each value over the workgroup, returned in every lane.
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

// Publish this workgroup's sums of (a, b) as entry lb of the partial arrays (pb may be null).
__device__ __forceinline__ void publish(double a, double b, double* pa, double* pb, int lb) {
  double v[2] = {a, b};
  block_sum<2>(v);
  if (threadIdx.x == 0) {
    pa[lb] = v[0];
    if (pb) pb[lb] = v[1];
  }
}

// Every consuming workgroup reduces the producer's n partials in the same fixed order, so all of
// them (and every run) obtain bit-identical totals.
__device__ __forceinline__ void gather(const double* pa, const double* pb, int n, double& ta, double& tb) {
  double v[2] = {0.0, 0.0};
  for (int i = threadIdx.x; i < n; i += kBlock) {
    v[0] += pa[i];
    if (pb) v[1] += pb[i];
  }
  block_sum<2>(v);
  ta = v[0];
  tb = v[1];
}

__device__ __forceinline__ bool nonfinite(double v) { return !(v == v) || v > DBL_MAX || v < -DBL_MAX; }

// An instance whose current inner solve has ended and has not yet been handled.
__device__ __forceinline__ bool solve_ended(const Ctl* c) {
  return c->phase == PHASE_SOLVING && c->sc[0].done;
}

static inline int grid_of(const Batch& b, int cap) { return cap > 0 && cap < b.total ? cap : b.total; }

// ---------------------------------------------------------------------------------------------
// step bookkeeping
// ---------------------------------------------------------------------------------------------
// _positionOld = _position; _velocityOld = _velocity; applyKinematicConstraints
// (BACKWARD_EULER_POSITION.cpp:152-157, TIMESTEPPER.cpp:105-122)
__global__ __launch_bounds__(kBlock) void begin_step_kernel(Batch b) {
  for (int vb = blockIdx.x; vb < b.total; vb += gridDim.x) {
    int lb, lg;
    const InstDev* __restrict__ I = seg_find(b, vb, lb, lg);
    const int r = lb * kBlock + threadIdx.x;
    if (r < I->nv) {
      double* u = I->u;
      const double* v = I->v;
      double* uo = I->u_old;
      double* vo = I->v_old;
      const bool pin = I->pinned[r];
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const size_t i = 3 * (size_t)r + c;
        uo[i] = u[i];
        vo[i] = v[i];
        if (pin) u[i] = I->pin_target[i];
      }
    }
  }
}

// _position += xDelta (BACKWARD_EULER_POSITION.cpp:262-266)
template <typename T>
__global__ __launch_bounds__(kBlock) void apply_dx_kernel(Batch b, int use64) {
  for (int vb = blockIdx.x; vb < b.total; vb += gridDim.x) {
    int lb, lg;
    const InstDev* __restrict__ I = seg_find(b, vb, lb, lg);
    const int i = lb * kBlock + threadIdx.x;
    if (i < 3 * I->nv) I->u[i] += use64 ? I->x64[i] : (double)((const T*)I->x)[i];
  }
}

// _velocity = (u - u_old)/dt, then S-filtered (BACKWARD_EULER_POSITION.cpp:300-314)
__global__ __launch_bounds__(kBlock) void end_step_kernel(Batch b) {
  for (int vb = blockIdx.x; vb < b.total; vb += gridDim.x) {
    int lb, lg;
    const InstDev* __restrict__ I = seg_find(b, vb, lb, lg);
    const int r = lb * kBlock + threadIdx.x;
    if (r < I->nv) {
      const double invDt = I->invDt;
      const bool pin = I->pinned[r];
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const size_t i = 3 * (size_t)r + c;
        I->v[i] = pin ? 0.0 : invDt * (I->u[i] - I->u_old[i]);
      }
    }
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

// One lane per block slot.  ASM_REST / ASM_PLAIN_A: out = sum of staged contributions (the
// stiffness, -K) into kneg / A.  ASM_SYSTEM: A = S (D - K) S + (I - S) (+ the ground penalty on
// the diagonal); the diagonal slot also writes its preconditioner.
template <typename T>
__global__ __launch_bounds__(kBlock) void assemble_kernel(Batch b, int target) {
  for (int vb = blockIdx.x; vb < b.total; vb += gridDim.x) {
    int lb, lg;
    const InstDev* __restrict__ I = seg_find(b, vb, lb, lg);
    const int slot = lb * kBlock + threadIdx.x;
    const int nslots = I->nslots;
    if (slot >= nslots) continue;
    const int r = I->slot_row[slot];
    if (r < 0) continue;
    const int c = I->col[slot];
    const size_t ns = (size_t)nslots;
    T acc[9];
#pragma unroll
    for (int e = 0; e < 9; e++) acc[e] = T(0);
    // the slot's contributions are contiguous and in ascending tet order
    const T* src = (const T*)I->stage_h + 9 * (size_t)I->cptr[slot];
    const T* end = (const T*)I->stage_h + 9 * (size_t)I->cptr[slot + 1];
    for (; src < end; src += 9) {
#pragma unroll
      for (int e = 0; e < 9; e++) acc[e] += src[e];
    }
    if (target != ASM_SYSTEM) {
      T* out = (T*)(target == ASM_REST ? I->kneg : I->A);
#pragma unroll
      for (int e = 0; e < 9; e++) out[e * ns + slot] = acc[e];
      continue;
    }
    constexpr bool kDiff = sizeof(T) < sizeof(double);
    // standard form: pinned rows and columns are zeroed; difference form: only pinned rows
    const bool pin = kDiff ? (bool)I->pinned[r] : (I->pinned[r] || I->pinned[c]);
    bool ground = false;
    if (I->ground_on && c == r && !pin) {
      const double y = I->rest[3 * (size_t)r + 1] + I->u[3 * (size_t)r + 1];
      ground = y < I->ground_y;
      if (ground) acc[4] += (T)I->ground_k;
    }
    const T* D = (const T*)I->D;
    T* out = (T*)I->A;
    T val[9];
#pragma unroll
    for (int e = 0; e < 9; e++) {
      val[e] = pin ? ((c == r && (e % 4) == 0) ? T(1) : T(0)) : D[e * ns + slot] + acc[e];
      out[e * ns + slot] = val[e];
    }
    if (c == r) {
      write_precond<T>(val, (T*)I->minv, r, I->precond);
      if (kDiff) {
#pragma unroll
        for (int k = 0; k < 3; k++) {
          const double mk = I->mass[3 * (size_t)r + k] * (I->invDt2 - I->ra * I->invDt);
          I->dg[3 * (size_t)r + k] = pin ? 1.0 : mk + ((k == 1 && ground) ? I->ground_k : 0.0);
        }
      }
    }
  }
}

// D = M/dt^2 - C/dt, E = M/dt - C, C = alpha M + beta K_rest, from the assembled rest stiffness
// (-K_rest in kneg).  Order of operations as BACKWARD_EULER_POSITION.cpp:198, 208 and
// TIMESTEPPER.cpp:232-233.
template <typename T>
__global__ __launch_bounds__(kBlock) void constant_parts_kernel(Batch b) {
  for (int vb = blockIdx.x; vb < b.total; vb += gridDim.x) {
    int lb, lg;
    const InstDev* __restrict__ I = seg_find(b, vb, lb, lg);
    const int slot = lb * kBlock + threadIdx.x;
    const int nslots = I->nslots;
    if (slot >= nslots) continue;
    const int r = I->slot_row[slot];
    if (r < 0) continue;
    const int c = I->col[slot];
    const size_t ns = (size_t)nslots;
    const double ra = I->ra, rb = I->rb, invDt = I->invDt, invDt2 = I->invDt2;
#pragma unroll
    for (int e = 0; e < 9; e++) {
      const double krest = -(double)((const T*)I->kneg)[e * ns + slot];
      const bool dg = (c == r) && (e % 4 == 0);
      const double m = dg ? I->mass[3 * (size_t)r + e / 4] : 0.0;
      const double cc = dg ? ra * m + rb * krest : rb * krest;
      const double d = dg ? m * invDt2 - cc * invDt : -(cc * invDt);
      const double ee = dg ? invDt * m - cc : -cc;
      ((T*)I->D)[e * ns + slot] = (T)d;
      ((T*)I->E)[e * ns + slot] = (T)ee;
      ((T*)I->C)[e * ns + slot] = (T)cc;
    }
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
// b = C vb + ((M a + R) + f_ext) (BDF_1); rhs = S b.  With init_cg also x = 0, r = rhs,
// z = M^-1 r; the partial sums of |rhs|^2 and r.z seed Eigen's CG (ConjugateGradient.h:43-66).
template <typename T>
__global__ __launch_bounds__(kBlock) void rhs_kernel(Batch b, int init_cg) {
  for (int vb = blockIdx.x; vb < b.total; vb += gridDim.x) {
    int lb, lg;
    const InstDev* __restrict__ I = seg_find(b, vb, lb, lg);
    const int r = lb * kBlock + threadIdx.x;
    double part[2] = {0.0, 0.0};
    if (r < I->nv) {
      const T* sf = (const T*)I->stage_f;
      const int* vptr = I->vptr;
      double R[3] = {0.0, 0.0, 0.0};
      for (int q = vptr[r]; q < vptr[r + 1]; q++) {
#pragma unroll
        for (int c = 0; c < 3; c++) R[c] += (double)sf[3 * (size_t)q + c];
      }
      const double* u = I->u;
      if (I->ground_on) {
        const double y = I->rest[3 * (size_t)r + 1] + u[3 * (size_t)r + 1];
        if (y < I->ground_y) R[1] += I->ground_k * (I->ground_y - y);
      }
      const int bdf = I->bdf;
      const double invDt = I->invDt, invDt2 = I->invDt2;
      const double* uOld = I->u_old;
      const double* vOld = I->v_old;
      const double* vel = I->v;
      const size_t ns = (size_t)I->nslots;
      const T* Mx = (const T*)(bdf ? I->C : I->E);
      const int sh = I->shift;
      const int* col = I->col;
      constexpr bool kDiff = sizeof(T) < sizeof(double);
      // the row's own velocity (difference form: E v = e_m v_i + sum_j E_ij (v_j - v_i), with
      // e_m = m (1/dt - alpha) for E and m alpha for C)
      double own[3] = {0.0, 0.0, 0.0};
      double s[3] = {0.0, 0.0, 0.0};
      if (kDiff) {
        const double em = bdf ? I->ra : invDt - I->ra;
#pragma unroll
        for (int j = 0; j < 3; j++) {
          const size_t i = 3 * (size_t)r + j;
          own[j] = bdf ? invDt * (u[i] - uOld[i]) - vOld[i] : vel[i];
          s[j] = I->mass[i] * em * own[j];
        }
      }
      const int len = I->row_len[r];
      const int base = I->slice_off[r >> sh] + (r & ((1 << sh) - 1));
      for (int k = 0; k < len; k++) {
        const int slot = base + (k << sh);
        const int c = col[slot];
        double xv[3];
#pragma unroll
        for (int j = 0; j < 3; j++) {
          const size_t i = 3 * (size_t)c + j;
          xv[j] = (bdf ? invDt * (u[i] - uOld[i]) - vOld[i] : vel[i]) - own[j];
        }
#pragma unroll
        for (int i = 0; i < 3; i++) {
          s[i] += (double)Mx[(3 * i + 0) * ns + slot] * xv[0];
          s[i] += (double)Mx[(3 * i + 1) * ns + slot] * xv[1];
          s[i] += (double)Mx[(3 * i + 2) * ns + slot] * xv[2];
        }
      }
      const bool pin = I->pinned[r];
      const double orv = I->one_ring[r];
      double bv[3];
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const size_t i = 3 * (size_t)r + c;
        const double fext = orv * I->g[c];
        if (bdf) {
          const double acc = -invDt2 * (u[i] - uOld[i]) + invDt * vOld[i];
          bv[c] = s[c] + (((I->mass[i] * acc + R[c]) + fext));
        } else {
          bv[c] = (s[c] + R[c]) + fext;
        }
        if (pin) bv[c] = 0.0;
        I->rhs64[i] = bv[c];
        part[0] += bv[c] * bv[c];
      }
      if (init_cg) {
        T rv[3], zv[3];
#pragma unroll
        for (int c = 0; c < 3; c++) rv[c] = (T)bv[c];
        apply_precond<T>((const T*)I->minv, I->precond, r, rv, zv);
#pragma unroll
        for (int c = 0; c < 3; c++) {
          const size_t i = 3 * (size_t)r + c;
          ((T*)I->r)[i] = rv[c];
          ((T*)I->x)[i] = T(0);
          ((T*)I->z)[i] = zv[c];
          part[1] += (double)rv[c] * (double)zv[c];
        }
      }
    }
    publish(part[0], part[1], I->part_row_a, I->part_row_b, lb);
  }
}

// sc[0] from the seeding kernel's partials (the right-hand side kernel, or the refinement's
// inner-solve seed): Eigen's early exits on a zero or already small right-hand side
// (ConjugateGradient.h:45-61) and the first r.z.  With gate, only instances whose inner solve
// has just ended and that still need another refinement sweep are (re)started.
__global__ __launch_bounds__(kBlock) void cg_init_kernel(Batch b, int refine, int gate) {
  for (int vb = blockIdx.x; vb < b.total; vb += gridDim.x) {
    int lb, lg;
    const InstDev* __restrict__ I = seg_find(b, vb, lb, lg);
    Ctl* c = I->ctl;
    if (gate && !solve_ended(c)) continue;
    double bb, rz;
    gather(I->part_row_a, I->part_row_b, I->nb[K_ROW], bb, rz);
    if (threadIdx.x != 0) continue;
    const double tol = refine ? c->rs.tau : I->tol;
    const int rdone = refine ? c->rs.done : 0;
    const double thr0 = tol * tol * bb;
    const double thr = thr0 > DBL_MIN ? thr0 : DBL_MIN;
    const int max_iters = I->max_iters;
    int done = (bb == 0.0 || bb < thr || max_iters <= 0 || rdone) ? 1 : 0;
    int status = 0;
    if (nonfinite(bb)) {
      done = 1;
      status = 2;
    }
    CgScalars* s = &c->sc[0];
    s->maxR = refine ? c->rs.b2 : bb;
    s->rhsNorm2 = bb;
    s->resNorm2 = bb;
    s->threshold = thr;
    s->absNew = rz;
    s->alpha = 0.0;
    s->beta = 0.0;
    s->iters = 0;
    s->maxIters = max_iters;
    s->done = done;
    s->status = status;
    s->first = 1;
    c->phase = PHASE_SOLVING;
  }
}

// ---------------------------------------------------------------------------------------------
// PCG iteration of parity p: two kernels (Eigen's loop body, ConjugateGradient.h:69-88, regrouped)
//   spmv(p):    S = sc[p].  Unless S is the first set of the solve, it also finishes the
//               previous iteration: from the update kernel's partials, |r|^2 < threshold -> done
//               (the converging pass is not counted); else beta = r.z / S.absNew, absNew = r.z,
//               iters + 1 (done at maxIters).  The instance's workgroup 0 writes that as the next
//               set sc[p ^ 1] (a finished set is carried forward).  The product t = A p_new with
//               p_new = z + beta p_old is formed as A z + beta (A p_old): the matrix streams
//               while the partials are still being reduced, and beta is applied at the end.
//               p_new is written for the kernel's own rows (p double-buffered by parity), and the
//               partials of p.t are published.
//   update(p):  S' = sc[p ^ 1]; alpha = S'.absNew / (p.t); x += alpha p_new; r -= alpha t;
//               z = M^-1 r; partials of r.r and r.z.
// Kernels of a finished or inactive solve return at once, so groups of iterations can be queued
// (or replayed from a captured graph) ahead and the host polls between groups.  Scalars are read
// and written field by field: a by-value copy of the struct becomes a per-lane LDS buffer.
// ---------------------------------------------------------------------------------------------
// One wavefront per H-row slice, L = 64 / H lanes per row: lane = H * kk + rr handles the row's
// blocks k = kk, kk + L, ...; a row's L partial sums are combined with cross-lane adds.
template <typename T, int H>
__global__ __launch_bounds__(kBlock) void cg_spmv_kernel(Batch b, int parity) {
  for (int vb = blockIdx.x; vb < b.total; vb += gridDim.x) {
    int lb, lg;
    const InstDev* __restrict__ I = seg_find(b, vb, lb, lg);
    Ctl* ctl = I->ctl;
    const CgScalars* S = &ctl->sc[parity];
    CgScalars* N = &ctl->sc[parity ^ 1];
    const bool lead = lb == 0 && threadIdx.x == 0;
    const int done = S->done;
    const int first = S->first;
    const double absOld = S->absNew;
    const int iters = S->iters;
    const double threshold = S->threshold;
    const int maxIters = S->maxIters;
    if (lead) {
      N->rhsNorm2 = S->rhsNorm2;
      N->threshold = threshold;
      N->absNew = absOld;
      N->alpha = S->alpha;
      N->beta = S->beta;
      N->resNorm2 = S->resNorm2;
      N->maxR = S->maxR;
      N->iters = iters;
      N->maxIters = maxIters;
      N->done = done;
      N->status = S->status;
      N->first = 0;
    }
    if (done) continue;
    constexpr int L = 64 / H;
    const int slice = (int)((lb * (unsigned)kBlock + threadIdx.x) >> 6);
    const int lane = threadIdx.x & 63;
    const int kk = lane / H;
    const int rr = lane % H;
    const int r = slice * H + rr;
    const T* pold = (const T*)(parity ? I->p1 : I->p0);
    T* pnew = (T*)(parity ? I->p0 : I->p1);
    const T* z = (const T*)I->z;
    const T* A = (const T*)I->A;
    const int* col = I->col;
    const size_t ns = (size_t)I->nslots;
    const int nrows = I->nv;
    // a0..a2 = (A z)_r, c0..c2 = (A p_old)_r; in the difference form (T = float) the column
    // values are taken relative to the row's own (zo, qo) and dg adds the diagonal term
    constexpr bool kDiff = sizeof(T) < sizeof(double);
    T a0 = T(0), a1 = T(0), a2 = T(0), c0 = T(0), c1 = T(0), c2 = T(0);
    T zo0 = T(0), zo1 = T(0), zo2 = T(0), qo0 = T(0), qo1 = T(0), qo2 = T(0);
    const bool live = r < nrows;
    if (live) {
      const int len = I->row_len[r];
      const int base = I->slice_off[slice] + rr;
      if (kDiff) {
        const size_t i = 3 * (size_t)r;
        zo0 = z[i];
        zo1 = z[i + 1];
        zo2 = z[i + 2];
        if (!first) {
          qo0 = pold[i];
          qo1 = pold[i + 1];
          qo2 = pold[i + 2];
        }
      }
      if (first) {
        for (int k = kk; k < len; k += L) {
          const int slot = base + k * H;
          const size_t c = 3 * (size_t)col[slot];
          const T z0 = z[c] - zo0, z1 = z[c + 1] - zo1, z2 = z[c + 2] - zo2;
          a0 += A[0 * ns + slot] * z0 + A[1 * ns + slot] * z1 + A[2 * ns + slot] * z2;
          a1 += A[3 * ns + slot] * z0 + A[4 * ns + slot] * z1 + A[5 * ns + slot] * z2;
          a2 += A[6 * ns + slot] * z0 + A[7 * ns + slot] * z1 + A[8 * ns + slot] * z2;
        }
      } else {
        for (int k = kk; k < len; k += L) {
          const int slot = base + k * H;
          const size_t c = 3 * (size_t)col[slot];
          const T z0 = z[c] - zo0, z1 = z[c + 1] - zo1, z2 = z[c + 2] - zo2;
          const T q0 = pold[c] - qo0, q1 = pold[c + 1] - qo1, q2 = pold[c + 2] - qo2;
          const T m0 = A[0 * ns + slot], m1 = A[1 * ns + slot], m2 = A[2 * ns + slot];
          const T m3 = A[3 * ns + slot], m4 = A[4 * ns + slot], m5 = A[5 * ns + slot];
          const T m6 = A[6 * ns + slot], m7 = A[7 * ns + slot], m8 = A[8 * ns + slot];
          a0 += m0 * z0 + m1 * z1 + m2 * z2;
          a1 += m3 * z0 + m4 * z1 + m5 * z2;
          a2 += m6 * z0 + m7 * z1 + m8 * z2;
          c0 += m0 * q0 + m1 * q1 + m2 * q2;
          c1 += m3 * q0 + m4 * q1 + m5 * q2;
          c2 += m6 * q0 + m7 * q1 + m8 * q2;
        }
      }
    }
#pragma unroll
    for (int off = H; off < 64; off <<= 1) {
      a0 += __shfl_xor(a0, off, 64);
      a1 += __shfl_xor(a1, off, 64);
      a2 += __shfl_xor(a2, off, 64);
      if (!first) {
        c0 += __shfl_xor(c0, off, 64);
        c1 += __shfl_xor(c1, off, 64);
        c2 += __shfl_xor(c2, off, 64);
      }
    }
    T beta = T(0);
    if (!first) {
      double rrv, rz;
      gather(I->part_rr, I->part_rz, I->nb[K_UPD], rrv, rz);
      if (nonfinite(rrv) || !(rz == rz)) {
        if (lead) {
          N->resNorm2 = rrv;
          N->done = 1;
          N->status = 2;
        }
        continue;
      }
      if (rrv < threshold) {
        if (lead) {
          N->resNorm2 = rrv;
          N->done = 1;
        }
        continue;
      }
      const double bt = rz / absOld;
      if (lead) {
        N->resNorm2 = rrv;
        N->beta = bt;
        N->absNew = rz;
        N->iters = iters + 1;
        if (iters + 1 >= maxIters) {
          N->done = 1;
          N->status = 1;
        }
      }
      if (iters + 1 >= maxIters) continue;
      beta = (T)bt;
    }
    double pt = 0.0;
    if (live && kk == 0) {
      const size_t i = 3 * (size_t)r;
      if (kDiff) {
        const T d0 = (T)I->dg[i], d1 = (T)I->dg[i + 1], d2 = (T)I->dg[i + 2];
        a0 += d0 * zo0;
        a1 += d1 * zo1;
        a2 += d2 * zo2;
        c0 += d0 * qo0;
        c1 += d1 * qo1;
        c2 += d2 * qo2;
      }
      const T s0 = a0 + beta * c0, s1 = a1 + beta * c1, s2 = a2 + beta * c2;
      T* t = (T*)I->t;
      t[i] = s0;
      t[i + 1] = s1;
      t[i + 2] = s2;
      T q0 = z[i], q1 = z[i + 1], q2 = z[i + 2];
      if (!first) {
        q0 += beta * pold[i];
        q1 += beta * pold[i + 1];
        q2 += beta * pold[i + 2];
      }
      pnew[i] = q0;
      pnew[i + 1] = q1;
      pnew[i + 2] = q2;
      pt = (double)q0 * (double)s0 + (double)q1 * (double)s1 + (double)q2 * (double)s2;
    }
    publish(pt, 0.0, I->part_pt, nullptr, lb);
  }
}

// One lane per DOF, 21 rows (63 lanes) per wavefront so that a row's three DOFs share a wave and
// the 3x3 preconditioner apply can read its siblings' residuals with cross-lane moves.  Each
// instance's workgroups stride over its waves (at most kUpdateCap workgroups per instance), which
// keeps the partial-sum count (and the work of every workgroup that re-reduces them) small.  The
// first wave's vectors are loaded before the p.t partials are reduced, so the two overlap.
template <typename T>
__global__ __launch_bounds__(kBlock) void cg_update_kernel(Batch b, int parity) {
  for (int vb = blockIdx.x; vb < b.total; vb += gridDim.x) {
    int lb, lg;
    const InstDev* __restrict__ I = seg_find(b, vb, lb, lg);
    const CgScalars* S = &I->ctl->sc[parity ^ 1];
    if (S->done) continue;
    const double absNew = S->absNew;
    const int lane = threadIdx.x & 63;
    const int c = lane - 3 * (lane / 3);
    const int bse = lane - c;
    const int nrows = I->nv;
    const int nwaves = (nrows + 20) / 21;
    const int stride = lg * (kBlock / 64);
    const T* p = (const T*)(parity ? I->p0 : I->p1);  // the p_new this iteration's SpMV wrote
    const T* mi = (const T*)I->minv;
    T* x = (T*)I->x;
    T* rv_ = (T*)I->r;
    T* z = (T*)I->z;
    const T* t = (const T*)I->t;
    const int precond = I->precond;
    const int w0 = (int)((lb * (unsigned)kBlock + threadIdx.x) >> 6);
    T pv = T(0), tv = T(0), rv0 = T(0), xv = T(0);
    {
      const int row = w0 * 21 + lane / 3;
      if (w0 < nwaves && lane < 63 && row < nrows) {
        const size_t i = 3 * (size_t)row + c;
        pv = p[i];
        tv = t[i];
        rv0 = rv_[i];
        xv = x[i];
      }
    }
    double pt, unused;
    gather(I->part_pt, nullptr, I->nb[K_SPMV], pt, unused);
    const T alpha = (T)(absNew / pt);
    double rr2 = 0.0, rz = 0.0;
    for (int wave = w0; wave < nwaves; wave += stride) {
      const int row = wave * 21 + lane / 3;
      const bool live = lane < 63 && row < nrows;
      const size_t i = 3 * (size_t)row + c;
      if (wave != w0 && live) {
        pv = p[i];
        tv = t[i];
        rv0 = rv_[i];
        xv = x[i];
      }
      T rv = T(0);
      if (live) {
        x[i] = xv + alpha * pv;
        rv = rv0 - alpha * tv;
        rv_[i] = rv;
      }
      const T r0 = __shfl(rv, bse, 64);
      const T r1 = __shfl(rv, bse + 1, 64);
      const T r2 = __shfl(rv, bse + 2, 64);
      if (live) {
        T zv;
        if (precond == 0) {
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
    publish(rr2, rz, I->part_rr, I->part_rz, lb);
  }
}

// ---------------------------------------------------------------------------------------------
// FP64 iterative refinement of the reduced-precision solve
//   begin:     b2 = |S b|^2; x64 = 0; res64 = S b; tau = first inner tolerance
//   init:      r = (T) res64, x = 0, z = M^-1 r -> the inner PCG (same kernels as above)
//   residual:  res64 = rhs64 - A (x64 + x) with A's T values promoted to FP64; then x64 += x
//   check:     res2 = |res64|^2; done when res2 < tol^2 b2; else the next inner tolerance
// The residual / check / re-init kernels are gated per instance, so they can follow every group
// of PCG iterations of a batch and act only on instances whose inner solve has just ended.
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void refine_begin_kernel(Batch b) {
  for (int vb = blockIdx.x; vb < b.total; vb += gridDim.x) {
    int lb, lg;
    const InstDev* __restrict__ I = seg_find(b, vb, lb, lg);
    const int r = lb * kBlock + threadIdx.x;
    if (r < I->nv) {
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const size_t i = 3 * (size_t)r + c;
        I->x64[i] = 0.0;
        I->res64[i] = I->rhs64[i];
      }
    }
    if (lb != 0) continue;
    double b2, unused;
    gather(I->part_row_a, nullptr, I->nb[K_ROW], b2, unused);
    if (threadIdx.x == 0) {
      RefineState* st = &I->ctl->rs;
      const int bad = nonfinite(b2) ? 1 : 0;
      st->b2 = b2;
      st->res2 = b2;
      st->tau = I->tau0;
      st->sweeps = 0;
      st->inner_iters = 0;
      st->status = bad ? 2 : 0;
      st->done = (b2 == 0.0 || bad) ? 1 : 0;
    }
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void refine_init_kernel(Batch b, int gate) {
  for (int vb = blockIdx.x; vb < b.total; vb += gridDim.x) {
    int lb, lg;
    const InstDev* __restrict__ I = seg_find(b, vb, lb, lg);
    if (gate && !solve_ended(I->ctl)) continue;
    const int r = lb * kBlock + threadIdx.x;
    double rr = 0.0, rz = 0.0;
    if (r < I->nv) {
      T rv[3], zv[3];
#pragma unroll
      for (int c = 0; c < 3; c++) rv[c] = (T)I->res64[3 * (size_t)r + c];
      apply_precond<T>((const T*)I->minv, I->precond, r, rv, zv);
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const size_t i = 3 * (size_t)r + c;
        ((T*)I->r)[i] = rv[c];
        ((T*)I->x)[i] = T(0);
        ((T*)I->z)[i] = zv[c];
        rr += (double)rv[c] * (double)rv[c];
        rz += (double)rv[c] * (double)zv[c];
      }
    }
    publish(rr, rz, I->part_row_a, I->part_row_b, lb);
  }
}

// One wavefront per H-row slice as in the SpMV; x64 + x is formed on the fly for every column.
template <typename T, int H>
__global__ __launch_bounds__(kBlock) void refine_residual_kernel(Batch b) {
  for (int vb = blockIdx.x; vb < b.total; vb += gridDim.x) {
    int lb, lg;
    const InstDev* __restrict__ I = seg_find(b, vb, lb, lg);
    const Ctl* ctl = I->ctl;
    if (!solve_ended(ctl) || ctl->rs.done) continue;
    constexpr int L = 64 / H;
    const int slice = (int)((lb * (unsigned)kBlock + threadIdx.x) >> 6);
    const int lane = threadIdx.x & 63;
    const int kk = lane / H;
    const int rr = lane % H;
    const int r = slice * H + rr;
    const T* A = (const T*)I->A;
    const T* xi = (const T*)I->x;
    const double* x64 = I->x64;
    const int* col = I->col;
    const size_t ns = (size_t)I->nslots;
    constexpr bool kDiff = sizeof(T) < sizeof(double);
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    double o0 = 0.0, o1 = 0.0, o2 = 0.0;  // the row's own x64 + x (difference form)
    const bool live = r < I->nv;
    if (live) {
      const int len = I->row_len[r];
      const int base = I->slice_off[slice] + rr;
      if (kDiff) {
        const size_t i = 3 * (size_t)r;
        o0 = x64[i] + (double)xi[i];
        o1 = x64[i + 1] + (double)xi[i + 1];
        o2 = x64[i + 2] + (double)xi[i + 2];
      }
      for (int k = kk; k < len; k += L) {
        const int slot = base + k * H;
        const size_t c = 3 * (size_t)col[slot];
        const double x0 = (x64[c] + (double)xi[c]) - o0;
        const double x1 = (x64[c + 1] + (double)xi[c + 1]) - o1;
        const double x2 = (x64[c + 2] + (double)xi[c + 2]) - o2;
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
      if (kDiff) {
        const size_t i = 3 * (size_t)r;
        s0 += I->dg[i] * o0;
        s1 += I->dg[i + 1] * o1;
        s2 += I->dg[i + 2] * o2;
      }
      const double s[3] = {s0, s1, s2};
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const size_t i = 3 * (size_t)r + c;
        const double res = I->rhs64[i] - s[c];
        I->res64[i] = res;
        res2 += res * res;
      }
    }
    publish(res2, 0.0, I->part_res, nullptr, lb);
  }
}

// x64 += x after every residual workgroup has read the old x64: a separate pass keeps the
// on-the-fly sums above race-free.
template <typename T>
__global__ __launch_bounds__(kBlock) void refine_accumulate_kernel(Batch b) {
  for (int vb = blockIdx.x; vb < b.total; vb += gridDim.x) {
    int lb, lg;
    const InstDev* __restrict__ I = seg_find(b, vb, lb, lg);
    const Ctl* ctl = I->ctl;
    if (!solve_ended(ctl) || ctl->rs.done) continue;
    const int i = lb * kBlock + threadIdx.x;
    if (i < 3 * I->nv) I->x64[i] += (double)((const T*)I->x)[i];
  }
}

// One workgroup per instance whose inner solve has just ended.  Plain form: the solve is
// finished.  Refinement: record the sweep and decide whether another one is needed; the next
// inner tolerance aims at the outer target with a 5x margin (0.2 tol |b| / |res|), floored at
// 1e-6 (FP32's reach) and capped at the first inner tolerance.
__global__ __launch_bounds__(kBlock) void finish_check_kernel(Batch b, int refine) {
  for (int vb = blockIdx.x; vb < b.total; vb += gridDim.x) {
    int lb, lg;
    const InstDev* __restrict__ I = seg_find(b, vb, lb, lg);
    Ctl* c = I->ctl;
    if (!solve_ended(c)) continue;
    if (!refine) {
      if (threadIdx.x == 0) c->phase = PHASE_FINISHED;
      continue;
    }
    if (c->rs.done) {  // nothing to solve (b == 0 or non-finite b)
      if (threadIdx.x == 0) c->phase = PHASE_FINISHED;
      continue;
    }
    double res2, unused;
    gather(I->part_res, nullptr, I->nb[K_SPMV], res2, unused);
    if (threadIdx.x != 0) continue;
    RefineState* st = &c->rs;
    const double b2 = st->b2;
    const double tol = I->tol;
    const int sweeps = st->sweeps + 1;
    st->res2 = res2;
    st->sweeps = sweeps;
    st->inner_iters += c->sc[0].iters;
    int done = 0;
    if (nonfinite(res2) || c->sc[0].status == 2) {
      done = 1;
      st->status = 2;
    } else if (res2 < tol * tol * b2) {
      done = 1;
    } else if (sweeps >= kMaxSweeps) {
      done = 1;
      st->status = 1;
    } else {
      const double want = 0.2 * tol * sqrt(b2 / res2);
      const double lo = want > 1e-6 ? want : 1e-6;
      st->tau = lo < I->tau0 ? lo : I->tau0;
    }
    st->done = done;
    if (done) c->phase = PHASE_FINISHED;
  }
}

// One lane per instance of the batch: what the host needs to schedule and report.
__global__ __launch_bounds__(kBlock) void status_kernel(Batch b, InstStatus* out) {
  for (int vb = blockIdx.x; vb < b.total; vb += gridDim.x) {
    int lb, lg;
    const InstDev* __restrict__ I = seg_find(b, vb, lb, lg);
    if (threadIdx.x != 0) continue;
    const Ctl* c = I->ctl;
    const CgScalars* s = &c->sc[0];
    InstStatus o;
    o.phase = c->phase;
    o.iters = s->iters;
    o.inner_total = c->rs.inner_iters;
    o.sweeps = c->rs.sweeps;
    o.cg_status = s->status;
    o.ref_status = c->rs.status;
    o.ref_done = c->rs.done;
    o.pad = 0;
    if (I->refine) {
      o.err = c->rs.b2 > 0 ? sqrt(c->rs.res2 / c->rs.b2) : 0.0;
      o.maxR = c->rs.b2;
    } else {
      o.err = s->rhsNorm2 > 0 ? sqrt(s->resNorm2 / s->rhsNorm2) : 0.0;
      o.maxR = s->maxR;
    }
    out[vb] = o;
  }
}

// computeHyperelasticForces' scatter as a per-vertex gather in ascending tet order
// (TET_MESH.cpp:756-767); output in FP64 into rhs64.
template <typename T>
__global__ __launch_bounds__(kBlock) void force_gather_kernel(Batch b) {
  for (int vb = blockIdx.x; vb < b.total; vb += gridDim.x) {
    int lb, lg;
    const InstDev* __restrict__ I = seg_find(b, vb, lb, lg);
    const int r = lb * kBlock + threadIdx.x;
    if (r >= I->nv) continue;
    const T* sf = (const T*)I->stage_f;
    double s[3] = {0.0, 0.0, 0.0};
    for (int q = I->vptr[r]; q < I->vptr[r + 1]; q++) {
#pragma unroll
      for (int c = 0; c < 3; c++) s[c] += (double)sf[3 * (size_t)q + c];
    }
#pragma unroll
    for (int c = 0; c < 3; c++) I->rhs64[3 * (size_t)r + c] = s[c];
  }
}

// Plain streaming copy, used only to calibrate hardware byte counters against a known size.
__global__ void copy_kernel(const double4* __restrict__ src, double4* __restrict__ dst, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// ---------------------------------------------------------------------------------------------
// workgroup counts and launchers
// ---------------------------------------------------------------------------------------------
static inline int cdiv(long a, long b) { return (int)((a + b - 1) / b); }

int blocks_of(int kind, int nv, int nt, int nslots, int shift) {
  switch (kind) {
    case K_ELEM: return cdiv(nt, kElementBlock);
    case K_SLOT: return cdiv(nslots, kBlock);
    case K_ROW: return cdiv(nv, kBlock);
    case K_DOF: return cdiv(3L * nv, kBlock);
    case K_SPMV: return cdiv((long)cdiv(nv, 1 << shift) * 64, kBlock);
    case K_UPD: {
      const int g = cdiv((long)cdiv(nv, 21) * 64, kBlock);
      return g < kUpdateCap ? g : kUpdateCap;
    }
    case K_ONE: return 1;
    default: return 0;
  }
}

#define SEG_LAUNCH(kernel, b, cap, s, ...)                                                  \
  do {                                                                                      \
    if ((b).total <= 0) return hipSuccess;                                                  \
    hipLaunchKernelGGL(kernel, dim3(grid_of((b), (cap))), dim3(kBlock), 0, (s), (b), ##__VA_ARGS__); \
    return hipGetLastError();                                                               \
  } while (0)

hipError_t launch_begin_step(const Batch& b, int cap, hipStream_t s) { SEG_LAUNCH(begin_step_kernel, b, cap, s); }
hipError_t launch_end_step(const Batch& b, int cap, hipStream_t s) { SEG_LAUNCH(end_step_kernel, b, cap, s); }

template <typename T>
hipError_t launch_apply_dx(const Batch& b, int use64, int cap, hipStream_t s) {
  SEG_LAUNCH((apply_dx_kernel<T>), b, cap, s, use64);
}
template <typename T>
hipError_t launch_assemble(const Batch& b, int target, int cap, hipStream_t s) {
  SEG_LAUNCH((assemble_kernel<T>), b, cap, s, target);
}
template <typename T>
hipError_t launch_constant_parts(const Batch& b, int cap, hipStream_t s) {
  SEG_LAUNCH((constant_parts_kernel<T>), b, cap, s);
}
template <typename T>
hipError_t launch_rhs(const Batch& b, int init_cg, int cap, hipStream_t s) {
  SEG_LAUNCH((rhs_kernel<T>), b, cap, s, init_cg);
}
hipError_t launch_cg_init(const Batch& b, int refine, int gate, hipStream_t s) {
  SEG_LAUNCH(cg_init_kernel, b, 0, s, refine, gate);
}

// With events, the launch carries them in its own dispatch (kernel start and end), so the
// measured interval is the kernel's execution, not the queue's.  All instances of a batch share
// one slice height (1 << shift rows).
template <typename T, int H>
static hipError_t spmv_h(const Batch& b, int parity, int cap, hipStream_t s, hipEvent_t start, hipEvent_t stop) {
  const dim3 g(grid_of(b, cap)), blk(kBlock);
  if (start)
    hipExtLaunchKernelGGL((cg_spmv_kernel<T, H>), g, blk, 0, s, start, stop, 0, b, parity);
  else
    hipLaunchKernelGGL((cg_spmv_kernel<T, H>), g, blk, 0, s, b, parity);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_cg_spmv(const Batch& b, int shift, int parity, int cap, hipStream_t s, hipEvent_t start,
                          hipEvent_t stop) {
  if (b.total <= 0) return hipSuccess;
  switch (shift) {
    case 3: return spmv_h<T, 8>(b, parity, cap, s, start, stop);
    case 4: return spmv_h<T, 16>(b, parity, cap, s, start, stop);
    case 5: return spmv_h<T, 32>(b, parity, cap, s, start, stop);
    case 6: return spmv_h<T, 64>(b, parity, cap, s, start, stop);
    default: return hipErrorInvalidValue;
  }
}

template <typename T>
hipError_t launch_cg_update(const Batch& b, int parity, int cap, hipStream_t s) {
  SEG_LAUNCH((cg_update_kernel<T>), b, cap, s, parity);
}
hipError_t launch_refine_begin(const Batch& b, int cap, hipStream_t s) { SEG_LAUNCH(refine_begin_kernel, b, cap, s); }
template <typename T>
hipError_t launch_refine_init(const Batch& b, int gate, int cap, hipStream_t s) {
  SEG_LAUNCH((refine_init_kernel<T>), b, cap, s, gate);
}
template <typename T>
hipError_t launch_refine_residual(const Batch& spmv, const Batch& dof, int shift, int cap, hipStream_t s) {
  if (spmv.total <= 0) return hipSuccess;
  const dim3 g(grid_of(spmv, cap)), blk(kBlock);
  switch (shift) {
    case 3: hipLaunchKernelGGL((refine_residual_kernel<T, 8>), g, blk, 0, s, spmv); break;
    case 4: hipLaunchKernelGGL((refine_residual_kernel<T, 16>), g, blk, 0, s, spmv); break;
    case 5: hipLaunchKernelGGL((refine_residual_kernel<T, 32>), g, blk, 0, s, spmv); break;
    case 6: hipLaunchKernelGGL((refine_residual_kernel<T, 64>), g, blk, 0, s, spmv); break;
    default: return hipErrorInvalidValue;
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  SEG_LAUNCH((refine_accumulate_kernel<T>), dof, cap, s);
}
hipError_t launch_finish_check(const Batch& b, int refine, hipStream_t s) {
  SEG_LAUNCH(finish_check_kernel, b, 0, s, refine);
}
hipError_t launch_status(const Batch& b, InstStatus* out, hipStream_t s) { SEG_LAUNCH(status_kernel, b, 0, s, out); }
template <typename T>
hipError_t launch_force_gather(const Batch& b, int cap, hipStream_t s) {
  SEG_LAUNCH((force_gather_kernel<T>), b, cap, s);
}

hipError_t launch_copy(const void* src, void* dst, size_t bytes, hipStream_t s) {
  const size_t n = bytes / sizeof(double4);
  hipLaunchKernelGGL(copy_kernel, dim3(4096), dim3(kBlock), 0, s, (const double4*)src, (double4*)dst, n);
  return hipGetLastError();
}

#define INST(T)                                                                              \
  template hipError_t launch_apply_dx<T>(const Batch&, int, int, hipStream_t);               \
  template hipError_t launch_assemble<T>(const Batch&, int, int, hipStream_t);               \
  template hipError_t launch_constant_parts<T>(const Batch&, int, hipStream_t);              \
  template hipError_t launch_rhs<T>(const Batch&, int, int, hipStream_t);                    \
  template hipError_t launch_cg_spmv<T>(const Batch&, int, int, int, hipStream_t, hipEvent_t, hipEvent_t); \
  template hipError_t launch_cg_update<T>(const Batch&, int, int, hipStream_t);              \
  template hipError_t launch_refine_init<T>(const Batch&, int, int, hipStream_t);            \
  template hipError_t launch_refine_residual<T>(const Batch&, const Batch&, int, int, hipStream_t); \
  template hipError_t launch_force_gather<T>(const Batch&, int, hipStream_t);
INST(double)
INST(float)
#undef INST

}  // namespace hobak
