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
// (butterfly: each step adds the same two values in both lanes, so every lane ends with the same
// bits)
__device__ __forceinline__ float wave_sum(float v) {
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
__device__ __forceinline__ void publish(double a, double b, HOBAK_GLOBAL double* pa, HOBAK_GLOBAL double* pb,
                                        int lb) {
  double v[2] = {a, b};
  block_sum<2>(v);
  if (threadIdx.x == 0) {
    pa[lb] = v[0];
    if (pb) pb[lb] = v[1];
  }
}

// Every consuming workgroup reduces the producer's n partials in the same fixed order, so all of
// them (and every run) obtain bit-identical totals.
__device__ __forceinline__ void gather(const HOBAK_GLOBAL double* pa, const HOBAK_GLOBAL double* pb, int n,
                                       double& ta, double& tb) {
  double v[2] = {0.0, 0.0};
  for (int i = threadIdx.x; i < n; i += kBlock) {
    v[0] += pa[i];
    if (pb) v[1] += pb[i];
  }
  block_sum<2>(v);
  ta = v[0];
  tb = v[1];
}

// The same reduction in two halves: pre_load issues the loads of up to kPre partials per thread
// (all in flight at once, before the caller's other work), pre_sum adds them (and any beyond
// kPre * kBlock) and reduces over the workgroup.  Summation order is fixed, as in gather().  TWO
// selects one array (pa) or two (pa, pb); n >= 1, so the loads need no branch.
constexpr int kPre = 4;
template <int KP = kPre>
struct PreParts {
  double a[KP], b[KP];
};
template <bool TWO, int KP>
__device__ __forceinline__ void pre_load(const HOBAK_GLOBAL double* pa, const HOBAK_GLOBAL double* pb, int n,
                                         PreParts<KP>& q) {
#pragma unroll
  for (int k = 0; k < KP; k++) {
    const int i = threadIdx.x + k * kBlock;
    const int ic = i < n ? i : 0;
    const double a = pa[ic];
    const double bv = TWO ? pb[ic] : 0.0;
    q.a[k] = i < n ? a : 0.0;
    q.b[k] = i < n ? bv : 0.0;
  }
}
template <bool TWO, int KP>
__device__ __forceinline__ void pre_sum(const HOBAK_GLOBAL double* pa, const HOBAK_GLOBAL double* pb, int n,
                                        const PreParts<KP>& q, double& ta, double& tb) {
  double v[2] = {0.0, 0.0};
#pragma unroll
  for (int k = 0; k < KP; k++) {
    v[0] += q.a[k];
    v[1] += q.b[k];
  }
  for (int i = threadIdx.x + KP * kBlock; i < n; i += kBlock) {
    v[0] += pa[i];
    if (TWO) v[1] += pb[i];
  }
  block_sum<2>(v);
  ta = v[0];
  tb = v[1];
}

// A lane's load at a 32-bit byte offset from a uniform base (plus a uniform offset): the address
// is a scalar base and one 32-bit lane register, so loads that share the lane offset share it.
template <typename T>
__device__ __forceinline__ T ld_at(const HOBAK_GLOBAL T* base, unsigned off) {
  return *(const HOBAK_GLOBAL T*)((const HOBAK_GLOBAL char*)base + off);
}
template <typename T>
__device__ __forceinline__ T ld_at(const HOBAK_GLOBAL T* base, size_t uniform_off, unsigned off) {
  return *(const HOBAK_GLOBAL T*)((const HOBAK_GLOBAL char*)base + uniform_off + off);
}
// The same as a streaming (non-temporal) load: data read once, which should not displace the
// gathered vectors from the cache.
template <typename T>
__device__ __forceinline__ T ld_stream(const HOBAK_GLOBAL T* base, size_t uniform_off, unsigned off) {
  return __builtin_nontemporal_load((const HOBAK_GLOBAL T*)((const HOBAK_GLOBAL char*)base + uniform_off + off));
}

// A value that is the same in every lane, moved to a scalar register.
__device__ __forceinline__ float uniform_f(float v) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
__device__ __forceinline__ double uniform_f(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffffll));
  const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// s + (m0 p0 + m1 p1 + m2 p2) with its two fused multiply-adds spelled out.  Kernels that must
// agree bitwise (the SpMV's variants, the FP64 two-kernel and p-form iterations) then contract
// the products the same way, whatever the compiler would pick for each.
template <typename T>
__device__ __forceinline__ T row3(T s, T m0, T m1, T m2, T p0, T p1, T p2) {
  return s + fma(m2, p2, fma(m0, p0, m1 * p1));
}
// a0 b0 + a1 b1 + a2 b2 in FP64 with its fused multiply-adds spelled out (for the same reason).
__device__ __forceinline__ double dot3(double a0, double a1, double a2, double b0, double b1, double b2) {
  return fma(a2, b2, fma(a1, b1, a0 * b0));
}

template <typename T>
__device__ __forceinline__ T d_own(bool rec, T z_own, T p_own) {
  return rec ? z_own : p_own;
}

__device__ __forceinline__ bool nonfinite(double v) { return !(v == v) || v > DBL_MAX || v < -DBL_MAX; }

// An instance whose current inner solve has ended and has not yet been handled.
__device__ __forceinline__ bool solve_ended(const HOBAK_GLOBAL Ctl* c) {
  return c->phase == PHASE_SOLVING && c->sc[0].done;
}

static inline int grid_of(const Batch& b, int cap) { return cap > 0 && cap < b.total ? cap : b.total; }

// ---------------------------------------------------------------------------------------------
// Kinematic collision cubes and Baraff-Witkin plane constraints (the FP64 form)
//   contact_begin (after begin_step): surface vertices that are inside a cube and move into it
//     become constrained (TIMESTEPPER::findNewSurfaceConstraints, TIMESTEPPER.cpp:446-526); the
//     filter S (I - n n^T, or 0 at pinned vertices, buildConstraintMatrix :128-165) and
//     z = (I - S) (closest point - x) (updateConstraintTargets, BACKWARD_EULER_POSITION.cpp:85-110)
//   assembly writes A unprojected; rhs forms c = b - A z, S c and S A S + (I - S) (:212, :242-243)
//   apply_dx adds z (:252)
//   contact_sep: the separation test at the step's positions with the unfiltered b
//     (findSeparatingSurfaceConstraints, TIMESTEPPER.cpp:252-329)
//   end_step: deletes separating constraints, re-targets the rest (:531-586) and filters the
//     velocity with S, rebuilt only when a constraint went away (BACKWARD_EULER_POSITION.cpp:283-314)
// The cube has no rotation: local = s^-1 (w - c), world = s l + c.  Products and sums are
// rounded one by one (no fused multiply-add), as the reference evaluates them.
// ---------------------------------------------------------------------------------------------
struct CubeD {
  double c[3], s, sinv;
};
__device__ __forceinline__ CubeD cube_of(const HOBAK_GLOBAL InstDev* I, int o) {
  CubeD q;
  for (int k = 0; k < 3; k++) q.c[k] = I->shape[o][k];
  q.s = I->shape[o][3];
  q.sinv = I->shape[o][4];
  return q;
}
__device__ __forceinline__ void cube_local(const CubeD& q, const double* w, double* l) {
  for (int k = 0; k < 3; k++) l[k] = __dmul_rn(q.sinv, __dsub_rn(w[k], q.c[k]));
}
__device__ __forceinline__ bool cube_in(const double* l) {  // CUBE.cpp:44-55
  return l[0] <= 0.5 && l[0] >= -0.5 && l[1] <= 0.5 && l[1] >= -0.5 && l[2] <= 0.5 && l[2] >= -0.5;
}
__device__ __forceinline__ double cube_sd(const CubeD& q, const double* w) {  // CUBE.cpp:90-119
  double l[3];
  cube_local(q, w, l);
  if (cube_in(l)) {
    const double xm = fmin(__dsub_rn(0.5, l[0]), __dadd_rn(l[0], 0.5));
    const double ym = fmin(__dsub_rn(0.5, l[1]), __dadd_rn(l[1], 0.5));
    const double zm = fmin(__dsub_rn(0.5, l[2]), __dadd_rn(l[2], 0.5));
    return __dmul_rn(-fabs(fmin(fmin(xm, ym), zm)), q.s);
  }
  double d[3] = {0.0, 0.0, 0.0};
  for (int k = 0; k < 3; k++) {
    if (l[k] > 0.5)
      d[k] = __dsub_rn(l[k], 0.5);
    else if (l[k] < -0.5)
      d[k] = __dsub_rn(-0.5, l[k]);
  }
  const double n2 = __dadd_rn(__dadd_rn(__dmul_rn(d[0], d[0]), __dmul_rn(d[1], d[1])), __dmul_rn(d[2], d[2]));
  return __dmul_rn(sqrt(n2), q.s);
}
// CUBE.cpp:124-191: the nearest face (the first of the six smallest distances)
__device__ __forceinline__ void cube_closest(const CubeD& q, const double* w, double* cl, double* nl) {
  double l[3];
  cube_local(q, w, l);
  const double diffs[6] = {__dadd_rn(0.5, l[0]), __dsub_rn(0.5, l[0]), __dadd_rn(0.5, l[1]),
                           __dsub_rn(0.5, l[1]), __dadd_rn(0.5, l[2]), __dsub_rn(0.5, l[2])};
  int mi = 0;
  double mf = diffs[0];
  for (int x = 1; x < 6; x++)
    if (diffs[x] < mf) {
      mf = diffs[x];
      mi = x;
    }
  const int axis = mi >> 1;
  const double sign = (mi & 1) ? 1.0 : -1.0;
  for (int k = 0; k < 3; k++) {
    cl[k] = k == axis ? 0.5 * sign : l[k];
    nl[k] = k == axis ? sign : 0.0;
  }
}
__device__ __forceinline__ double dot3_rn(const double* a, const double* b) {
  return __dadd_rn(__dadd_rn(__dmul_rn(a[0], b[0]), __dmul_rn(a[1], b[1])), __dmul_rn(a[2], b[2]));
}
// the filter block: I, I - n n^T or 0
__device__ __forceinline__ void filter_block(int mode, const double* n, double S[3][3]) {
  for (int a = 0; a < 3; a++)
    for (int c = 0; c < 3; c++) {
      const double id = a == c ? 1.0 : 0.0;
      S[a][c] = mode == 2 ? 0.0 : mode == 1 ? __dsub_rn(id, __dmul_rn(n[a], n[c])) : id;
    }
}
// the unit normal of the filter (normalized, as buildConstraintMatrix does)
__device__ __forceinline__ void unit_normal(const double* nl, double* n) {
  const double len = sqrt(dot3_rn(nl, nl));
  for (int k = 0; k < 3; k++) n[k] = __ddiv_rn(nl[k], len);
}

__global__ __launch_bounds__(kBlock) void contact_begin_kernel(Batch b) {
  for (int ws = blockIdx.x; ws < b.total; ws += gridDim.x) {
    const int vb = xcd_vb(ws, b.total);
    int lb, lg;
    const HOBAK_GLOBAL InstDev* __restrict__ I = seg_find(b, vb, lb, lg);
    if (!I->contact) continue;
    const int r = lb * kBlock + threadIdx.x;
    if (r >= I->nv) continue;
    double x[3];
    for (int k = 0; k < 3; k++) x[k] = gbl(I->rest)[3 * (size_t)r + k] + gbl(I->u)[3 * (size_t)r + k];
    const bool pin = gbl(I->pinned)[r];
    int o = gbl(I->pc_obj)[r];
    auto pcl = gbl(I->pc_cl) + 3 * (size_t)r;
    auto pnl = gbl(I->pc_nl) + 3 * (size_t)r;
    if (gbl(I->surf)[r] && o < 0) {
      double v[3];
      for (int k = 0; k < 3; k++) v[k] = gbl(I->v)[3 * (size_t)r + k];
      for (int q = 0; q < I->nshapes; q++) {
        const CubeD cb = cube_of(I, q);
        double l[3];
        cube_local(cb, x, l);
        if (!cube_in(l)) continue;
        double cl[3], nl[3];
        cube_closest(cb, x, cl, nl);
        if (dot3_rn(v, nl) >= -FLT_EPSILON) continue;  // moving away (or along): no constraint
        o = q;
        gbl(I->pc_obj)[r] = q;
        for (int k = 0; k < 3; k++) {
          pcl[k] = cl[k];
          pnl[k] = nl[k];
        }
        break;
      }
    }
    const int mode = pin ? 2 : (o >= 0 ? 1 : 0);
    double n[3] = {0.0, 0.0, 0.0}, tg[3] = {0.0, 0.0, 0.0};
    if (o >= 0) {
      double nl[3], cl[3], cw[3];
      for (int k = 0; k < 3; k++) {
        nl[k] = pnl[k];
        cl[k] = pcl[k];
      }
      unit_normal(nl, n);
      const CubeD cb = cube_of(I, o);
      for (int k = 0; k < 3; k++) {
        cw[k] = __dadd_rn(__dmul_rn(cb.s, cl[k]), cb.c[k]);
        tg[k] = __dsub_rn(cw[k], x[k]);
        // velocity form: the target is the velocity change (cl - x)/dt - v
        // (BACKWARD_EULER_VELOCITY.cpp:51-78)
        if (I->velform) tg[k] = __dsub_rn(__ddiv_rn(tg[k], I->dt), gbl(I->v)[3 * (size_t)r + k]);
      }
    }
    double S[3][3];
    filter_block(mode, n, S);
    gbl(I->fmode)[r] = (unsigned char)mode;
    for (int a = 0; a < 3; a++) {
      double ims[3];
      for (int c = 0; c < 3; c++) ims[c] = __dsub_rn(a == c ? 1.0 : 0.0, S[a][c]);
      gbl(I->fnrm)[3 * (size_t)r + a] = n[a];
      gbl(I->zc)[3 * (size_t)r + a] = dot3_rn(ims, tg);
    }
  }
}

// The positions the constraint logic saw during this step: rest + u after the pins were applied.
__device__ __forceinline__ void step_positions(const HOBAK_GLOBAL InstDev* I, int r, double* x) {
  const bool pin = gbl(I->pinned)[r];
  for (int k = 0; k < 3; k++) {
    const size_t i = 3 * (size_t)r + k;
    x[k] = gbl(I->rest)[i] + (pin ? gbl(I->pin_target)[i] : gbl(I->u_old)[i]);
  }
}

__global__ __launch_bounds__(kBlock) void contact_sep_kernel(Batch b) {
  for (int ws = blockIdx.x; ws < b.total; ws += gridDim.x) {
    const int vb = xcd_vb(ws, b.total);
    int lb, lg;
    const HOBAK_GLOBAL InstDev* __restrict__ I = seg_find(b, vb, lb, lg);
    if (!I->contact) continue;
    const int r = lb * kBlock + threadIdx.x;
    double count = 0.0;
    if (r < I->nv) {
      const int o = gbl(I->pc_obj)[r];
      int sep = 0;
      if (o >= 0) {
        double x[3];
        step_positions(I, r, x);
        const CubeD cb = cube_of(I, o);
        if (cube_sd(cb, x) > 1e-6) {
          sep = 1;
        } else {
          double xd[3], nl[3];
          for (int k = 0; k < 3; k++) {
            xd[k] = gbl(I->bunf)[3 * (size_t)r + k];
            nl[k] = gbl(I->pc_nl)[3 * (size_t)r + k];
          }
          const double len = sqrt(dot3_rn(xd, xd));
          if (len > 1.0)
            for (int k = 0; k < 3; k++) xd[k] = __ddiv_rn(xd[k], len);
          sep = dot3_rn(xd, nl) > 1e-6;
        }
      }
      gbl(I->pc_sep)[r] = (unsigned char)sep;
      count = sep;
    }
    publish(count, 0.0, gbl(I->part_sep), nullptr, lb);
  }
}

// End of a step for an instance with plane constraints: separating constraints go (when any
// does), the others take the closest point and normal at this step's positions, and
// v = S (u - u_old) / dt with S rebuilt only if a constraint went away; in the velocity form
// v = S v with this step's S.
__device__ void contact_end(const HOBAK_GLOBAL InstDev* I, int r, bool changed) {
  double x[3];
  step_positions(I, r, x);
  const bool pin = gbl(I->pinned)[r];
  int o = gbl(I->pc_obj)[r];
  int mode = gbl(I->fmode)[r];
  double n[3];
  for (int k = 0; k < 3; k++) n[k] = gbl(I->fnrm)[3 * (size_t)r + k];
  const int step_mode = mode;
  const double step_n[3] = {n[0], n[1], n[2]};
  if (o >= 0) {
    if (changed && gbl(I->pc_sep)[r]) {
      gbl(I->pc_obj)[r] = -1;
      o = -1;
    } else {
      double cl[3], nl[3];
      cube_closest(cube_of(I, o), x, cl, nl);
      for (int k = 0; k < 3; k++) {
        gbl(I->pc_cl)[3 * (size_t)r + k] = cl[k];
        gbl(I->pc_nl)[3 * (size_t)r + k] = nl[k];
      }
      if (changed) unit_normal(nl, n);
    }
  }
  if (I->velform) {
    // v = S v with this step's filter, before the constraints change (BACKWARD_EULER_VELOCITY.cpp:200)
    double S[3][3], w[3];
    filter_block(step_mode, step_n, S);
    for (int k = 0; k < 3; k++) w[k] = gbl(I->v)[3 * (size_t)r + k];
    for (int a = 0; a < 3; a++) gbl(I->v)[3 * (size_t)r + a] = dot3_rn(S[a], w);
    return;
  }
  if (changed) mode = pin ? 2 : (o >= 0 ? 1 : 0);
  double S[3][3], w[3];
  filter_block(mode, n, S);
  for (int k = 0; k < 3; k++) {
    const size_t i = 3 * (size_t)r + k;
    w[k] = I->invDt * (gbl(I->u)[i] - gbl(I->u_old)[i]);
  }
  for (int a = 0; a < 3; a++) gbl(I->v)[3 * (size_t)r + a] = dot3_rn(S[a], w);
}

// ---------------------------------------------------------------------------------------------
// step bookkeeping
// ---------------------------------------------------------------------------------------------
// _positionOld = _position; _velocityOld = _velocity; applyKinematicConstraints
// (BACKWARD_EULER_POSITION.cpp:152-157, TIMESTEPPER.cpp:105-122)
__global__ __launch_bounds__(kBlock) void begin_step_kernel(Batch b) {
  for (int ws = blockIdx.x; ws < b.total; ws += gridDim.x) {
    const int vb = xcd_vb(ws, b.total);
    int lb, lg;
    const HOBAK_GLOBAL InstDev* __restrict__ I = seg_find(b, vb, lb, lg);
    const int r = lb * kBlock + threadIdx.x;
    if (r < I->nv) {
      auto u = gbl(I->u);
      auto v = gbl(I->v);
      auto uo = gbl(I->u_old);
      auto vo = gbl(I->v_old);
      const bool pin = gbl(I->pinned)[r];
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const size_t i = 3 * (size_t)r + c;
        uo[i] = u[i];
        vo[i] = v[i];
        if (pin) u[i] = gbl(I->pin_target)[i];
      }
    }
  }
}

// _position += xDelta (BACKWARD_EULER_POSITION.cpp:262-266).  Velocity form
// (BACKWARD_EULER_VELOCITY.cpp:190-194): vDelta = y + z, v += vDelta, u += dt v; the S filter of
// v follows in end_step.
template <typename T>
__global__ __launch_bounds__(kBlock) void apply_dx_kernel(Batch b, int use64) {
  for (int ws = blockIdx.x; ws < b.total; ws += gridDim.x) {
    const int vb = xcd_vb(ws, b.total);
    int lb, lg;
    const HOBAK_GLOBAL InstDev* __restrict__ I = seg_find(b, vb, lb, lg);
    const int i = lb * kBlock + threadIdx.x;
    if (i < 3 * I->nv) {
      const double dx = use64 ? gbl(I->x64)[i] : (double)(gbl((const T*)I->x))[i];
      const double d = I->contact ? dx + gbl(I->zc)[i] : dx;  // y + z
      if (I->velform) {
        const double vn = gbl(I->v)[i] + d;
        gbl(I->v)[i] = vn;
        gbl(I->u)[i] = __dadd_rn(gbl(I->u)[i], __dmul_rn(I->dt, vn));
      } else {
        gbl(I->u)[i] += d;
      }
    }
  }
}

// _velocity = (u - u_old)/dt, then S-filtered (BACKWARD_EULER_POSITION.cpp:300-314); the velocity
// form filters its updated v with this step's S (BACKWARD_EULER_VELOCITY.cpp:200)
__global__ __launch_bounds__(kBlock) void end_step_kernel(Batch b) {
  for (int ws = blockIdx.x; ws < b.total; ws += gridDim.x) {
    const int vb = xcd_vb(ws, b.total);
    int lb, lg;
    const HOBAK_GLOBAL InstDev* __restrict__ I = seg_find(b, vb, lb, lg);
    const int r = lb * kBlock + threadIdx.x;
    if (I->contact) {
      double nsep, unused;
      gather(gbl(I->part_sep), nullptr, I->nb[K_ROW], nsep, unused);
      if (r < I->nv) contact_end(I, r, nsep > 0.0);
      continue;
    }
    if (r < I->nv) {
      const double invDt = I->invDt;
      const bool pin = gbl(I->pinned)[r];
      const bool vel = I->velform != 0;  // v already holds v + vDelta: only S filters it
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const size_t i = 3 * (size_t)r + c;
        gbl(I->v)[i] = pin ? 0.0 : vel ? gbl(I->v)[i] : invDt * (gbl(I->u)[i] - gbl(I->u_old)[i]);
      }
    }
  }
}

// ---------------------------------------------------------------------------------------------
// assembly
// ---------------------------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void write_precond(const T (&a)[9], HOBAK_GLOBAL T* minv, int r, int precond) {
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
  auto m = minv + (size_t)kMinvStride * r;
  m[0] = c0 * invdet;
  m[1] = c1 * invdet;
  m[2] = c2 * invdet;
#pragma unroll
  for (int row = 1; row < 3; row++)
#pragma unroll
    for (int k = 0; k < 3; k++) m[3 * row + k] = cof(k, row) * invdet;
}

// Entry e of a block's inverse by cofactors (write_precond's arithmetic, one entry per lane):
// m = cof(k, row) / det with row = e / 3, k = e % 3.
template <typename T>
__device__ __forceinline__ T precond_entry(const T (&a)[9], int e, int precond) {
  if (precond == 0) {  // DiagonalPreconditioner: 1/a_ii, 1 where a_ii == 0 (entries 0..2)
    const T d = a[4 * e];
    return (d != T(0)) ? T(1) / d : T(1);
  }
  auto cof = [&](int i, int j) {
    const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
    return a[3 * i1 + j1] * a[3 * i2 + j2] - a[3 * i1 + j2] * a[3 * i2 + j1];
  };
  const T c0 = cof(0, 0), c1 = cof(1, 0), c2 = cof(2, 0);
  const T det = (c0 * a[0] + c1 * a[3]) + c2 * a[6];
  const T invdet = T(1) / det;
  const int row = e / 3, k = e - 3 * row;
  return (row == 0 ? (k == 0 ? c0 : k == 1 ? c1 : c2) : cof(k, row)) * invdet;
}

// Nine lanes per upper block slot (row <= column), 7 slots per wavefront: lane 9 g + e holds
// entry e of the slot aslot[7 w + g] (the upper slots in slot order, so no lane waits on a lower
// slot, which its mirror writes).  A slot sums its
// staged contributions entry by entry in list order (the nine lanes read one staged block, 72 or
// 36 contiguous bytes, with one load instruction) and writes the slot and, off the diagonal, the
// transposed block into its mirror (lower) slot.
// ASM_REST / ASM_PLAIN_A: out = that sum (the stiffness, -K) into kneg / A.  ASM_SYSTEM:
// A = S (D - K) S + (I - S) (+ the ground penalty on the diagonal); the diagonal slot also
// writes its preconditioner.
template <typename T>
__global__ __launch_bounds__(kBlock) void assemble_kernel(Batch b, int target) {
  for (int ws = blockIdx.x; ws < b.total; ws += gridDim.x) {
    const int vb = xcd_vb(ws, b.total);
    int lb, lg;
    const HOBAK_GLOBAL InstDev* __restrict__ I = seg_find(b, vb, lb, lg);
    const int lane = threadIdx.x & 63;
    const int g = lane / 9, e = lane - 9 * g;
    const int k = (lb * (kBlock / 64) + (int)(threadIdx.x >> 6)) * kAsmSlotsPerWave + g;
    if (g >= kAsmSlotsPerWave || k >= I->nasm) continue;
    // from here on the nine lanes of the group are active together (the tests above are per slot)
    const int nslots = I->nslots;
    const int slot = gbl(I->aslot)[k];
    // the slot's record in one load round
    const int r = gbl(I->slot_row)[slot];
    const int c = gbl(I->col)[slot];
    const int kbeg = gbl(I->cptr)[slot], kend = gbl(I->cptr)[slot + 1];
    const int mirror = gbl(I->mirror)[slot];
    const int mslot = c == r ? -1 : mirror;
    const size_t ns = (size_t)nslots;
    const int te = 3 * (e % 3) + e / 3;  // entry (a, b) of a transposed block is entry (b, a)
    const int gl = lane - e;             // the group's first lane
    // the slot's contributions in ascending tet order, a flagged block summed transposed; eight
    // blocks are in flight per pass and the next pass's list entries are fetched with them
    auto stage = gbl((const T*)I->stage_h);
    auto ci = gbl(I->cidx);
    constexpr int kW = 8;
    int w[kW];
#pragma unroll
    for (int j = 0; j < kW; j++) w[j] = kbeg + j < kend ? ci[kbeg + j] : 0;
    T acc = T(0);
    for (int k = kbeg; k < kend; k += kW) {
      T v[kW];
#pragma unroll
      for (int j = 0; j < kW; j++)
        v[j] = stage[9 * (size_t)(w[j] & kCidxIndex) + ((w[j] & kCidxTranspose) ? te : e)];
#pragma unroll
      for (int j = 0; j < kW; j++) w[j] = k + kW + j < kend ? ci[k + kW + j] : 0;
#pragma unroll
      for (int j = 0; j < kW; j++)
        if (k + j < kend) acc += v[j];
    }
    if (target != ASM_SYSTEM) {
      auto out = gbl((T*)(target == ASM_REST ? I->kneg : I->A));
      const T accT = __shfl(acc, gl + te, 64);
      out[e * ns + slot] = acc;
      if (mslot >= 0) out[e * ns + mslot] = accT;
      continue;
    }
    constexpr bool kDiff = sizeof(T) < sizeof(double);
    // standard form: pinned rows and columns are zeroed; difference form: only pinned rows.  With
    // plane constraints A stays unprojected here: the right-hand side kernel needs A z, then
    // projects each row in place.
    const bool contact = I->contact != 0;
    const bool pin_r = gbl(I->pinned)[r] != 0, pin_c = gbl(I->pinned)[c] != 0;
    const bool pin = contact ? false : kDiff ? pin_r : (pin_r || pin_c);
    bool ground = false;
    if (I->ground_on && c == r && !pin) {
      const double y = gbl(I->rest)[3 * (size_t)r + 1] + gbl(I->u)[3 * (size_t)r + 1];
      ground = y < I->ground_y;
      if (ground && e == 4) acc += (T)I->ground_k;
    }
    const T accT = __shfl(acc, gl + te, 64);
    auto D = gbl((const T*)I->D);
    auto out = gbl((T*)I->A);
    // velocity form: A = M - dt^2 K (BACKWARD_EULER_VELOCITY.cpp:161), with K = -acc; the
    // unprojected stiffness also goes to E for b's K v
    const bool vel = I->velform != 0;
    const T ks = (T)I->kscale;
    const T a = vel ? D[e * ns + slot] + ks * acc : D[e * ns + slot] + acc;
    const T val = pin ? ((c == r && (e % 4) == 0) ? T(1) : T(0)) : a;
    out[e * ns + slot] = val;
    if (mslot >= 0) {
      const bool pin_m = contact ? false : kDiff ? pin_c : (pin_r || pin_c);
      const T am = vel ? D[e * ns + mslot] + ks * accT : D[e * ns + mslot] + accT;
      out[e * ns + mslot] = pin_m ? T(0) : am;
    }
    if (vel) {
      auto kst = gbl((T*)I->E);
      kst[e * ns + slot] = acc;
      if (mslot >= 0) kst[e * ns + mslot] = accT;
    }
    if (c == r && !contact) {
      T blk[9];
#pragma unroll
      for (int j = 0; j < 9; j++) blk[j] = __shfl(val, gl + j, 64);
      const int precond = I->precond;
      if (precond == 0) {
        if (e < 3) gbl((T*)I->minv)[3 * (size_t)r + e] = precond_entry<T>(blk, e, 0);
      } else {
        gbl((T*)I->minv)[(size_t)kMinvStride * r + e] = precond_entry<T>(blk, e, precond);
      }
      if (kDiff && e < 3) {
        const double mk = gbl(I->mass)[3 * (size_t)r + e] * (I->invDt2 - I->ra * I->invDt);
        gbl(I->dg)[3 * (size_t)r + e] = pin ? 1.0 : mk + ((e == 1 && ground) ? I->ground_k : 0.0);
      }
    }
  }
}

// D = M/dt^2 - C/dt, E = M/dt - C, C = alpha M + beta K_rest, from the assembled rest stiffness
// (-K_rest in kneg).  Order of operations as BACKWARD_EULER_POSITION.cpp:198, 208 and
// TIMESTEPPER.cpp:232-233.  Velocity form: C is emptied before the system is formed
// (BACKWARD_EULER_VELOCITY.cpp:153, TIMESTEPPER.cpp:657), so D = M and C = E = 0.
template <typename T>
__global__ __launch_bounds__(kBlock) void constant_parts_kernel(Batch b) {
  for (int ws = blockIdx.x; ws < b.total; ws += gridDim.x) {
    const int vb = xcd_vb(ws, b.total);
    int lb, lg;
    const HOBAK_GLOBAL InstDev* __restrict__ I = seg_find(b, vb, lb, lg);
    const int slot = lb * kBlock + threadIdx.x;
    const int nslots = I->nslots;
    if (slot >= nslots) continue;
    const int r = gbl(I->slot_row)[slot];
    if (r < 0) continue;
    const int c = gbl(I->col)[slot];
    const size_t ns = (size_t)nslots;
    const double ra = I->ra, rb = I->rb, invDt = I->invDt, invDt2 = I->invDt2;
#pragma unroll
    for (int e = 0; e < 9; e++) {
      const double krest = -(double)(gbl((const T*)I->kneg))[e * ns + slot];
      const bool dg = (c == r) && (e % 4 == 0);
      const double m = dg ? gbl(I->mass)[3 * (size_t)r + e / 4] : 0.0;
      const double cc = dg ? ra * m + rb * krest : rb * krest;
      double d = dg ? m * invDt2 - cc * invDt : -(cc * invDt);
      double ee = dg ? invDt * m - cc : -cc;
      double cw = cc;
      if (I->velform) {  // C is emptied before A = M - dt C - dt^2 K is formed: D = M
        d = m;
        ee = 0.0;
        cw = 0.0;
      }
      (gbl((T*)I->D))[e * ns + slot] = (T)d;
      (gbl((T*)I->E))[e * ns + slot] = (T)ee;
      (gbl((T*)I->C))[e * ns + slot] = (T)cw;
    }
  }
}

// ---------------------------------------------------------------------------------------------
// right-hand side + PCG initialisation
// ---------------------------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void apply_precond(const HOBAK_GLOBAL T* minv, int precond, int r, const T (&rv)[3],
                                              T (&zv)[3]) {
  if (precond == 0) {
#pragma unroll
    for (int c = 0; c < 3; c++) zv[c] = minv[3 * (size_t)r + c] * rv[c];
  } else {
    auto m = minv + (size_t)kMinvStride * r;
#pragma unroll
    for (int i = 0; i < 3; i++) zv[i] = m[3 * i] * rv[0] + m[3 * i + 1] * rv[1] + m[3 * i + 2] * rv[2];
  }
}

// Per block row: R = ascending-tet gather of staged forces; b = (E v + R) + f_ext (BEP),
// b = C vb + ((M a + R) + f_ext) (BDF_1) or b = dt (R + dt K v + f_ext) (the velocity form, with
// E holding -K); rhs = S b.  With init_cg also x = 0, r = rhs,
// z = M^-1 r; the partial sums of |rhs|^2 and r.z seed Eigen's CG (ConjugateGradient.h:43-66).
// Row r of an instance with plane constraints (FP64 form): b (unfiltered, kept for the
// separation test) becomes S (b - A z), and the row's blocks of the unprojected A become
// S_r A_rj S_j (+ (I - S_r) on the diagonal), in place: the row's thread owns its slots.  The
// preconditioner is taken from the projected diagonal block.
template <typename T>
__device__ void contact_rhs(const HOBAK_GLOBAL InstDev* I, int r, double (&bv)[3]) {
  const size_t ns = (size_t)I->nslots;
  const int sh = I->shift;
  auto col = gbl(I->col);
  auto A = gbl((T*)I->A);
  auto zc = gbl(I->zc);
  for (int k = 0; k < 3; k++) gbl(I->bunf)[3 * (size_t)r + k] = bv[k];
  double Sr[3][3], nr[3];
  for (int k = 0; k < 3; k++) nr[k] = gbl(I->fnrm)[3 * (size_t)r + k];
  filter_block(gbl(I->fmode)[r], nr, Sr);
  double az[3] = {0.0, 0.0, 0.0};
  T diag[9];
  bool has_diag = false;
  const int len = gbl(I->row_len)[r];
  const int base = gbl(I->slice_off)[r >> sh] + (r & ((1 << sh) - 1));
  for (int k = 0; k < len; k++) {
    const int slot = base + (k << sh);
    const int c = col[slot];
    double a[3][3], Sc[3][3], nc[3], zj[3];
    for (int e = 0; e < 9; e++) a[e / 3][e % 3] = (double)A[e * ns + slot];
    for (int j = 0; j < 3; j++) {
      nc[j] = gbl(I->fnrm)[3 * (size_t)c + j];
      zj[j] = zc[3 * (size_t)c + j];
    }
    for (int i = 0; i < 3; i++) az[i] += (a[i][0] * zj[0] + a[i][1] * zj[1]) + a[i][2] * zj[2];
    filter_block(gbl(I->fmode)[c], nc, Sc);
    double sa[3][3];
    for (int i = 0; i < 3; i++)
      for (int d = 0; d < 3; d++) sa[i][d] = (Sr[i][0] * a[0][d] + Sr[i][1] * a[1][d]) + Sr[i][2] * a[2][d];
    for (int i = 0; i < 3; i++)
      for (int d = 0; d < 3; d++) {
        double v = (sa[i][0] * Sc[0][d] + sa[i][1] * Sc[1][d]) + sa[i][2] * Sc[2][d];
        if (c == r) v += (i == d ? 1.0 : 0.0) - Sr[i][d];
        A[(3 * i + d) * ns + slot] = (T)v;
        if (c == r) diag[3 * i + d] = (T)v;
      }
    has_diag |= c == r;
  }
  double cv[3];
  for (int i = 0; i < 3; i++) cv[i] = bv[i] - az[i];
  for (int i = 0; i < 3; i++) bv[i] = (Sr[i][0] * cv[0] + Sr[i][1] * cv[1]) + Sr[i][2] * cv[2];
  if (has_diag) write_precond<T>(diag, gbl((T*)I->minv), r, I->precond);
}

// A vertex's force: the staged corner forces of its list fidx[q0 .. q1) (ascending tet), summed
// in FP64 in list order.  Eight corners are in flight per pass and the next pass's list entries
// are fetched with them.
template <typename T>
__device__ __forceinline__ void gather_forces(const HOBAK_GLOBAL T* sf, const HOBAK_GLOBAL int* fi, int q0,
                                              int q1, double (&f)[3]) {
  f[0] = f[1] = f[2] = 0.0;
  constexpr int kW = 8;
  int w[kW];
#pragma unroll
  for (int j = 0; j < kW; j++) w[j] = q0 + j < q1 ? fi[q0 + j] : 0;
  for (int q = q0; q < q1; q += kW) {
    T v[kW][3];
#pragma unroll
    for (int j = 0; j < kW; j++)
#pragma unroll
      for (int c = 0; c < 3; c++) v[j][c] = sf[3 * (size_t)w[j] + c];
#pragma unroll
    for (int j = 0; j < kW; j++) w[j] = q + kW + j < q1 ? fi[q + kW + j] : 0;
#pragma unroll
    for (int j = 0; j < kW; j++)
      if (q + j < q1) {
#pragma unroll
        for (int c = 0; c < 3; c++) f[c] += (double)v[j][c];
      }
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void rhs_kernel(Batch b, int init_cg) {
  for (int ws = blockIdx.x; ws < b.total; ws += gridDim.x) {
    const int vb = xcd_vb(ws, b.total);
    int lb, lg;
    const HOBAK_GLOBAL InstDev* __restrict__ I = seg_find(b, vb, lb, lg);
    const int r = lb * kBlock + threadIdx.x;
    double part[2] = {0.0, 0.0};
    if (r < I->nv) {
      auto sf = gbl((const T*)I->stage_f);
      auto vptr = gbl(I->vptr);
      double R[3];
      gather_forces(sf, gbl(I->fidx), vptr[r], vptr[r + 1], R);
      auto u = gbl(I->u);
      if (I->ground_on) {
        const double y = gbl(I->rest)[3 * (size_t)r + 1] + u[3 * (size_t)r + 1];
        if (y < I->ground_y) R[1] += I->ground_k * (I->ground_y - y);
      }
      const int bdf = I->bdf;
      const double invDt = I->invDt, invDt2 = I->invDt2;
      auto uOld = gbl(I->u_old);
      auto vOld = gbl(I->v_old);
      auto vel = gbl(I->v);
      const size_t ns = (size_t)I->nslots;
      auto Mx = gbl((const T*)(bdf ? I->C : I->E));
      const int sh = I->shift;
      auto col = gbl(I->col);
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
          s[j] = gbl(I->mass)[i] * em * own[j];
        }
      }
      const int len = gbl(I->row_len)[r];
      const int base = gbl(I->slice_off)[r >> sh] + (r & ((1 << sh) - 1));
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
      const bool pin = gbl(I->pinned)[r];
      const double orv = gbl(I->one_ring)[r];
      double bv[3];
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const size_t i = 3 * (size_t)r + c;
        const double fext = orv * I->g[c];
        if (bdf) {
          const double acc = -invDt2 * (u[i] - uOld[i]) + invDt * vOld[i];
          bv[c] = s[c] + (((gbl(I->mass)[i] * acc + R[c]) + fext));
        } else if (I->velform) {
          // b = dt (R + dt K v + f_ext) (BACKWARD_EULER_VELOCITY.cpp:157); s = -K v here
          const double dt = I->dt;
          bv[c] = __dmul_rn(dt, __dadd_rn(__dsub_rn(R[c], __dmul_rn(dt, s[c])), fext));
        } else {
          bv[c] = (s[c] + R[c]) + fext;
        }
        if (pin && !I->contact) bv[c] = 0.0;
      }
      if (!kDiff && I->contact) contact_rhs<T>(I, r, bv);
#pragma unroll
      for (int c = 0; c < 3; c++) {
        gbl(I->rhs64)[3 * (size_t)r + c] = bv[c];
        part[0] += bv[c] * bv[c];
      }
      if (init_cg) {
        T rv[3], zv[3];
#pragma unroll
        for (int c = 0; c < 3; c++) rv[c] = (T)bv[c];
        apply_precond<T>(gbl((const T*)I->minv), I->precond, r, rv, zv);
#pragma unroll
        for (int c = 0; c < 3; c++) {
          const size_t i = 3 * (size_t)r + c;
          (gbl((T*)I->r))[i] = rv[c];
          (gbl((T*)I->x))[i] = T(0);
          (gbl((T*)I->z))[i] = zv[c];
          part[1] += (double)rv[c] * (double)zv[c];
        }
      }
    }
    publish(part[0], part[1], gbl(I->part_row_a), gbl(I->part_row_b), lb);
  }
}

// sc[0] from the seeding kernel's partials (the right-hand side kernel, or the refinement's
// inner-solve seed): Eigen's early exits on a zero or already small right-hand side
// (ConjugateGradient.h:45-61) and the first r.z.  With gate, only instances whose inner solve
// has just ended and that still need another refinement sweep are (re)started.
__global__ __launch_bounds__(kBlock) void cg_init_kernel(Batch b, int refine, int gate) {
  for (int ws = blockIdx.x; ws < b.total; ws += gridDim.x) {
    const int vb = xcd_vb(ws, b.total);
    int lb, lg;
    const HOBAK_GLOBAL InstDev* __restrict__ I = seg_find(b, vb, lb, lg);
    auto c = ctl_of(I);
    if (gate && !solve_ended(c)) continue;
    double bb, rz;
    gather(gbl(I->part_row_a), gbl(I->part_row_b), I->nb[K_ROW], bb, rz);
    // coarse correction of the seed (refinement sweeps of an instance with I->coarse):
    // e = (sum dg)^-1 (sum r) per component, and r.z gains (P^T r).e
    double ce[3] = {0.0, 0.0, 0.0}, ai[3] = {0.0, 0.0, 0.0};
    const bool coarse = refine && I->coarse != 0;
    if (coarse) {
      const int nb = I->nb[K_ROW];
      auto pc = gbl(I->part_row_c);
      double g[3], d[3];
      gather(pc, pc + nb, nb, g[0], g[1]);
      gather(pc + 2 * (size_t)nb, pc + 3 * (size_t)nb, nb, g[2], d[0]);
      gather(pc + 4 * (size_t)nb, pc + 5 * (size_t)nb, nb, d[1], d[2]);
#pragma unroll
      for (int k = 0; k < 3; k++) {
        ai[k] = d[k] > 0.0 ? 1.0 / d[k] : 0.0;
        ce[k] = ai[k] * g[k];
      }
      rz += g[0] * ce[0] + g[1] * ce[1] + g[2] * ce[2];
    }
    if (threadIdx.x != 0) continue;
    for (int k = 0; k < 3; k++) c->ac_inv[k] = ai[k];
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
    auto s = &c->sc[0];
    s->maxR = refine ? c->rs.b2 : bb;
    s->rhsNorm2 = bb;
    s->resNorm2 = bb;
    s->threshold = thr;
    s->absNew = rz;
    s->e[0] = ce[0];
    s->e[1] = ce[1];
    s->e[2] = ce[2];
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
template <typename T, int H, bool PAD, bool NT, bool CO>
__global__ __launch_bounds__(kBlock) void cg_spmv_kernel(Batch b, int parity) {
  for (int ws = blockIdx.x; ws < b.total; ws += gridDim.x) {
    const int vb = xcd_vb(ws, b.total);
    int lb, lg;
    const HOBAK_GLOBAL InstDev* __restrict__ I = seg_find(b, vb, lb, lg);
    const int lane = threadIdx.x & 63;
    const int slice = (int)((lb * (unsigned)kBlock + threadIdx.x) >> 6);
    const int rr = lane % H;
    const int r = slice * H + rr;
    // First load group: every field of the record and of the scalar set that this kernel uses.
    const int nrows = I->nv;
    const int nupd = I->nb[K_UPD];
    const size_t ns = (size_t)I->nslots;
    const auto slice_off = gbl(I->slice_off);
    const auto row_len = gbl(I->row_len);
    const auto prr = gbl(I->part_rr), prz = gbl(I->part_rz);
    const auto p0 = gbl((T*)I->p0), p1 = gbl((T*)I->p1);
    const auto z = gbl((const T*)I->z);
    const auto A = gbl((const T*)I->A);
    const auto col = gbl(I->col);
    const auto t = gbl((T*)I->t);
    const auto dg = gbl(I->dg);
    const auto ppt = gbl(I->part_pt);
    auto ctl = ctl_of(I);
    const auto S = &ctl->sc[parity];
    auto N = &ctl->sc[parity ^ 1];
    const int done = S->done;
    const int first = S->first;
    const double absOld = S->absNew;
    const int iters = S->iters;
    const double threshold = S->threshold;
    const int maxIters = S->maxIters;
    // coarse correction (reduced-precision form only): e of the first set, or 1 / P^T A P and the
    // update kernel's sums of r, from which e = (P^T A P)^-1 P^T r
    const bool coarse = CO && I->coarse != 0;
    const auto pg = gbl(I->part_upd_g);
    double ce0 = 0.0, ce1 = 0.0, ce2 = 0.0, ai0 = 0.0, ai1 = 0.0, ai2 = 0.0;
    if (coarse) {
      ce0 = S->e[0];
      ce1 = S->e[1];
      ce2 = S->e[2];
      ai0 = ctl->ac_inv[0];
      ai1 = ctl->ac_inv[1];
      ai2 = ctl->ac_inv[2];
    }
    const bool lead = lb == 0 && threadIdx.x == 0;
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
    const int nslices = (nrows + H - 1) / H;
    const bool wave_live = slice < nslices;
    const bool live = r < nrows;
    const auto pold = parity ? p1 : p0;
    const auto pnew = parity ? p0 : p1;
    // Second load group: the slice extent, the row length, the update kernel's partials and the row's own
    // z and p_old, all before any of them is used.
    // Indices are clamped instead of branched on, so these loads are issued together.
    const int sc = wave_live ? slice : nslices - 1;
    const int rc = live ? r : nrows - 1;
    int sbeg = slice_off[sc];
    int send = slice_off[sc + 1];
    int len = PAD ? 0 : row_len[rc];
    PreParts<> q;
    pre_load<true>(prr, prz, nupd, q);
    // The reduced-precision form (T = float) uses the difference-form operator and forms
    // t = A p_new as A z + beta t_old (kRec): it gathers z only, where the FP64 form gathers z and
    // p_old and forms p_new at every column, Eigen's arithmetic.
    constexpr bool kDiff = sizeof(T) < sizeof(double);
    constexpr bool kRec = kDiff;
    const size_t io = 3 * (size_t)rc;
    const auto pq = first ? z : pold;  // p_old is not read for the first iteration (beta = 0)
    T zo0 = z[io], zo1 = z[io + 1], zo2 = z[io + 2];
    const T qo0 = pq[io], qo1 = pq[io + 1], qo2 = pq[io + 2];
    T to0 = T(0), to1 = T(0), to2 = T(0);
    if (kRec) {  // t_old of the own row; used only after the first iteration
      to0 = t[io];
      to1 = t[io + 1];
      to2 = t[io + 2];
    }
    if (!wave_live) sbeg = send = 0;
    if (!live) {
      len = 0;
      zo0 = zo1 = zo2 = T(0);
    }
    T po0 = zo0, po1 = zo1, po2 = zo2;
    T beta = T(0);
    // The first pass's column indices depend only on the slice bounds: they are issued before the
    // partial-sum reduction below, so its LDS work overlaps their latency.
    const int nit = wave_live ? (send - sbeg) >> 6 : 0;
    // per-lane step count: the padded width (PAD), or just the lane's own blocks of its row
    const int nj = PAD ? nit : (live ? (len - lane / H + (64 / H) - 1) / (64 / H) : 0);
    // Step j of this lane reads slot sbeg + lane + 64 j.  A step past the lane's own blocks
    // re-reads step 0's slot (a cache hit, no HBM traffic) and its column values are zeroed, so
    // every pass issues the same loads without branches.  Each pass also fetches the next
    // pass's two column indices, so its gathers never wait on a column load.  All addresses
    // are 32-bit byte offsets from uniform bases (the instance limits keep them in range).
    const unsigned slot0 = (unsigned)(sbeg + lane);
    auto slot_of = [&](int j) -> unsigned { return j < nj ? slot0 + 64u * (unsigned)j : slot0; };
    unsigned ca = 0u, cb = 0u;
    if (wave_live) {
      ca = ld_at(col, slot_of(0) * 4u);
      cb = ld_at(col, slot_of(1) * 4u);
    }
    if (!first) {
      double rrv, rz;
      pre_sum<true>(prr, prz, nupd, q, rrv, rz);
      if (coarse) {
        double g0, g1, g2, unused;
        gather(pg, pg + nupd, nupd, g0, g1);
        gather(pg + 2 * (size_t)nupd, nullptr, nupd, g2, unused);
        ce0 = ai0 * g0;
        ce1 = ai1 * g1;
        ce2 = ai2 * g2;
        rz += g0 * ce0 + g1 * ce1 + g2 * ce2;  // r.(M^-1 r) + (P^T r).e
      }
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
      if (coarse && live) {  // the whole z of the own row: M^-1 r + e
        po0 += (T)ce0;
        po1 += (T)ce1;
        po2 += (T)ce2;
      }
      po0 = fma(beta, qo0, po0);
      po1 = fma(beta, qo1, po1);
      po2 = fma(beta, qo2, po2);
    } else if (coarse && live) {
      po0 += (T)ce0;
      po1 += (T)ce1;
      po2 += (T)ce2;
    }
    // e as uniform values for the diagonal term below (the columns' differences need only
    // M^-1 r: e cancels in them)
    const T cf0 = uniform_f((T)ce0), cf1 = uniform_f((T)ce1), cf2 = uniform_f((T)ce2);
    if (lead) ctl->spmv_work += 1u;  // this launch streams A (a measurement count)
    // FP64 form: p_new = z + beta p_old at every column (p_old is not read for the first
    // iteration).  FP32 form: A z in the difference form (columns relative to the row's own z, dg
    // adds the diagonal term), then + beta t_old.
    T s0 = T(0), s1 = T(0), s2 = T(0);
    if (wave_live) {
      // difference form: columns relative to the row's own entry of the vector multiplied
      const T d0 = kRec ? zo0 : kDiff ? po0 : T(0);
      const T d1 = kRec ? zo1 : kDiff ? po1 : T(0);
      const T d2 = kRec ? zo2 : kDiff ? po2 : T(0);
      const T bt = first ? T(0) : beta;
      constexpr unsigned ts = sizeof(T), vs = 3 * sizeof(T);
      const size_t plane = ns * ts;  // bytes between value planes
      for (int j = 0; j < nit; j += 2) {
        const unsigned oa = slot_of(j) * ts, ob = slot_of(j + 1) * ts;
        const unsigned ga = ca * vs, gb = cb * vs;
        T ma[9], mb[9];
#pragma unroll
        for (int e = 0; e < 9; e++) {
          ma[e] = NT ? ld_stream(A, (size_t)e * plane, oa) : ld_at(A, (size_t)e * plane, oa);
          mb[e] = NT ? ld_stream(A, (size_t)e * plane, ob) : ld_at(A, (size_t)e * plane, ob);
        }
        // (component offsets are added in 64 bits, so the three loads of a vertex are known to
        // be adjacent and merge into wide loads)
        const T za0 = ld_at(z, 0, ga), za1 = ld_at(z, ts, ga), za2 = ld_at(z, 2 * ts, ga);
        const T qa0 = kRec ? T(0) : ld_at(pq, 0, ga), qa1 = kRec ? T(0) : ld_at(pq, ts, ga),
                qa2 = kRec ? T(0) : ld_at(pq, 2 * ts, ga);
        const T zb0 = ld_at(z, 0, gb), zb1 = ld_at(z, ts, gb), zb2 = ld_at(z, 2 * ts, gb);
        const T qb0 = kRec ? T(0) : ld_at(pq, 0, gb), qb1 = kRec ? T(0) : ld_at(pq, ts, gb),
                qb2 = kRec ? T(0) : ld_at(pq, 2 * ts, gb);
        ca = ld_at(col, slot_of(j + 2) * 4u);
        cb = ld_at(col, slot_of(j + 3) * 4u);
        const bool va = j < nj, vb = j + 1 < nj;
        const T pa0 = va ? (kRec ? za0 : fma(bt, qa0, za0)) - d0 : T(0);
        const T pa1 = va ? (kRec ? za1 : fma(bt, qa1, za1)) - d1 : T(0);
        const T pa2 = va ? (kRec ? za2 : fma(bt, qa2, za2)) - d2 : T(0);
        const T pb0 = vb ? (kRec ? zb0 : fma(bt, qb0, zb0)) - d0 : T(0);
        const T pb1 = vb ? (kRec ? zb1 : fma(bt, qb1, zb1)) - d1 : T(0);
        const T pb2 = vb ? (kRec ? zb2 : fma(bt, qb2, zb2)) - d2 : T(0);
        s0 = row3(s0, ma[0], ma[1], ma[2], pa0, pa1, pa2);
        s1 = row3(s1, ma[3], ma[4], ma[5], pa0, pa1, pa2);
        s2 = row3(s2, ma[6], ma[7], ma[8], pa0, pa1, pa2);
        s0 = row3(s0, mb[0], mb[1], mb[2], pb0, pb1, pb2);
        s1 = row3(s1, mb[3], mb[4], mb[5], pb0, pb1, pb2);
        s2 = row3(s2, mb[6], mb[7], mb[8], pb0, pb1, pb2);
        // The prefetched indices are consumed here, at the end of the pass: otherwise the
        // compiler merges this load with the one before the loop and issues it at the top of the
        // next pass, where the gathers would wait for it.
        asm volatile("" : "+v"(ca), "+v"(cb));
      }
    }
#pragma unroll
    for (int off = H; off < 64; off <<= 1) {
      s0 += __shfl_xor(s0, off, 64);
      s1 += __shfl_xor(s1, off, 64);
      s2 += __shfl_xor(s2, off, 64);
    }
    double pt = 0.0;
    if (live && lane < H) {
      const size_t i = 3 * (size_t)r;
      if (kDiff) {
        s0 = fma((T)dg[i], d_own(kRec, coarse ? zo0 + cf0 : zo0, po0), s0);
        s1 = fma((T)dg[i + 1], d_own(kRec, coarse ? zo1 + cf1 : zo1, po1), s1);
        s2 = fma((T)dg[i + 2], d_own(kRec, coarse ? zo2 + cf2 : zo2, po2), s2);
      }
      if (kRec && !first) {
        s0 = fma(beta, to0, s0);
        s1 = fma(beta, to1, s1);
        s2 = fma(beta, to2, s2);
      }
      pt = dot3((double)po0, (double)po1, (double)po2, (double)s0, (double)s1, (double)s2);
    }
    // Lane rho < H now holds row slice*H + rho.  The slice's rows own the 3H consecutive entries
    // of t and p_new from 3*slice*H, and lane l stores entries l, l + 64, ...: every store
    // instruction writes whole lines instead of three strided partial-line stores per row.
    if (wave_live) {
      const size_t base = 3 * (size_t)slice * H;
      const int nval = 3 * min(H, nrows - slice * H);
#pragma unroll
      for (int d0 = 0; d0 < 3 * H; d0 += 64) {
        const int d = d0 + lane;
        const int rho = d / 3;
        const int comp = d - 3 * rho;
        const int src = rho < H ? rho : 0;
        const T v0 = __shfl(s0, src, 64), v1 = __shfl(s1, src, 64), v2 = __shfl(s2, src, 64);
        const T q0 = __shfl(po0, src, 64), q1 = __shfl(po1, src, 64), q2 = __shfl(po2, src, 64);
        if (d < nval) {
          t[base + d] = comp == 0 ? v0 : comp == 1 ? v1 : v2;
          pnew[base + d] = comp == 0 ? q0 : comp == 1 ? q1 : q2;
        }
      }
    }
    publish(pt, 0.0, ppt, nullptr, lb);
  }
}

// ---------------------------------------------------------------------------------------------
// The FP64 iteration in three kernels (the p-form), for solving sets whose matrices stream from
// HBM: the SpMV above gathers z and p_old at every column to form p_new = z + beta p_old on the
// fly, two gathered vectors per block; here
//   p(p):       the previous iteration is finished exactly as the SpMV's first part does it (the
//               same partial sums, tests and scalar set), and p_new = z + beta p_old is formed once
//               per entry, with the same fused multiply-add;
//   spmv_p(p):  t = A p_new gathering p_new only, the partials of p.t;
//   update(p):  unchanged.
// Every value is computed by the same operations in the same order as in the two-kernel form, so
// the results are bitwise those of the two-kernel form.
// ---------------------------------------------------------------------------------------------
// One thread per DOF entry of p, over the instance's update-kernel workgroups (capped per
// instance, so few workgroups re-reduce the update kernel's partials).
__global__ __launch_bounds__(kBlock) void cg_p_kernel(Batch b, int parity) {
  for (int ws = blockIdx.x; ws < b.total; ws += gridDim.x) {
    const int vb = xcd_vb(ws, b.total);
    int lb, lg;
    const HOBAK_GLOBAL InstDev* __restrict__ I = seg_find(b, vb, lb, lg);
    const int nupd = I->nb[K_UPD];
    const auto prr = gbl(I->part_rr), prz = gbl(I->part_rz);
    const auto p0 = gbl((double*)I->p0), p1 = gbl((double*)I->p1);
    const auto z = gbl((const double*)I->z);
    auto ctl = ctl_of(I);
    const auto S = &ctl->sc[parity];
    auto N = &ctl->sc[parity ^ 1];
    const int done = S->done;
    const int first = S->first;
    const double absOld = S->absNew;
    const int iters = S->iters;
    const double threshold = S->threshold;
    const int maxIters = S->maxIters;
    const bool lead = lb == 0 && threadIdx.x == 0;
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
    double beta = 0.0;
    if (!first) {
      PreParts<> q;
      pre_load<true>(prr, prz, nupd, q);
      double rrv, rz;
      pre_sum<true>(prr, prz, nupd, q, rrv, rz);
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
      beta = bt;
    }
    const auto pold = parity ? p1 : p0;
    const auto pnew = parity ? p0 : p1;
    const int n = 3 * I->nv;
    for (int i = lb * kBlock + threadIdx.x; i < n; i += lg * kBlock) pnew[i] = first ? z[i] : fma(beta, pold[i], z[i]);
  }
}

// t = A p_new gathering p_new only (FP64 standard form); the layout and loop are the SpMV's.
template <int H, bool PAD, bool NT>
__global__ __launch_bounds__(kBlock) void cg_spmv_p_kernel(Batch b, int parity) {
  typedef double T;
  for (int ws = blockIdx.x; ws < b.total; ws += gridDim.x) {
    const int vb = xcd_vb(ws, b.total);
    int lb, lg;
    const HOBAK_GLOBAL InstDev* __restrict__ I = seg_find(b, vb, lb, lg);
    const int lane = threadIdx.x & 63;
    const int slice = (int)((lb * (unsigned)kBlock + threadIdx.x) >> 6);
    const int rr = lane % H;
    const int r = slice * H + rr;
    const int nrows = I->nv;
    const size_t ns = (size_t)I->nslots;
    const auto slice_off = gbl(I->slice_off);
    const auto row_len = gbl(I->row_len);
    const auto pnew = gbl((const T*)(parity ? I->p0 : I->p1));
    const auto A = gbl((const T*)I->A);
    const auto col = gbl(I->col);
    const auto t = gbl((T*)I->t);
    const auto ppt = gbl(I->part_pt);
    if (ctl_of(I)->sc[parity ^ 1].done) continue;  // the p kernel ended the solve
    if (lb == 0 && threadIdx.x == 0) ctl_of(I)->spmv_work += 1u;  // this launch streams A
    const int nslices = (nrows + H - 1) / H;
    const bool wave_live = slice < nslices;
    const bool live = r < nrows;
    const int sc = wave_live ? slice : nslices - 1;
    const int rc = live ? r : nrows - 1;
    int sbeg = slice_off[sc];
    int send = slice_off[sc + 1];
    int len = PAD ? 0 : row_len[rc];
    const size_t io = 3 * (size_t)rc;
    T po0 = pnew[io], po1 = pnew[io + 1], po2 = pnew[io + 2];
    if (!wave_live) sbeg = send = 0;
    if (!live) {
      len = 0;
      po0 = po1 = po2 = T(0);
    }
    const int nit = wave_live ? (send - sbeg) >> 6 : 0;
    const int nj = PAD ? nit : (live ? (len - lane / H + (64 / H) - 1) / (64 / H) : 0);
    const unsigned slot0 = (unsigned)(sbeg + lane);
    auto slot_of = [&](int j) -> unsigned { return j < nj ? slot0 + 64u * (unsigned)j : slot0; };
    unsigned ca = 0u, cb = 0u;
    if (wave_live) {
      ca = ld_at(col, slot_of(0) * 4u);
      cb = ld_at(col, slot_of(1) * 4u);
    }
    T s0 = T(0), s1 = T(0), s2 = T(0);
    if (wave_live) {
      constexpr unsigned ts = sizeof(T), vs = 3 * sizeof(T);
      const size_t plane = ns * ts;
      for (int j = 0; j < nit; j += 2) {
        const unsigned oa = slot_of(j) * ts, ob = slot_of(j + 1) * ts;
        const unsigned ga = ca * vs, gb = cb * vs;
        T ma[9], mb[9];
#pragma unroll
        for (int e = 0; e < 9; e++) {
          ma[e] = NT ? ld_stream(A, (size_t)e * plane, oa) : ld_at(A, (size_t)e * plane, oa);
          mb[e] = NT ? ld_stream(A, (size_t)e * plane, ob) : ld_at(A, (size_t)e * plane, ob);
        }
        const T qa0 = ld_at(pnew, 0, ga), qa1 = ld_at(pnew, ts, ga), qa2 = ld_at(pnew, 2 * ts, ga);
        const T qb0 = ld_at(pnew, 0, gb), qb1 = ld_at(pnew, ts, gb), qb2 = ld_at(pnew, 2 * ts, gb);
        ca = ld_at(col, slot_of(j + 2) * 4u);
        cb = ld_at(col, slot_of(j + 3) * 4u);
        const bool va = j < nj, vb2 = j + 1 < nj;
        const T pa0 = va ? qa0 : T(0), pa1 = va ? qa1 : T(0), pa2 = va ? qa2 : T(0);
        const T pb0 = vb2 ? qb0 : T(0), pb1 = vb2 ? qb1 : T(0), pb2 = vb2 ? qb2 : T(0);
        s0 = row3(s0, ma[0], ma[1], ma[2], pa0, pa1, pa2);
        s1 = row3(s1, ma[3], ma[4], ma[5], pa0, pa1, pa2);
        s2 = row3(s2, ma[6], ma[7], ma[8], pa0, pa1, pa2);
        s0 = row3(s0, mb[0], mb[1], mb[2], pb0, pb1, pb2);
        s1 = row3(s1, mb[3], mb[4], mb[5], pb0, pb1, pb2);
        s2 = row3(s2, mb[6], mb[7], mb[8], pb0, pb1, pb2);
        asm volatile("" : "+v"(ca), "+v"(cb));
      }
    }
#pragma unroll
    for (int off = H; off < 64; off <<= 1) {
      s0 += __shfl_xor(s0, off, 64);
      s1 += __shfl_xor(s1, off, 64);
      s2 += __shfl_xor(s2, off, 64);
    }
    double pt = 0.0;
    if (live && lane < H) pt = dot3((double)po0, (double)po1, (double)po2, (double)s0, (double)s1, (double)s2);
    if (wave_live) {
      const size_t base = 3 * (size_t)slice * H;
      const int nval = 3 * min(H, nrows - slice * H);
#pragma unroll
      for (int d0 = 0; d0 < 3 * H; d0 += 64) {
        const int d = d0 + lane;
        const int rho = d / 3;
        const int comp = d - 3 * rho;
        const int src = rho < H ? rho : 0;
        const T v0 = __shfl(s0, src, 64), v1 = __shfl(s1, src, 64), v2 = __shfl(s2, src, 64);
        if (d < nval) t[base + d] = comp == 0 ? v0 : comp == 1 ? v1 : v2;
      }
    }
    publish(pt, 0.0, ppt, nullptr, lb);
  }
}

// Local aggregate correction of update chunk ch (rows (ch * R + q) * 21 + lane / 3, component
// c = lane % 3 of each, as in the update kernel): g = P_I^T r over the chunk's free rows, summed
// in a fixed lane order to the same value in every lane, then z += (P_I^T A P_I)^-1 g at the free
// rows.  pn[q] != 0: row q is pinned or past the end.  Every lane of the wave takes part.
template <typename T, int R>
__device__ __forceinline__ void agg_correct(const HOBAK_GLOBAL T* ainv, int ch, int c, const T (&rq)[R],
                                            const unsigned char (&pn)[R], T (&zq)[R]) {
  T s = T(0);
#pragma unroll
  for (int q = 0; q < R; q++) s += pn[q] ? T(0) : rq[q];
  const T g0 = wave_sum(c == 0 ? s : T(0));
  const T g1 = wave_sum(c == 1 ? s : T(0));
  const T g2 = wave_sum(c == 2 ? s : T(0));
  const HOBAK_GLOBAL T* m = ainv + 9 * (size_t)ch + 3 * c;
  const T e = fma(m[2], g2, fma(m[1], g1, m[0] * g0));
#pragma unroll
  for (int q = 0; q < R; q++)
    if (!pn[q]) zq[q] += e;
}

// One lane per DOF of each of R groups of 21 rows (63 lanes) per wavefront, so that a row's
// three DOFs share a wave and the 3x3 preconditioner apply can read its siblings' residuals with
// cross-lane moves.  A wave covers 21 R consecutive rows and issues the loads of all R groups
// (the preconditioner's included) before it uses any, so each wave keeps R times the bytes in
// flight and a launch needs R times fewer waves.  Each instance's workgroups stride over its
// waves (at most a tuned cap per instance), which keeps the partial-sum count (and the work of
// every workgroup that re-reduces them) small.  The first wave's loads are issued before the
// p.t partials are reduced, so the two overlap.  A lane sums r.r and r.z over its groups in
// order, so the sums do not depend on the batch.
// CM bit 0: the coarse correction (InstDev::coarse) may be on for an instance of the batch; bit 1:
// local aggregates (InstDev::agg).  The variants without bit 1 keep the single pass over a wave's
// rows; with it, z is written after the chunk's correction.
template <typename T, int R, int CM>
__global__ __launch_bounds__(kBlock) void cg_update_kernel(Batch b, int parity) {
  constexpr bool CO = (CM & 1) != 0;
  constexpr bool LO = (CM & 2) != 0;
  for (int ws = blockIdx.x; ws < b.total; ws += gridDim.x) {
    const int vb = xcd_vb(ws, b.total);
    int lb, lg;
    const HOBAK_GLOBAL InstDev* __restrict__ I = seg_find(b, vb, lb, lg);
    const auto S = &ctl_of(I)->sc[parity ^ 1];
    if (S->done) continue;
    const double absNew = S->absNew;
    const int lane = threadIdx.x & 63;
    const int c = lane - 3 * (lane / 3);
    const int bse = lane - c;
    const int nrows = I->nv;
    const int nwaves = (nrows + 21 * R - 1) / (21 * R);
    const int stride = lg * (kBlock / 64);
    auto p = gbl((const T*)(parity ? I->p0 : I->p1));  // the p_new this iteration's SpMV wrote
    auto mi = gbl((const T*)I->minv);
    auto x = gbl((T*)I->x);
    auto rv_ = gbl((T*)I->r);
    auto z = gbl((T*)I->z);
    auto t = gbl((const T*)I->t);
    const int precond = I->precond;
    const int w0 = (int)((lb * (unsigned)kBlock + threadIdx.x) >> 6);
    const bool local = LO && I->agg != 0;
    const auto pin = gbl(I->pinned);
    T pv[R], tv[R], rv0[R], xv[R], m0[R], m1[R], m2[R];
    unsigned char pn[R];
    auto load = [&](int wave) {
#pragma unroll
      for (int q = 0; q < R; q++) {
        const int row = (wave * R + q) * 21 + lane / 3;
        pv[q] = tv[q] = rv0[q] = xv[q] = m0[q] = m1[q] = m2[q] = T(0);
        if constexpr (LO) pn[q] = 1;
        if (lane < 63 && row < nrows) {
          const size_t i = 3 * (size_t)row + c;
          if constexpr (LO) {
            if (local) pn[q] = pin[row];
          }
          pv[q] = p[i];
          tv[q] = t[i];
          rv0[q] = rv_[i];
          xv[q] = x[i];
          if (precond == 0) {
            m0[q] = mi[i];
          } else {
            auto m = mi + (size_t)kMinvStride * row + 3 * c;
            m0[q] = m[0];
            m1[q] = m[1];
            m2[q] = m[2];
          }
        }
      }
    };
    if (w0 < nwaves) load(w0);
    PreParts<> qp;
    pre_load<false>(gbl(I->part_pt), nullptr, I->nb[K_SPMV], qp);
    double pt, unused;
    pre_sum<false>(gbl(I->part_pt), nullptr, I->nb[K_SPMV], qp, pt, unused);
    const T alpha = (T)(absNew / pt);
    double rr2 = 0.0, rz = 0.0;
    double gc = 0.0;  // coarse correction: this lane's sum of r (component c)
    const bool coarse = CO && I->coarse != 0;
    for (int wave = w0; wave < nwaves; wave += stride) {
      if (wave != w0) load(wave);
      if constexpr (LO) {
        T rq[R], zq[R];
#pragma unroll
        for (int q = 0; q < R; q++) {
          const int row = (wave * R + q) * 21 + lane / 3;
          const bool live = lane < 63 && row < nrows;
          const size_t i = 3 * (size_t)row + c;
          T rv = T(0);
          if (live) {
            x[i] = fma(alpha, pv[q], xv[q]);
            rv = fma(-alpha, tv[q], rv0[q]);
            rv_[i] = rv;
          }
          const T r0 = __shfl(rv, bse, 64);
          const T r1 = __shfl(rv, bse + 1, 64);
          const T r2 = __shfl(rv, bse + 2, 64);
          rq[q] = rv;
          zq[q] = precond == 0 ? m0[q] * rv : fma(m2[q], r2, fma(m0[q], r0, m1[q] * r1));
        }
        if (local) agg_correct<T, R>(gbl((const T*)I->agg_inv), wave, c, rq, pn, zq);
#pragma unroll
        for (int q = 0; q < R; q++) {
          const int row = (wave * R + q) * 21 + lane / 3;
          if (lane < 63 && row < nrows) {
            z[3 * (size_t)row + c] = zq[q];
            rr2 += (double)rq[q] * (double)rq[q];
            rz += (double)rq[q] * (double)zq[q];
            gc += (double)rq[q];
          }
        }
      } else {
#pragma unroll
        for (int q = 0; q < R; q++) {
          const int row = (wave * R + q) * 21 + lane / 3;
          const bool live = lane < 63 && row < nrows;
          const size_t i = 3 * (size_t)row + c;
          T rv = T(0);
          if (live) {
            x[i] = fma(alpha, pv[q], xv[q]);
            rv = fma(-alpha, tv[q], rv0[q]);
            rv_[i] = rv;
          }
          const T r0 = __shfl(rv, bse, 64);
          const T r1 = __shfl(rv, bse + 1, 64);
          const T r2 = __shfl(rv, bse + 2, 64);
          if (live) {
            const T zv = precond == 0 ? m0[q] * rv : fma(m2[q], r2, fma(m0[q], r0, m1[q] * r1));
            z[i] = zv;
            rr2 += (double)rv * (double)rv;
            rz += (double)rv * (double)zv;
            gc += (double)rv;
          }
        }
      }
    }
    if (coarse) {
      // r.r, r.z and the three components' sums of r (P^T r) in one reduction
      double v[5] = {rr2, rz, c == 0 ? gc : 0.0, c == 1 ? gc : 0.0, c == 2 ? gc : 0.0};
      block_sum<5>(v);
      if (threadIdx.x == 0) {
        const int nupd = I->nb[K_UPD];
        gbl(I->part_rr)[lb] = v[0];
        gbl(I->part_rz)[lb] = v[1];
        auto pg = gbl(I->part_upd_g);
        pg[lb] = v[2];
        pg[nupd + lb] = v[3];
        pg[2 * (size_t)nupd + lb] = v[4];
      }
    } else {
      publish(rr2, rz, gbl(I->part_rr), gbl(I->part_rz), lb);
    }
  }
}

// The local aggregates' matrices P_I^T A P_I = sum over the chunk's free rows i of (dg_i - sum of
// A_ij over the row's columns j outside the chunk's free vertices), the operator the difference
// form applies to a translation of the chunk's free vertices, summed in FP64 with one wave per
// chunk, symmetrised and inverted (cofactors; zero for a chunk without free vertices).
template <typename T, int R>
__device__ void agg_build_chunks(const HOBAK_GLOBAL InstDev* I, int w0, int stride) {
  constexpr int G = 21 * R;
  const int lane = threadIdx.x & 63;
  const int nrows = I->nv;
  const int nch = (nrows + G - 1) / G;
  const int sh = I->shift;
  const int hm = (1 << sh) - 1;
  const size_t ns = (size_t)I->nslots;
  const auto A = gbl((const T*)I->A);
  const auto col = gbl(I->col);
  const auto pin = gbl(I->pinned);
  const auto dg = gbl(I->dg);
  const auto so = gbl(I->slice_off);
  const auto rl = gbl(I->row_len);
  auto out = gbl((T*)I->agg_inv);
  for (int ch = w0; ch < nch; ch += stride) {
    double a[9];
#pragma unroll
    for (int e = 0; e < 9; e++) a[e] = 0.0;
    for (int k0 = lane; k0 < G; k0 += 64) {
      const int row = ch * G + k0;
      if (row >= nrows || pin[row]) continue;
      a[0] += dg[3 * (size_t)row];
      a[4] += dg[3 * (size_t)row + 1];
      a[8] += dg[3 * (size_t)row + 2];
      const int sb = so[row >> sh] + (row & hm);
      const int len = rl[row];
      for (int k = 0; k < len; k++) {
        const int slot = sb + (k << sh);
        const int j = col[slot];
        if (j / G == ch && !pin[j]) continue;  // inside the aggregate (the own block included)
#pragma unroll
        for (int e = 0; e < 9; e++) a[e] -= (double)A[e * ns + slot];
      }
    }
#pragma unroll
    for (int e = 0; e < 9; e++) a[e] = wave_sum(a[e]);
    if (lane == 0) {
      const double s01 = 0.5 * (a[1] + a[3]), s02 = 0.5 * (a[2] + a[6]), s12 = 0.5 * (a[5] + a[7]);
      const double c00 = a[4] * a[8] - s12 * s12, c01 = s02 * s12 - s01 * a[8], c02 = s01 * s12 - s02 * a[4];
      const double c11 = a[0] * a[8] - s02 * s02, c12 = s01 * s02 - a[0] * s12, c22 = a[0] * a[4] - s01 * s01;
      const double det = a[0] * c00 + s01 * c01 + s02 * c02;
      const double id = (det > 0.0 && det < HUGE_VAL) ? 1.0 / det : 0.0;
      auto o = out + 9 * (size_t)ch;
      o[0] = (T)(c00 * id);
      o[1] = (T)(c01 * id);
      o[2] = (T)(c02 * id);
      o[3] = (T)(c01 * id);
      o[4] = (T)(c11 * id);
      o[5] = (T)(c12 * id);
      o[6] = (T)(c02 * id);
      o[7] = (T)(c12 * id);
      o[8] = (T)(c22 * id);
    }
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void agg_build_kernel(Batch b, int groups) {
  for (int ws = blockIdx.x; ws < b.total; ws += gridDim.x) {
    const int vb = xcd_vb(ws, b.total);
    int lb, lg;
    const HOBAK_GLOBAL InstDev* __restrict__ I = seg_find(b, vb, lb, lg);
    if (I->agg == 0) continue;
    const int w0 = (int)((lb * (unsigned)kBlock + threadIdx.x) >> 6);
    const int stride = lg * (kBlock / 64);
    if (groups == 1)
      agg_build_chunks<T, 1>(I, w0, stride);
    else if (groups == 2)
      agg_build_chunks<T, 2>(I, w0, stride);
    else
      agg_build_chunks<T, 4>(I, w0, stride);
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
  for (int ws = blockIdx.x; ws < b.total; ws += gridDim.x) {
    const int vb = xcd_vb(ws, b.total);
    int lb, lg;
    const HOBAK_GLOBAL InstDev* __restrict__ I = seg_find(b, vb, lb, lg);
    const int r = lb * kBlock + threadIdx.x;
    if (r < I->nv) {
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const size_t i = 3 * (size_t)r + c;
        gbl(I->x64)[i] = 0.0;
        gbl(I->res64)[i] = gbl(I->rhs64)[i];
      }
    }
    if (lb != 0) continue;
    double b2, unused;
    gather(gbl(I->part_row_a), nullptr, I->nb[K_ROW], b2, unused);
    if (threadIdx.x == 0) {
      auto st = &ctl_of(I)->rs;
      const int bad = nonfinite(b2) ? 1 : 0;
      st->b2 = b2;
      st->res2 = b2;
      st->tau = I->tau0;
      st->sweeps = 0;
      st->inner_iters = 0;
      st->first_iters = 0;
      st->status = bad ? 2 : 0;
      st->done = (b2 == 0.0 || bad) ? 1 : 0;
    }
  }
}

// The inner-solve seed of an instance with local aggregates: the update kernel's row mapping (one
// wave per chunk at a time), so that z = M^-1 r gains the chunk's correction before it is written.
// gs and ds: this lane's sums of r and dg (its component), for an instance that also has the
// coarse correction.
template <typename T, int R>
__device__ void agg_seed_chunks(const HOBAK_GLOBAL InstDev* I, int w0, int stride, double& rr, double& rz,
                                double& gs, double& ds) {
  const int lane = threadIdx.x & 63;
  const int c = lane - 3 * (lane / 3);
  const int bse = lane - c;
  const int nrows = I->nv;
  const int nch = (nrows + 21 * R - 1) / (21 * R);
  const auto res = gbl(I->res64);
  const auto mi = gbl((const T*)I->minv);
  const auto pin = gbl(I->pinned);
  auto rvec = gbl((T*)I->r);
  auto x = gbl((T*)I->x);
  auto z = gbl((T*)I->z);
  const int precond = I->precond;
  const bool coarse = I->coarse != 0;
  const auto dg = gbl(I->dg);
  for (int ch = w0; ch < nch; ch += stride) {
    T rq[R], zq[R];
    unsigned char pn[R];
#pragma unroll
    for (int q = 0; q < R; q++) {
      const int row = (ch * R + q) * 21 + lane / 3;
      const bool live = lane < 63 && row < nrows;
      const size_t i = 3 * (size_t)row + c;
      const T rv = live ? (T)res[i] : T(0);
      pn[q] = live ? pin[row] : 1;
      const T r0 = __shfl(rv, bse, 64);
      const T r1 = __shfl(rv, bse + 1, 64);
      const T r2 = __shfl(rv, bse + 2, 64);
      T zv = T(0);
      if (live) {
        if (precond == 0) {
          zv = mi[i] * rv;
        } else {
          auto m = mi + (size_t)kMinvStride * row + 3 * c;
          zv = fma(m[2], r2, fma(m[0], r0, m[1] * r1));
        }
      }
      rq[q] = rv;
      zq[q] = zv;
    }
    agg_correct<T, R>(gbl((const T*)I->agg_inv), ch, c, rq, pn, zq);
#pragma unroll
    for (int q = 0; q < R; q++) {
      const int row = (ch * R + q) * 21 + lane / 3;
      if (lane < 63 && row < nrows) {
        const size_t i = 3 * (size_t)row + c;
        rvec[i] = rq[q];
        x[i] = T(0);
        z[i] = zq[q];
        rr += (double)rq[q] * (double)rq[q];
        rz += (double)rq[q] * (double)zq[q];
        if (coarse) {
          gs += (double)rq[q];
          ds += dg[i];
        }
      }
    }
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void refine_init_kernel(Batch b, int gate, int groups) {
  for (int ws = blockIdx.x; ws < b.total; ws += gridDim.x) {
    const int vb = xcd_vb(ws, b.total);
    int lb, lg;
    const HOBAK_GLOBAL InstDev* __restrict__ I = seg_find(b, vb, lb, lg);
    if (gate && !solve_ended(ctl_of(I))) continue;
    if (sizeof(T) < sizeof(double) && I->agg != 0) {
      double rr = 0.0, rz = 0.0, gs = 0.0, ds = 0.0;
      const int w0 = (int)((lb * (unsigned)kBlock + threadIdx.x) >> 6);
      const int stride = lg * (kBlock / 64);
      if (groups == 1)
        agg_seed_chunks<T, 1>(I, w0, stride, rr, rz, gs, ds);
      else if (groups == 2)
        agg_seed_chunks<T, 2>(I, w0, stride, rr, rz, gs, ds);
      else
        agg_seed_chunks<T, 4>(I, w0, stride, rr, rz, gs, ds);
      if (I->coarse != 0) {  // both corrections: the coarse one's seed sums too (component c)
        const int c = (threadIdx.x & 63) % 3;
        double v[8] = {rr, rz, c == 0 ? gs : 0.0, c == 1 ? gs : 0.0, c == 2 ? gs : 0.0,
                       c == 0 ? ds : 0.0, c == 1 ? ds : 0.0, c == 2 ? ds : 0.0};
        block_sum<8>(v);
        if (threadIdx.x == 0) {
          const size_t nb = (size_t)I->nb[K_ROW];
          gbl(I->part_row_a)[lb] = v[0];
          gbl(I->part_row_b)[lb] = v[1];
          auto pc = gbl(I->part_row_c);
#pragma unroll
          for (int k = 0; k < 6; k++) pc[k * nb + lb] = v[2 + k];
        }
      } else {
        publish(rr, rz, gbl(I->part_row_a), gbl(I->part_row_b), lb);
      }
      continue;
    }
    const int r = lb * kBlock + threadIdx.x;
    double rr = 0.0, rz = 0.0;
    double gs[3] = {0.0, 0.0, 0.0}, ds[3] = {0.0, 0.0, 0.0};  // coarse correction: sums of r, dg
    const bool coarse = sizeof(T) < sizeof(double) && I->coarse != 0;
    if (r < I->nv) {
      T rv[3], zv[3];
#pragma unroll
      for (int c = 0; c < 3; c++) rv[c] = (T)gbl(I->res64)[3 * (size_t)r + c];
      apply_precond<T>(gbl((const T*)I->minv), I->precond, r, rv, zv);
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const size_t i = 3 * (size_t)r + c;
        (gbl((T*)I->r))[i] = rv[c];
        (gbl((T*)I->x))[i] = T(0);
        (gbl((T*)I->z))[i] = zv[c];
        rr += (double)rv[c] * (double)rv[c];
        rz += (double)rv[c] * (double)zv[c];
        if (coarse) {
          gs[c] = (double)rv[c];
          ds[c] = gbl(I->dg)[i];
        }
      }
    }
    if (coarse) {
      double v[8] = {rr, rz, gs[0], gs[1], gs[2], ds[0], ds[1], ds[2]};
      block_sum<8>(v);
      if (threadIdx.x == 0) {
        const size_t nb = (size_t)I->nb[K_ROW];
        gbl(I->part_row_a)[lb] = v[0];
        gbl(I->part_row_b)[lb] = v[1];
        auto pc = gbl(I->part_row_c);
#pragma unroll
        for (int k = 0; k < 6; k++) pc[k * nb + lb] = v[2 + k];
      }
    } else {
      publish(rr, rz, gbl(I->part_row_a), gbl(I->part_row_b), lb);
    }
  }
}

// One wavefront per H-row slice as in the SpMV; x64 + x is formed on the fly for every column.
template <typename T, int H>
__global__ __launch_bounds__(kBlock) void refine_residual_kernel(Batch b) {
  for (int ws = blockIdx.x; ws < b.total; ws += gridDim.x) {
    const int vb = xcd_vb(ws, b.total);
    int lb, lg;
    const HOBAK_GLOBAL InstDev* __restrict__ I = seg_find(b, vb, lb, lg);
    const auto ctl = ctl_of(I);
    if (!solve_ended(ctl) || ctl->rs.done) continue;
    constexpr int L = 64 / H;
    const int slice = (int)((lb * (unsigned)kBlock + threadIdx.x) >> 6);
    const int lane = threadIdx.x & 63;
    const int kk = lane / H;
    const int rr = lane % H;
    const int r = slice * H + rr;
    auto A = gbl((const T*)I->A);
    auto xi = gbl((const T*)I->x);
    auto x64 = gbl(I->x64);
    auto col = gbl(I->col);
    const size_t ns = (size_t)I->nslots;
    constexpr bool kDiff = sizeof(T) < sizeof(double);
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    double o0 = 0.0, o1 = 0.0, o2 = 0.0;  // the row's own x64 + x (difference form)
    const bool live = r < I->nv;
    const int nslices = (I->nv + H - 1) / H;
    if (slice < nslices) {
      // the slice's padded width, as in the SpMV
      const int sbeg = gbl(I->slice_off)[slice];
      const int nit = (gbl(I->slice_off)[slice + 1] - sbeg) >> 6;
      auto cp = col + sbeg + lane;
      auto Ap = A + sbeg + lane;
      if (kDiff && live) {
        const size_t i = 3 * (size_t)r;
        o0 = x64[i] + (double)xi[i];
        o1 = x64[i + 1] + (double)xi[i + 1];
        o2 = x64[i + 2] + (double)xi[i + 2];
      }
#pragma unroll 4
      for (int j = 0; j < nit; j++) {
        const int o = 64 * j;
        const size_t c = 3 * (size_t)cp[o];
        const double x0 = (x64[c] + (double)xi[c]) - o0;
        const double x1 = (x64[c + 1] + (double)xi[c + 1]) - o1;
        const double x2 = (x64[c + 2] + (double)xi[c + 2]) - o2;
        s0 = row3(s0, (double)Ap[0 * ns + o], (double)Ap[1 * ns + o], (double)Ap[2 * ns + o], x0, x1, x2);
        s1 = row3(s1, (double)Ap[3 * ns + o], (double)Ap[4 * ns + o], (double)Ap[5 * ns + o], x0, x1, x2);
        s2 = row3(s2, (double)Ap[6 * ns + o], (double)Ap[7 * ns + o], (double)Ap[8 * ns + o], x0, x1, x2);
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
        s0 = fma(gbl(I->dg)[i], o0, s0);
        s1 = fma(gbl(I->dg)[i + 1], o1, s1);
        s2 = fma(gbl(I->dg)[i + 2], o2, s2);
      }
      const double s[3] = {s0, s1, s2};
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const size_t i = 3 * (size_t)r + c;
        const double res = gbl(I->rhs64)[i] - s[c];
        gbl(I->res64)[i] = res;
        res2 += res * res;
      }
    }
    publish(res2, 0.0, gbl(I->part_res), nullptr, lb);
  }
}

// x64 += x after every residual workgroup has read the old x64: a separate pass keeps the
// on-the-fly sums above race-free.
template <typename T>
__global__ __launch_bounds__(kBlock) void refine_accumulate_kernel(Batch b) {
  for (int ws = blockIdx.x; ws < b.total; ws += gridDim.x) {
    const int vb = xcd_vb(ws, b.total);
    int lb, lg;
    const HOBAK_GLOBAL InstDev* __restrict__ I = seg_find(b, vb, lb, lg);
    const auto ctl = ctl_of(I);
    if (!solve_ended(ctl) || ctl->rs.done) continue;
    const int i = lb * kBlock + threadIdx.x;
    if (i < 3 * I->nv) gbl(I->x64)[i] += (double)(gbl((const T*)I->x))[i];
  }
}

// One workgroup per instance whose inner solve has just ended.  Plain form: the solve is
// finished.  Refinement: record the sweep and decide whether another one is needed; the next
// inner tolerance aims at the outer target with a 5x margin (0.2 tol |b| / |res|), floored at
// 1e-6 (FP32's reach) and capped at the first inner tolerance.
__global__ __launch_bounds__(kBlock) void finish_check_kernel(Batch b, int refine) {
  for (int ws = blockIdx.x; ws < b.total; ws += gridDim.x) {
    const int vb = xcd_vb(ws, b.total);
    int lb, lg;
    const HOBAK_GLOBAL InstDev* __restrict__ I = seg_find(b, vb, lb, lg);
    auto c = ctl_of(I);
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
    gather(gbl(I->part_res), nullptr, I->nb[K_SPMV], res2, unused);
    if (threadIdx.x != 0) continue;
    auto st = &c->rs;
    const double b2 = st->b2;
    const double tol = I->tol;
    const int sweeps = st->sweeps + 1;
    st->res2 = res2;
    st->sweeps = sweeps;
    st->inner_iters += c->sc[0].iters;
    if (sweeps == 1) st->first_iters = c->sc[0].iters;
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
  for (int ws = blockIdx.x; ws < b.total; ws += gridDim.x) {
    const int vb = xcd_vb(ws, b.total);
    int lb, lg;
    const HOBAK_GLOBAL InstDev* __restrict__ I = seg_find(b, vb, lb, lg);
    if (threadIdx.x != 0) continue;
    const auto c = ctl_of(I);
    const auto s = &c->sc[0];
    InstStatus o;
    o.phase = c->phase;
    o.iters = s->iters;
    o.inner_total = c->rs.inner_iters;
    o.sweeps = c->rs.sweeps;
    o.cg_status = s->status;
    o.ref_status = c->rs.status;
    o.ref_done = c->rs.done;
    o.first_iters = I->refine ? c->rs.first_iters : s->iters;
    o.spmv_work = c->spmv_work;
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
// (TET_MESH.cpp:756-767); output in FP64 into rhs64, or from sf_ext into out_ext.
template <typename T>
__global__ __launch_bounds__(kBlock) void force_gather_kernel(Batch b, const void* sf_ext, double* out_ext) {
  for (int ws = blockIdx.x; ws < b.total; ws += gridDim.x) {
    const int vb = xcd_vb(ws, b.total);
    int lb, lg;
    const HOBAK_GLOBAL InstDev* __restrict__ I = seg_find(b, vb, lb, lg);
    const int r = lb * kBlock + threadIdx.x;
    if (r >= I->nv) continue;
    auto sf = gbl((const T*)(sf_ext ? sf_ext : I->stage_f));
    auto out = gbl(out_ext ? out_ext : I->rhs64);
    double s[3];
    gather_forces(sf, gbl(I->fidx), gbl(I->vptr)[r], gbl(I->vptr)[r + 1], s);
#pragma unroll
    for (int c = 0; c < 3; c++) out[3 * (size_t)r + c] = s[c];
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

int blocks_of(int kind, int nv, int nt, int nslots, int shift, int update_cap, int update_groups) {
  switch (kind) {
    case K_ELEM: return cdiv(nt, kElementBlock);
    case K_SLOT: return cdiv(nslots, kBlock);
    case K_ROW: return cdiv(nv, kBlock);
    case K_DOF: return cdiv(3L * nv, kBlock);
    case K_SPMV: return cdiv((long)cdiv(nv, 1 << shift) * 64, kBlock);
    case K_UPD: {
      const int g = cdiv((long)cdiv(nv, 21 * update_groups) * 64, kBlock);
      return g < update_cap ? g : update_cap;
    }
    case K_ONE: return 1;
    case K_ASM: return cdiv(nslots, kAsmSlotsPerWave * (kBlock / 64));  // nslots: here the upper slots
    default: return 0;
  }
}

#define SEG_LAUNCH(kernel, b, cap, s, ...)                                                  \
  do {                                                                                      \
    if ((b).total <= 0) return hipSuccess;                                                  \
    hipLaunchKernelGGL(kernel, dim3(grid_of((b), (cap))), dim3(kBlock), 0, (s), (b), ##__VA_ARGS__); \
    return hipGetLastError();                                                               \
  } while (0)

hipError_t launch_contact_begin(const Batch& b, int cap, hipStream_t s) {
  SEG_LAUNCH(contact_begin_kernel, b, cap, s);
}
hipError_t launch_contact_sep(const Batch& b, int cap, hipStream_t s) { SEG_LAUNCH(contact_sep_kernel, b, cap, s); }
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
template <typename T, int H, bool PAD, bool NT, bool CO>
static hipError_t spmv_hp(const Batch& b, int parity, int cap, hipStream_t s, hipEvent_t start, hipEvent_t stop) {
  const dim3 g(grid_of(b, cap)), blk(kBlock);
  if (start)
    hipExtLaunchKernelGGL((cg_spmv_kernel<T, H, PAD, NT, CO>), g, blk, 0, s, start, stop, 0, b, parity);
  else
    hipLaunchKernelGGL((cg_spmv_kernel<T, H, PAD, NT, CO>), g, blk, 0, s, b, parity);
  return hipGetLastError();
}
template <typename T, int H, bool CO>
static hipError_t spmv_hc(const Batch& b, int parity, int cap, int pad, int stream, hipStream_t s,
                          hipEvent_t start, hipEvent_t stop) {
  if (stream)
    return pad ? spmv_hp<T, H, true, true, CO>(b, parity, cap, s, start, stop)
               : spmv_hp<T, H, false, true, CO>(b, parity, cap, s, start, stop);
  return pad ? spmv_hp<T, H, true, false, CO>(b, parity, cap, s, start, stop)
             : spmv_hp<T, H, false, false, CO>(b, parity, cap, s, start, stop);
}
// The coarse-correction variant exists for the reduced-precision form only.
template <typename T, int H>
static hipError_t spmv_h(const Batch& b, int parity, int cap, int pad, int stream, int coarse, hipStream_t s,
                         hipEvent_t start, hipEvent_t stop) {
  if constexpr (sizeof(T) < sizeof(double)) {
    if (coarse) return spmv_hc<T, H, true>(b, parity, cap, pad, stream, s, start, stop);
  }
  return spmv_hc<T, H, false>(b, parity, cap, pad, stream, s, start, stop);
}

template <int H, bool PAD, bool NT>
static hipError_t spmv_p_hp(const Batch& b, int parity, int cap, hipStream_t s, hipEvent_t start, hipEvent_t stop) {
  const dim3 g(grid_of(b, cap)), blk(kBlock);
  if (start)
    hipExtLaunchKernelGGL((cg_spmv_p_kernel<H, PAD, NT>), g, blk, 0, s, start, stop, 0, b, parity);
  else
    hipLaunchKernelGGL((cg_spmv_p_kernel<H, PAD, NT>), g, blk, 0, s, b, parity);
  return hipGetLastError();
}
template <int H>
static hipError_t spmv_p_h(const Batch& b, int parity, int cap, int pad, int stream, hipStream_t s,
                           hipEvent_t start, hipEvent_t stop) {
  if (stream)
    return pad ? spmv_p_hp<H, true, true>(b, parity, cap, s, start, stop)
               : spmv_p_hp<H, false, true>(b, parity, cap, s, start, stop);
  return pad ? spmv_p_hp<H, true, false>(b, parity, cap, s, start, stop)
             : spmv_p_hp<H, false, false>(b, parity, cap, s, start, stop);
}

hipError_t launch_cg_p(const Batch& upd, int parity, int cap, hipStream_t s) {
  SEG_LAUNCH(cg_p_kernel, upd, cap, s, parity);
}

hipError_t launch_cg_spmv_p(const Batch& b, int shift, int pad, int stream, int parity, int cap, hipStream_t s,
                            hipEvent_t start, hipEvent_t stop) {
  if (b.total <= 0) return hipSuccess;
  switch (shift) {
    case 3: return spmv_p_h<8>(b, parity, cap, pad, stream, s, start, stop);
    case 4: return spmv_p_h<16>(b, parity, cap, pad, stream, s, start, stop);
    case 5: return spmv_p_h<32>(b, parity, cap, pad, stream, s, start, stop);
    case 6: return spmv_p_h<64>(b, parity, cap, pad, stream, s, start, stop);
    default: return hipErrorInvalidValue;
  }
}

template <typename T>
hipError_t launch_cg_spmv(const Batch& b, int shift, int pad, int stream, int coarse, int parity, int cap,
                          hipStream_t s, hipEvent_t start, hipEvent_t stop) {
  if (b.total <= 0) return hipSuccess;
  switch (shift) {
    case 3: return spmv_h<T, 8>(b, parity, cap, pad, stream, coarse, s, start, stop);
    case 4: return spmv_h<T, 16>(b, parity, cap, pad, stream, coarse, s, start, stop);
    case 5: return spmv_h<T, 32>(b, parity, cap, pad, stream, coarse, s, start, stop);
    case 6: return spmv_h<T, 64>(b, parity, cap, pad, stream, coarse, s, start, stop);
    default: return hipErrorInvalidValue;
  }
}

template <typename T, int CM>
static hipError_t update_c(const Batch& b, int groups, int parity, int cap, hipStream_t s) {
  switch (groups) {
    case 1: SEG_LAUNCH((cg_update_kernel<T, 1, CM>), b, cap, s, parity);
    case 2: SEG_LAUNCH((cg_update_kernel<T, 2, CM>), b, cap, s, parity);
    case 4: SEG_LAUNCH((cg_update_kernel<T, 4, CM>), b, cap, s, parity);
    default: return hipErrorInvalidValue;
  }
}
template <typename T>
hipError_t launch_cg_update(const Batch& b, int groups, int coarse, int parity, int cap, hipStream_t s) {
  if constexpr (sizeof(T) < sizeof(double)) {
    switch (coarse & 3) {
      case 1: return update_c<T, 1>(b, groups, parity, cap, s);
      case 2: return update_c<T, 2>(b, groups, parity, cap, s);
      case 3: return update_c<T, 3>(b, groups, parity, cap, s);
      default: break;
    }
  }
  return update_c<T, 0>(b, groups, parity, cap, s);
}
template <typename T>
hipError_t launch_agg_build(const Batch& b, int groups, int cap, hipStream_t s) {
  if (groups != 1 && groups != 2 && groups != 4) return hipErrorInvalidValue;
  SEG_LAUNCH((agg_build_kernel<T>), b, cap, s, groups);
}
hipError_t launch_refine_begin(const Batch& b, int cap, hipStream_t s) { SEG_LAUNCH(refine_begin_kernel, b, cap, s); }
template <typename T>
hipError_t launch_refine_init(const Batch& b, int gate, int groups, int cap, hipStream_t s) {
  SEG_LAUNCH((refine_init_kernel<T>), b, cap, s, gate, groups);
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
hipError_t launch_force_gather(const Batch& b, int cap, hipStream_t s, const void* stage_f_ext,
                               double* out_ext) {
  SEG_LAUNCH((force_gather_kernel<T>), b, cap, s, stage_f_ext, out_ext);
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
  template hipError_t launch_cg_spmv<T>(const Batch&, int, int, int, int, int, int, hipStream_t, hipEvent_t, \
                                         hipEvent_t); \
  template hipError_t launch_cg_update<T>(const Batch&, int, int, int, int, hipStream_t);    \
  template hipError_t launch_refine_init<T>(const Batch&, int, int, int, hipStream_t);       \
  template hipError_t launch_agg_build<T>(const Batch&, int, int, hipStream_t);             \
  template hipError_t launch_refine_residual<T>(const Batch&, const Batch&, int, int, hipStream_t); \
  template hipError_t launch_force_gather<T>(const Batch&, int, hipStream_t, const void*, double*);
INST(double)
INST(float)
#undef INST

}  // namespace hobak
