// Host side of libhobak_hip.so: mesh set-up (sparsity, scatter map), device residency, step
// orchestration and the C ABI declared in include/hobak_c.h.
//
// Reference interfaces replaced (relative to the reference root):
//   TET_MESH::TET_MESH                    src/Geometry/TET_MESH.cpp:43-107 (volumes, one-rings, DmInv)
//   TET_MESH_FASTER::TET_MESH_FASTER      src/Geometry/TET_MESH_FASTER.cpp:30-84, 185-241
//                                         (fixed sparsity + per-entry gather lists)
//   TIMESTEPPER::initialize, buildMassMatrix, scaleMass, addGravity
//                                         src/Timestepper/TIMESTEPPER.cpp:48-92, 176-208, 334-345
//   BACKWARD_EULER_POSITION::solveRayleighDamped
//                                         src/Timestepper/BACKWARD_EULER_POSITION.cpp:138-327
//   BDF_1::solveRayleighDamped            src/Timestepper/BDF_1.cpp:78-239
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <set>
#include <string>
#include <vector>

#include "../../include/hobak_c.h"
#include "kernels.h"
#include "solver_kernels.h"

using namespace hobak;

struct hobak_ctx {
  int device = 0;
  int slice_rows = hobak::kDefaultSliceRows;  // tuning, applied to instances created afterwards
  int poll_chunk = 8;
  hipStream_t stream = nullptr;
  size_t budget = 0;
  size_t in_use = 0;
  std::string err;
  std::set<hobak_inst*> insts;
};

namespace {

struct Pattern {
  int nv = 0, nt = 0, nslices = 0, nslots = 0, nrows_pad = 0, nnzb = 0;
  int H = 16, shift = 4;  // rows per slice and its log2
  std::vector<int> row_ptr, col_idx;  // BSR
  std::vector<int> slice_off, row_len, col, slot_row, bsr_slot, diag_slot;
  std::vector<int> cptr, hdest;       // per slot contribution ranges; per (tet, x, y) position
  std::vector<int> vptr, fdest;       // per vertex force ranges; per (tet, x) position
};

void build_pattern(int nv, int nt, const int32_t* tets, int H, Pattern& P) {
  P.nv = nv;
  P.nt = nt;
  P.H = H;
  P.shift = 0;
  while ((1 << P.shift) < H) P.shift++;
  const int kSlice = H;
  const int kLanesPerRow = 64 / H;
  std::vector<std::vector<int>> nbr(nv);
  for (int t = 0; t < nt; t++)
    for (int a = 0; a < 4; a++)
      for (int b = 0; b < 4; b++) nbr[tets[4 * t + a]].push_back(tets[4 * t + b]);
  P.row_ptr.assign(nv + 1, 0);
  P.col_idx.clear();
  for (int i = 0; i < nv; i++) {
    auto& n = nbr[i];
    std::sort(n.begin(), n.end());
    n.erase(std::unique(n.begin(), n.end()), n.end());
    P.row_ptr[i + 1] = P.row_ptr[i] + (int)n.size();
    P.col_idx.insert(P.col_idx.end(), n.begin(), n.end());
    std::vector<int>().swap(n);
  }
  P.nnzb = P.row_ptr[nv];
  // SELL-H, k-interleaved: slice width rounded up to a multiple of the lanes per row so that
  // each slice's slots form whole 64-slot wave loads
  P.nslices = (nv + kSlice - 1) / kSlice;
  P.nrows_pad = P.nslices * kSlice;
  P.slice_off.assign(P.nslices + 1, 0);
  P.row_len.assign(nv, 0);
  for (int s = 0; s < P.nslices; s++) {
    int w = 0;
    for (int r = s * kSlice; r < std::min(nv, (s + 1) * kSlice); r++) {
      P.row_len[r] = P.row_ptr[r + 1] - P.row_ptr[r];
      w = std::max(w, P.row_len[r]);
    }
    w = (w + kLanesPerRow - 1) / kLanesPerRow * kLanesPerRow;
    P.slice_off[s + 1] = P.slice_off[s] + w * kSlice;
  }
  P.nslots = P.slice_off[P.nslices];
  P.col.assign(P.nslots, 0);
  P.slot_row.assign(P.nslots, -1);
  P.bsr_slot.assign(P.nnzb, 0);
  P.diag_slot.assign(nv, -1);
  for (int s = 0; s < P.nslices; s++) {
    const int w = (P.slice_off[s + 1] - P.slice_off[s]) / kSlice;
    for (int lane = 0; lane < kSlice; lane++) {
      const int r = s * kSlice + lane;
      for (int k = 0; k < w; k++) {
        const int slot = P.slice_off[s] + k * kSlice + lane;
        if (r < nv && k < P.row_len[r]) {
          const int j = P.col_idx[P.row_ptr[r] + k];
          P.col[slot] = j;
          P.slot_row[slot] = r;
          P.bsr_slot[P.row_ptr[r] + k] = slot;
          if (j == r) P.diag_slot[r] = slot;
        } else {
          P.col[slot] = std::min(r, nv - 1);
        }
      }
    }
  }
  auto slot_of = [&](int i, int j) {
    const int* b = P.col_idx.data() + P.row_ptr[i];
    const int* e = P.col_idx.data() + P.row_ptr[i + 1];
    const int k = (int)(std::lower_bound(b, e, j) - b);
    return P.slice_off[i / kSlice] + k * kSlice + (i % kSlice);
  };
  std::vector<int> cnt(P.nslots + 1, 0);
  std::vector<int> slots(16 * (size_t)nt);
  for (int t = 0; t < nt; t++)
    for (int x = 0; x < 4; x++)
      for (int y = 0; y < 4; y++) {
        const int s = slot_of(tets[4 * t + x], tets[4 * t + y]);
        slots[16 * (size_t)t + 4 * x + y] = s;
        cnt[s + 1]++;
      }
  P.cptr.assign(P.nslots + 1, 0);
  for (int s = 0; s < P.nslots; s++) P.cptr[s + 1] = P.cptr[s] + cnt[s + 1];
  // Position of every element block in its slot's contribution list: looping tets in ascending
  // order fills each list in ascending tet order (TET_MESH_FASTER.cpp:213-236 pushes the same way)
  P.hdest.assign(16 * (size_t)nt, 0);
  std::vector<int> fill(P.cptr.begin(), P.cptr.end() - 1);
  for (int t = 0; t < nt; t++)
    for (int x = 0; x < 4; x++)
      for (int y = 0; y < 4; y++) {
        const int s = slots[16 * (size_t)t + 4 * x + y];
        P.hdest[16 * (size_t)t + 4 * x + y] = fill[s]++;
      }
  std::vector<int> vc(nv + 1, 0);
  for (int t = 0; t < nt; t++)
    for (int a = 0; a < 4; a++) vc[tets[4 * t + a] + 1]++;
  P.vptr.assign(nv + 1, 0);
  for (int i = 0; i < nv; i++) P.vptr[i + 1] = P.vptr[i] + vc[i + 1];
  P.fdest.assign(4 * (size_t)nt, 0);
  std::vector<int> vf(P.vptr.begin(), P.vptr.end() - 1);
  for (int t = 0; t < nt; t++)
    for (int a = 0; a < 4; a++) P.fdest[4 * (size_t)t + a] = vf[tets[4 * t + a]]++;
}

}  // namespace

struct hobak_inst {
  hobak_ctx* ctx = nullptr;
  int nv = 0, nt = 0, dofs = 0;
  hobak_material mat{};
  hobak_step_params prm{};
  double lam_eff = 0, alpha = 0;
  int tsize = 8;  // bytes of the element/assembly/PCG real type
  Pattern P;
  std::vector<int> pinned_ids;
  std::vector<std::pair<void*, size_t>> allocs;
  size_t bytes = 0;

  // device, FP64
  double *rest = nullptr, *u = nullptr, *v = nullptr, *u_old = nullptr, *v_old = nullptr;
  double *dminv = nullptr, *vol = nullptr, *one_ring = nullptr, *mass = nullptr;
  double *pin_target = nullptr, *rhs64 = nullptr, *tmp64 = nullptr, *x64 = nullptr, *res64 = nullptr;
  void* kneg = nullptr;  // assembled -K_rest, T
  int4* tets = nullptr;
  int *slice_off = nullptr, *row_len = nullptr, *col = nullptr, *slot_row = nullptr;
  int *cptr = nullptr, *hdest = nullptr;
  int *vptr = nullptr, *fdest = nullptr;
  unsigned char* pinned = nullptr;
  // device, real type T
  void *stage_h = nullptr, *stage_f = nullptr, *A = nullptr, *D = nullptr, *E = nullptr, *C = nullptr;
  void *minv = nullptr, *r = nullptr, *x = nullptr, *z = nullptr, *t = nullptr;
  void *p0 = nullptr, *p1 = nullptr;
  RefineState* rstate = nullptr;       // FP64 refinement state (precision form)
  RefineState* rstate_host = nullptr;
  CgScalars* sc = nullptr;       // two ping-pong sets
  CgScalars* sc_host = nullptr;  // pinned host copy of one set
  // per-workgroup partial sums: rhs -> init, SpMV -> update, update -> p
  double *part_rhs_a = nullptr, *part_rhs_b = nullptr, *part_pt = nullptr;
  double *part_rr = nullptr, *part_rz = nullptr;
  int n_rhs = 0, n_spmv = 0, n_update = 0;
  int last_it = 0;  // iterations queued by the last solve
  hipEvent_t ev[6] = {};
  double last_ms[5] = {0, 0, 0, 0, 0};
  int prev_iters = -1;
  bool has_system = false;
  int poll_chunk = 8;
  // sampled SpMV timing (measurement hook)
  int sample_every = 0;
  std::vector<hipEvent_t> sample_ev;   // pairs
  double spmv_ms_sum = 0.0;
  long long spmv_count = 0;
};

namespace {

int fail(hobak_ctx* ctx, int code, const std::string& msg) {
  if (ctx) ctx->err = msg;
  return code;
}

#define HIPCHK(ctx, expr)                                                                   \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return fail((ctx), HOBAK_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <typename P>
int dalloc(hobak_inst* I, P** out, size_t count, size_t elem = sizeof(P)) {
  const size_t bytes = std::max<size_t>(count * elem, 16);
  if (I->ctx->budget && I->ctx->in_use + bytes > I->ctx->budget)
    return fail(I->ctx, HOBAK_E_NOMEM, "HBM budget exceeded");
  void* ptr = nullptr;
  if (hipMalloc(&ptr, bytes) != hipSuccess) return fail(I->ctx, HOBAK_E_NOMEM, "hipMalloc failed");
  if (hipMemsetAsync(ptr, 0, bytes, I->ctx->stream) != hipSuccess)
    return fail(I->ctx, HOBAK_E_HIP, "hipMemsetAsync failed");
  I->allocs.emplace_back(ptr, bytes);
  I->bytes += bytes;
  I->ctx->in_use += bytes;
  *out = (P*)ptr;
  return HOBAK_OK;
}

template <typename P>
int upload(hobak_inst* I, P* dst, const P* src, size_t count) {
  if (hipMemcpyAsync(dst, src, count * sizeof(P), hipMemcpyHostToDevice, I->ctx->stream) != hipSuccess)
    return fail(I->ctx, HOBAK_E_HIP, "host-to-device copy failed");
  return HOBAK_OK;
}

#define TRY(expr)              \
  do {                         \
    int rc_ = (expr);          \
    if (rc_ != HOBAK_OK) return rc_; \
  } while (0)

size_t estimate_bytes(int nv, int nt, int nnzb_est, int nslots_est, int precision) {
  const size_t ts = precision == HOBAK_PRECISION_FP64 ? 8 : 4;
  size_t b = 0;
  b += (size_t)nv * 3 * 8 * 10;                 // rest, u, v, olds, target, rhs, tmp, x64, res64, mass
  b += (size_t)nt * (16 + 9 * 8 + 8);           // tets, dminv, vol
  b += (size_t)nslots_est * (4 + 4 + 9 * ts);   // col + slot_row + kneg
  b += (size_t)nslots_est * 9 * ts * 4;         // A, D, E, C
  b += (size_t)nslots_est * 4 + (size_t)nt * 20 * 4;  // cptr + hdest + fdest
  b += (size_t)nt * (16 * 9 + 12) * ts;         // staging
  b += (size_t)nv * (9 + 15) * ts;              // minv + CG vectors
  b += (size_t)nt * 4 * 4 + (size_t)nv * 16;    // vertex lists, misc
  (void)nnzb_est;
  return b;
}

// ---- kernel argument builders ----
ElementArgs element_args(hobak_inst* I, const double* u, int with_h) {
  ElementArgs a{};
  a.nt = I->nt;
  a.material = I->mat.kind;
  a.mu = I->mat.mu;
  a.lam = I->lam_eff;
  a.alpha = I->alpha;
  a.tets = I->tets;
  a.rest = I->rest;
  a.u = u;
  a.dminv = I->dminv;
  a.vol = I->vol;
  a.stage_h = I->stage_h;
  a.stage_f = I->stage_f;
  a.hdest = I->hdest;
  a.fdest = I->fdest;
  a.with_hessian = with_h;
  return a;
}

AsmArgs asm_args(hobak_inst* I, int mode, void* out) {
  AsmArgs a{};
  a.nrows = I->nv;
  a.nslots = I->P.nslots;
  a.nt = I->nt;
  a.slot_row = I->slot_row;
  a.col = I->col;
  a.cptr = I->cptr;
  a.stage_h = I->stage_h;
  a.D = I->D;
  a.out = out;
  a.mode = mode;
  a.pinned = I->pinned;
  a.minv = I->minv;
  a.precond = I->prm.precond;
  a.ground_on = I->prm.ground_on;
  a.ground_y = I->prm.ground_y;
  a.ground_k = I->prm.ground_k;
  a.rest = I->rest;
  a.u = I->u;
  return a;
}

CgArgs cg_args(hobak_inst* I) {
  CgArgs a{};
  a.nrows = I->nv;
  a.nslots = I->P.nslots;
  a.slice_shift = I->P.shift;
  a.slice_off = I->slice_off;
  a.row_len = I->row_len;
  a.col = I->col;
  a.A = I->A;
  a.x = I->x;
  a.r = I->r;
  a.z = I->z;
  a.p0 = I->p0;
  a.p1 = I->p1;
  a.t = I->t;
  a.minv = I->minv;
  a.precond = I->prm.precond;
  a.sc = I->sc;
  a.it = 0;
  a.pt = Partials{I->part_pt, nullptr, I->n_spmv};
  a.rz = Partials{I->part_rr, I->part_rz, I->n_update};
  return a;
}

InitArgs init_args(hobak_inst* I) {
  InitArgs a{};
  a.in = Partials{I->part_rhs_a, I->part_rhs_b, I->n_rhs};
  a.sc = I->sc;
  a.tol = I->prm.pcg_tol > 0 ? I->prm.pcg_tol : 2.220446049250313e-16;
  a.max_iters = I->prm.pcg_max_iters > 0 ? I->prm.pcg_max_iters : 2 * I->dofs;
  return a;
}

RefineArgs refine_args(hobak_inst* I, bool residual_parts) {
  RefineArgs a{};
  a.nrows = I->nv;
  a.nslots = I->P.nslots;
  a.slice_shift = I->P.shift;
  a.slice_off = I->slice_off;
  a.row_len = I->row_len;
  a.col = I->col;
  a.A = I->A;
  a.rhs64 = I->rhs64;
  a.x64 = I->x64;
  a.res64 = I->res64;
  a.r = I->r;
  a.x = I->x;
  a.z = I->z;
  a.p0 = I->p0;
  a.minv = I->minv;
  a.precond = I->prm.precond;
  // the inner-solve seed publishes one partial per rhs-shaped workgroup, the residual one per
  // SpMV-shaped workgroup; both reuse arrays that are consumed before they are overwritten
  a.out = residual_parts ? Partials{I->part_pt, nullptr, I->n_spmv}
                         : Partials{I->part_rhs_a, I->part_rhs_b, I->n_rhs};
  a.state = I->rstate;
  a.tol = I->prm.pcg_tol > 0 ? I->prm.pcg_tol : 2.220446049250313e-16;
  return a;
}

RhsArgs rhs_args(hobak_inst* I, int bdf) {
  RhsArgs a{};
  a.nrows = I->nv;
  a.nslots = I->P.nslots;
  a.nt = I->nt;
  a.slice_shift = I->P.shift;
  a.slice_off = I->slice_off;
  a.row_len = I->row_len;
  a.col = I->col;
  a.vptr = I->vptr;
  a.stage_f = I->stage_f;
  a.Mvec = bdf ? I->C : I->E;
  a.vel = I->v;
  a.u = I->u;
  a.uOld = I->u_old;
  a.vOld = I->v_old;
  a.mass = I->mass;
  a.one_ring = I->one_ring;
  a.rest = I->rest;
  for (int c = 0; c < 3; c++) a.g[c] = I->prm.gravity[c];
  a.invDt = 1.0 / I->prm.dt;
  a.invDt2 = a.invDt * a.invDt;
  a.bdf = bdf;
  a.pinned = I->pinned;
  a.ground_on = I->prm.ground_on;
  a.ground_y = I->prm.ground_y;
  a.ground_k = I->prm.ground_k;
  a.rhs64 = I->rhs64;
  a.init_cg = 1;
  a.r = I->r;
  a.x = I->x;
  a.z = I->z;
  a.p0 = I->p0;
  a.minv = I->minv;
  a.precond = I->prm.precond;
  a.out = Partials{I->part_rhs_a, I->part_rhs_b, I->n_rhs};
  return a;
}

template <typename T>
int setup_constant_parts(hobak_inst* I) {
  hipStream_t s = I->ctx->stream;
  // rest-state stiffness (u = 0): TIMESTEPPER::buildRayleighDampingMatrix (TIMESTEPPER.cpp:213-236)
  HIPCHK(I->ctx, hipMemsetAsync(I->tmp64, 0, sizeof(double) * I->dofs, s));
  HIPCHK(I->ctx, launch_element<T>(element_args(I, I->tmp64, 1), s));
  HIPCHK(I->ctx, launch_assemble<T>(asm_args(I, 0, I->kneg), s));
  ConstArgs c{};
  c.nslots = I->P.nslots;
  c.slot_row = I->slot_row;
  c.col = I->col;
  c.kneg = I->kneg;
  c.mass = I->mass;
  c.ra = I->prm.rayleigh_alpha;
  c.rb = I->prm.rayleigh_beta;
  c.invDt = 1.0 / I->prm.dt;
  c.invDt2 = c.invDt * c.invDt;
  c.D = I->D;
  c.E = I->E;
  c.C = I->C;
  HIPCHK(I->ctx, launch_constant_parts<T>(c, s));
  return HOBAK_OK;
}

// Each partial array holds one entry per workgroup of its producing kernel; refuse to launch if
// the grids (functions of nv and the slice height) no longer match the arrays sized at creation.
int check_partials(hobak_inst* I) {
  if (rhs_grid(I->nv) != I->n_rhs || spmv_grid(I->nv, I->P.shift) != I->n_spmv ||
      update_grid(I->nv) != I->n_update)
    return fail(I->ctx, HOBAK_E_INVALID, "partial-sum arrays do not match the launch grids");
  return HOBAK_OK;
}

// Run PCG chunks until the device says the solve is done.  Reads back the scalars.
template <typename T>
int run_pcg(hobak_inst* I) {
  hipStream_t s = I->ctx->stream;
  CgArgs a = cg_args(I);
  const int maxIt = I->prm.pcg_max_iters > 0 ? I->prm.pcg_max_iters : 2 * I->dofs;
  int launched = 0;
  int chunk = I->prev_iters > I->poll_chunk ? I->prev_iters - I->poll_chunk / 2 : I->poll_chunk;
  // sampled launches: (iteration index, event pair slot)
  std::vector<std::pair<int, int>> samples;
  const int max_samples = (int)I->sample_ev.size() / 2;
  while (true) {
    for (int k = 0; k < chunk; k++) {
      a.it = launched + k;
      const bool sample = I->sample_every > 0 && (a.it % I->sample_every) == 0 &&
                          (int)samples.size() < max_samples;
      if (sample) {
        const int slot = (int)samples.size();
        HIPCHK(I->ctx, launch_cg_spmv<T>(a, s, I->sample_ev[2 * slot], I->sample_ev[2 * slot + 1]));
        samples.emplace_back(a.it, slot);
        HIPCHK(I->ctx, launch_cg_rest<T>(a, s));
      } else {
        HIPCHK(I->ctx, launch_cg_iteration<T>(a, s));
      }
    }
    launched += chunk;
    // after `launched` iterations the current scalars are in set launched & 1
    HIPCHK(I->ctx, hipMemcpyAsync(I->sc_host, I->sc + (launched & 1), sizeof(CgScalars),
                                  hipMemcpyDeviceToHost, s));
    HIPCHK(I->ctx, hipStreamSynchronize(s));
    if (I->sc_host->done) break;
    if (launched > maxIt + 64) return fail(I->ctx, HOBAK_E_HIP, "PCG loop did not terminate");
    chunk = I->poll_chunk;
  }
  I->last_it = launched;
  // SpMV launches 0..iters did work (the converging pass included); later ones returned at once.
  for (auto& sm : samples) {
    if (sm.first > I->sc_host->iters) continue;
    float ms = 0.f;
    HIPCHK(I->ctx, hipEventElapsedTime(&ms, I->sample_ev[2 * sm.second], I->sample_ev[2 * sm.second + 1]));
    I->spmv_ms_sum += ms;
    I->spmv_count += 1;
  }
  return HOBAK_OK;
}

// FP32 PCG inside FP64 iterative refinement: returns inner iterations, sweeps, status.
template <typename T>
int solve_refined(hobak_inst* I, int& iters, int& sweeps, double& err, int& status) {
  hipStream_t s = I->ctx->stream;
  hobak_ctx* ctx = I->ctx;
  const int max_sweeps = 32;
  iters = 0;
  sweeps = 0;
  status = HOBAK_OK;
  const RefineArgs ri = refine_args(I, false);
  const RefineArgs rr = refine_args(I, true);
  HIPCHK(ctx, launch_refine_begin(ri, Partials{I->part_rhs_a, nullptr, I->n_rhs}, s));
  HIPCHK(ctx, hipMemcpyAsync(I->rstate_host, I->rstate, sizeof(RefineState), hipMemcpyDeviceToHost, s));
  HIPCHK(ctx, hipStreamSynchronize(s));
  InitArgs ia = init_args(I);
  const double tau0 = I->prm.refine_inner_tol > 0 ? I->prm.refine_inner_tol : 1e-4;
  const double outer = I->prm.pcg_tol > 0 ? I->prm.pcg_tol : 2.220446049250313e-16;
  ia.tol = tau0;
  while (!I->rstate_host->done && sweeps < max_sweeps) {
    // After the first sweep, aim each correction solve at the outer target with a 5x margin
    // (0.2 tol |b| / |res|), floored at FP32's reach, so the last sweep overshoots the target
    // instead of stopping just short of it.
    if (sweeps > 0 && I->rstate_host->res2 > 0) {
      const double want = 0.2 * outer * std::sqrt(I->rstate_host->b2 / I->rstate_host->res2);
      ia.tol = std::min(tau0, std::max(want, 1e-6));
    }
    HIPCHK(ctx, launch_refine_init<T>(ri, s));
    HIPCHK(ctx, launch_cg_init(ia, s));
    TRY(run_pcg<T>(I));
    iters += I->sc_host->iters;
    if (I->sc_host->status == 2) status = HOBAK_E_NONFINITE;
    HIPCHK(ctx, launch_refine_residual<T>(rr, s));
    HIPCHK(ctx, launch_refine_check(rr, s));
    HIPCHK(ctx, hipMemcpyAsync(I->rstate_host, I->rstate, sizeof(RefineState), hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipStreamSynchronize(s));
    sweeps++;
    if (status != HOBAK_OK) break;
  }
  const RefineState& h = *I->rstate_host;
  err = h.b2 > 0 ? std::sqrt(h.res2 / h.b2) : 0.0;
  if (h.status == 2) status = HOBAK_E_NONFINITE;
  else if (!h.done && status == HOBAK_OK) status = HOBAK_E_NOT_CONVERGED;
  return HOBAK_OK;
}

template <typename T>
int step_impl(hobak_inst* I, hobak_step_stats* st) {
  hipStream_t s = I->ctx->stream;
  hobak_ctx* ctx = I->ctx;
  const int N = std::max(1, std::min(I->prm.newton_iters, HOBAK_MAX_NEWTON));
  const int bdf = N > 1;
  const bool refine = I->prm.precision == HOBAK_PRECISION_FP32_REFINE;
  const double invDt = 1.0 / I->prm.dt;
  double ms[5] = {0, 0, 0, 0, 0};
  TRY(check_partials(I));
  HIPCHK(ctx, launch_begin_step(I->nv, I->u, I->v, I->u_old, I->v_old, I->pinned, I->pin_target, s));
  int done = 0;
  double maxR = 10 * I->prm.newton_tol;
  int total = 0;
  int status = HOBAK_OK;
  while (done < N && (N == 1 || maxR > I->prm.newton_tol)) {
    HIPCHK(ctx, hipEventRecord(I->ev[0], s));
    HIPCHK(ctx, launch_element<T>(element_args(I, I->u, 1), s));
    HIPCHK(ctx, hipEventRecord(I->ev[1], s));
    HIPCHK(ctx, launch_assemble<T>(asm_args(I, 1, I->A), s));
    HIPCHK(ctx, hipEventRecord(I->ev[2], s));
    RhsArgs ra = rhs_args(I, bdf);
    ra.init_cg = refine ? 0 : 1;
    HIPCHK(ctx, launch_rhs_init<T>(ra, s));
    int iters = 0, sweeps = 0, st_solve = HOBAK_OK;
    double err = 0.0;
    if (refine) {
      HIPCHK(ctx, hipEventRecord(I->ev[3], s));
      TRY(solve_refined<T>(I, iters, sweeps, err, st_solve));
      maxR = I->rstate_host->b2;
      HIPCHK(ctx, hipEventRecord(I->ev[4], s));
      HIPCHK(ctx, launch_apply_dx<double>(I->dofs, I->u, I->x64, s));
    } else {
      HIPCHK(ctx, launch_cg_init(init_args(I), s));
      HIPCHK(ctx, hipEventRecord(I->ev[3], s));
      TRY(run_pcg<T>(I));
      HIPCHK(ctx, hipEventRecord(I->ev[4], s));
      HIPCHK(ctx, launch_apply_dx<T>(I->dofs, I->u, (const T*)I->x, s));
      const CgScalars& h = *I->sc_host;
      maxR = h.maxR;
      iters = h.iters;
      err = h.rhsNorm2 > 0 ? std::sqrt(h.resNorm2 / h.rhsNorm2) : 0.0;
      if (h.status == 2) st_solve = HOBAK_E_NONFINITE;
      else if (h.status == 1) st_solve = HOBAK_E_NOT_CONVERGED;
    }
    float e01, e12, e23, e34;
    HIPCHK(ctx, hipEventSynchronize(I->ev[4]));
    (void)hipEventElapsedTime(&e01, I->ev[0], I->ev[1]);
    (void)hipEventElapsedTime(&e12, I->ev[1], I->ev[2]);
    (void)hipEventElapsedTime(&e23, I->ev[2], I->ev[3]);
    (void)hipEventElapsedTime(&e34, I->ev[3], I->ev[4]);
    ms[0] += e01;
    ms[1] += e12;
    ms[2] += e23;
    ms[3] += e34;
    if (st) {
      st->pcg_iters[done] = iters;
      st->refine_iters[done] = sweeps;
      st->pcg_error[done] = err;
      st->newton_residual[done] = maxR;
    }
    I->prev_iters = refine ? -1 : iters;
    total += iters;
    if (st_solve == HOBAK_E_NONFINITE) status = HOBAK_E_NONFINITE;
    else if (st_solve == HOBAK_E_NOT_CONVERGED && status == HOBAK_OK) status = HOBAK_E_NOT_CONVERGED;
    done++;
  }
  HIPCHK(ctx, hipEventRecord(I->ev[0], s));
  HIPCHK(ctx, launch_end_step(I->nv, invDt, I->u, I->u_old, I->v, I->pinned, s));
  HIPCHK(ctx, hipEventRecord(I->ev[5], s));
  HIPCHK(ctx, hipEventSynchronize(I->ev[5]));
  float e;
  (void)hipEventElapsedTime(&e, I->ev[0], I->ev[5]);
  ms[4] = e;
  std::memcpy(I->last_ms, ms, sizeof(ms));
  I->has_system = true;
  if (st) {
    st->newton_iters = done;
    st->total_pcg_iters = total;
    st->status = status;
  }
  return status;
}

void free_inst(hobak_inst* I) {
  for (auto& a : I->allocs) {
    (void)hipFree(a.first);
    if (I->ctx) I->ctx->in_use -= a.second;
  }
  I->allocs.clear();
  if (I->sc_host) (void)hipHostFree(I->sc_host);
  if (I->rstate_host) (void)hipHostFree(I->rstate_host);
  for (auto& e : I->sample_ev)
    if (e) (void)hipEventDestroy(e);
  I->sample_ev.clear();
  for (auto& e : I->ev)
    if (e) (void)hipEventDestroy(e);
}

template <typename T>
int download_sell_as_bsr(hobak_inst* I, const void* dev, double* out, double sign) {
  const size_t ns = I->P.nslots;
  std::vector<T> h(9 * ns);
  HIPCHK(I->ctx, hipMemcpyAsync(h.data(), dev, sizeof(T) * 9 * ns, hipMemcpyDeviceToHost, I->ctx->stream));
  HIPCHK(I->ctx, hipStreamSynchronize(I->ctx->stream));
  for (int k = 0; k < I->P.nnzb; k++) {
    const int slot = I->P.bsr_slot[k];
    for (int e = 0; e < 9; e++) out[9 * (size_t)k + e] = sign * (double)h[e * ns + slot];
  }
  return HOBAK_OK;
}

}  // namespace

extern "C" {

void hobak_default_params(hobak_step_params* p) {
  std::memset(p, 0, sizeof(*p));
  p->dt = 1.0 / 60.0;          // BACKWARD_EULER_POSITION.cpp:54
  p->rayleigh_alpha = 0.01;    // :55
  p->rayleigh_beta = 0.01;     // :56
  p->mass_scale = 1.0;
  p->pcg_tol = 0.0;            // Eigen default: machine epsilon
  p->pcg_max_iters = 0;        // Eigen default: 2 * DOFs
  p->precond = HOBAK_PRECOND_JACOBI;
  p->newton_iters = 1;
  p->newton_tol = 1e-7;        // NEWMARK.cpp:43
  p->precision = HOBAK_PRECISION_FP64;
  p->refine_inner_tol = 1e-4;
}

int hobak_ctx_create(int device, size_t hbm_budget_bytes, hobak_ctx** out) {
  if (!out) return HOBAK_E_INVALID;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= device || device < 0) return HOBAK_E_NODEVICE;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return HOBAK_E_NODEVICE;
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return HOBAK_E_NODEVICE;
  if (hipSetDevice(device) != hipSuccess) return HOBAK_E_NODEVICE;
  hobak_ctx* ctx = new (std::nothrow) hobak_ctx;
  if (!ctx) return HOBAK_E_NOMEM;
  ctx->device = device;
  ctx->budget = hbm_budget_bytes;
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
    delete ctx;
    return HOBAK_E_HIP;
  }
  *out = ctx;
  return HOBAK_OK;
}

void hobak_ctx_destroy(hobak_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  std::vector<hobak_inst*> all(ctx->insts.begin(), ctx->insts.end());
  for (auto* I : all) hobak_inst_destroy(I);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

const char* hobak_last_error(hobak_ctx* ctx) { return ctx ? ctx->err.c_str() : "no context"; }

size_t hobak_ctx_bytes_in_use(hobak_ctx* ctx) { return ctx ? ctx->in_use : 0; }

size_t hobak_inst_bytes_estimate(int32_t nv, int32_t nt, int32_t precision) {
  const size_t nnzb = (size_t)nv * 16;
  return estimate_bytes(nv, nt, (int)nnzb, (int)(nnzb * 1.15), precision);
}

int hobak_inst_create(hobak_ctx* ctx, const double* rest_xyz, int32_t nv, const int32_t* tets,
                      int32_t nt, const hobak_material* material, const hobak_step_params* params,
                      hobak_inst** out) {
  if (!ctx || !out || !rest_xyz || !tets || !material || !params) return HOBAK_E_INVALID;
  *out = nullptr;
  if (nv <= 0 || nt <= 0) return fail(ctx, HOBAK_E_INVALID, "empty mesh");
  if ((size_t)nt * 32 >= (size_t)INT32_MAX) return fail(ctx, HOBAK_E_INVALID, "too many tets");
  if (material->kind != HOBAK_MAT_SNH && material->kind != HOBAK_MAT_ARAP)
    return fail(ctx, HOBAK_E_INVALID, "unknown material");
  if (!(params->dt > 0)) return fail(ctx, HOBAK_E_INVALID, "dt must be positive");
  if (params->precond != HOBAK_PRECOND_JACOBI && params->precond != HOBAK_PRECOND_BLOCK_JACOBI)
    return fail(ctx, HOBAK_E_INVALID, "unknown preconditioner");
  if (params->precision != HOBAK_PRECISION_FP64 && params->precision != HOBAK_PRECISION_FP32_REFINE)
    return fail(ctx, HOBAK_E_INVALID, "unknown precision form");
  for (size_t i = 0; i < 4 * (size_t)nt; i++)
    if (tets[i] < 0 || tets[i] >= nv) return fail(ctx, HOBAK_E_INVALID, "tet index out of range");
  // rest volumes must be positive (TET_MESH.cpp:132-137)
  std::vector<double> vol(nt), dminv(9 * (size_t)nt), one_ring(nv, 0.0);
  for (int t = 0; t < nt; t++) {
    const int32_t* T = tets + 4 * (size_t)t;
    double d[3][3];
    for (int k = 0; k < 3; k++)
      for (int c = 0; c < 3; c++) d[k][c] = rest_xyz[3 * (size_t)T[k + 1] + c] - rest_xyz[3 * (size_t)T[0] + c];
    const double cx = d[0][1] * d[1][2] - d[0][2] * d[1][1];
    const double cy = d[0][2] * d[1][0] - d[0][0] * d[1][2];
    const double cz = d[0][0] * d[1][1] - d[0][1] * d[1][0];
    vol[t] = (d[2][0] * cx + d[2][1] * cy + d[2][2] * cz) / 6.0;
    if (!(vol[t] > 0.0)) return fail(ctx, HOBAK_E_INVALID, "tet with non-positive rest volume");
    // Dm = [d0 d1 d2] as columns; inverse by cofactors (Eigen InverseImpl.h:126-169 order)
    double m[3][3];
    for (int r = 0; r < 3; r++)
      for (int c = 0; c < 3; c++) m[r][c] = d[c][r];
    auto cof = [&](int i, int j) {
      const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
      return m[i1][j1] * m[i2][j2] - m[i1][j2] * m[i2][j1];
    };
    const double c0 = cof(0, 0), c1 = cof(1, 0), c2 = cof(2, 0);
    const double det = (c0 * m[0][0] + c1 * m[1][0]) + c2 * m[2][0];
    const double invdet = 1.0 / det;
    double inv[3][3];
    inv[0][0] = c0 * invdet;
    inv[0][1] = c1 * invdet;
    inv[0][2] = c2 * invdet;
    for (int row = 1; row < 3; row++)
      for (int k = 0; k < 3; k++) inv[row][k] = cof(k, row) * invdet;
    for (int r = 0; r < 3; r++)
      for (int c = 0; c < 3; c++) dminv[(3 * r + c) * (size_t)nt + t] = inv[r][c];
  }
  for (int t = 0; t < nt; t++) {  // one-ring volumes, ascending tet (TET_MESH.cpp:156-173)
    const double q = 0.25 * vol[t];
    for (int k = 0; k < 4; k++) one_ring[tets[4 * (size_t)t + k]] += q;
  }
  (void)hipSetDevice(ctx->device);
  hobak_inst* I = new (std::nothrow) hobak_inst;
  if (!I) return fail(ctx, HOBAK_E_NOMEM, "host allocation failed");
  I->ctx = ctx;
  I->nv = nv;
  I->nt = nt;
  I->dofs = 3 * nv;
  I->mat = *material;
  I->prm = *params;
  I->prm.newton_iters = std::max(1, std::min(params->newton_iters, HOBAK_MAX_NEWTON));
  if (material->kind == HOBAK_MAT_SNH) {  // SNH.cpp:23-32 reparameterisation
    I->lam_eff = material->lambda + material->mu;
    I->alpha = 1.0 + material->mu / I->lam_eff;
  } else {
    I->lam_eff = material->lambda;
    I->alpha = 0.0;
  }
  I->tsize = params->precision == HOBAK_PRECISION_FP64 ? 8 : 4;
  build_pattern(nv, nt, tets, ctx->slice_rows, I->P);
  I->poll_chunk = ctx->poll_chunk;
  const Pattern& P = I->P;
  const size_t need = estimate_bytes(nv, nt, P.nnzb, P.nslots, params->precision);
  if (ctx->budget && ctx->in_use + need > ctx->budget) {
    delete I;
    return fail(ctx, HOBAK_E_NOMEM, "HBM budget would be exceeded");
  }
  auto bail = [&](int rc) {
    free_inst(I);
    delete I;
    return rc;
  };
  const size_t ts = I->tsize;
  const size_t np = P.nrows_pad;
  int rc = HOBAK_OK;
#define ALLOC(ptr, n) \
  if ((rc = dalloc(I, &I->ptr, (n))) != HOBAK_OK) return bail(rc)
#define ALLOCT(ptr, n) \
  if ((rc = dalloc(I, (char**)&I->ptr, (n), ts)) != HOBAK_OK) return bail(rc)
  ALLOC(rest, 3 * (size_t)nv);
  ALLOC(u, 3 * (size_t)nv);
  ALLOC(v, 3 * (size_t)nv);
  ALLOC(u_old, 3 * (size_t)nv);
  ALLOC(v_old, 3 * (size_t)nv);
  ALLOC(mass, 3 * (size_t)nv);
  ALLOC(one_ring, (size_t)nv);
  ALLOC(pin_target, 3 * (size_t)nv);
  ALLOC(rhs64, 3 * (size_t)nv);
  ALLOC(tmp64, 3 * (size_t)nv);
  ALLOC(dminv, 9 * (size_t)nt);
  ALLOC(vol, (size_t)nt);
  ALLOCT(kneg, 9 * (size_t)P.nslots);
  ALLOC(x64, 3 * (size_t)nv);
  ALLOC(res64, 3 * (size_t)nv);
  ALLOC(rstate, 1);
  ALLOC(tets, (size_t)nt);
  ALLOC(slice_off, P.slice_off.size());
  ALLOC(row_len, (size_t)nv);
  ALLOC(col, (size_t)P.nslots);
  ALLOC(slot_row, (size_t)P.nslots);
  ALLOC(cptr, P.cptr.size());
  ALLOC(hdest, P.hdest.size());
  ALLOC(vptr, P.vptr.size());
  ALLOC(fdest, P.fdest.size());
  ALLOC(pinned, (size_t)nv);
  ALLOCT(stage_h, (size_t)9 * P.cptr[P.nslots]);
  ALLOCT(stage_f, (size_t)3 * P.vptr[nv]);
  ALLOCT(A, 9 * (size_t)P.nslots);
  ALLOCT(D, 9 * (size_t)P.nslots);
  ALLOCT(E, 9 * (size_t)P.nslots);
  ALLOCT(C, 9 * (size_t)P.nslots);
  ALLOCT(minv, 9 * np);
  ALLOCT(r, 3 * np);
  ALLOCT(x, 3 * np);
  ALLOCT(z, 3 * np);
  ALLOCT(p0, 3 * np);
  ALLOCT(p1, 3 * np);
  ALLOCT(t, 3 * np);
  ALLOC(sc, 2);
  I->n_rhs = rhs_grid(nv);
  I->n_spmv = spmv_grid(nv, P.shift);
  I->n_update = update_grid(nv);
  ALLOC(part_rhs_a, (size_t)I->n_rhs);
  ALLOC(part_rhs_b, (size_t)I->n_rhs);
  ALLOC(part_pt, (size_t)I->n_spmv);
  ALLOC(part_rr, (size_t)I->n_update);
  ALLOC(part_rz, (size_t)I->n_update);
#undef ALLOC
#undef ALLOCT
  if (hipHostMalloc((void**)&I->sc_host, sizeof(CgScalars), hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&I->rstate_host, sizeof(RefineState), hipHostMallocDefault) != hipSuccess)
    return bail(fail(ctx, HOBAK_E_HIP, "hipHostMalloc failed"));
  for (auto& e : I->ev)
    if (hipEventCreate(&e) != hipSuccess) return bail(fail(ctx, HOBAK_E_HIP, "hipEventCreate failed"));
  std::vector<double> mass(3 * (size_t)nv);
  for (int i = 0; i < nv; i++)  // buildMassMatrix + scaleMass
    for (int c = 0; c < 3; c++) mass[3 * (size_t)i + c] = one_ring[i] * params->mass_scale;
  std::vector<int4> t4(nt);
  for (int t = 0; t < nt; t++) t4[t] = make_int4(tets[4 * t], tets[4 * t + 1], tets[4 * t + 2], tets[4 * t + 3]);
  if ((rc = upload(I, I->rest, rest_xyz, 3 * (size_t)nv)) ||
      (rc = upload(I, I->mass, mass.data(), mass.size())) ||
      (rc = upload(I, I->one_ring, one_ring.data(), one_ring.size())) ||
      (rc = upload(I, I->dminv, dminv.data(), dminv.size())) ||
      (rc = upload(I, I->vol, vol.data(), vol.size())) ||
      (rc = upload(I, I->tets, t4.data(), t4.size())) ||
      (rc = upload(I, I->slice_off, P.slice_off.data(), P.slice_off.size())) ||
      (rc = upload(I, I->row_len, P.row_len.data(), P.row_len.size())) ||
      (rc = upload(I, I->col, P.col.data(), P.col.size())) ||
      (rc = upload(I, I->slot_row, P.slot_row.data(), P.slot_row.size())) ||
      (rc = upload(I, I->cptr, P.cptr.data(), P.cptr.size())) ||
      (rc = upload(I, I->hdest, P.hdest.data(), P.hdest.size())) ||
      (rc = upload(I, I->vptr, P.vptr.data(), P.vptr.size())) ||
      (rc = upload(I, I->fdest, P.fdest.data(), P.fdest.size())))
    return bail(rc);
  if ((rc = (I->tsize == 8 ? setup_constant_parts<double>(I) : setup_constant_parts<float>(I))) != HOBAK_OK)
    return bail(rc);
  if (hipStreamSynchronize(ctx->stream) != hipSuccess)
    return bail(fail(ctx, HOBAK_E_HIP, "set-up kernels failed"));
  ctx->insts.insert(I);
  *out = I;
  return HOBAK_OK;
}

void hobak_inst_destroy(hobak_inst* I) {
  if (!I) return;
  (void)hipSetDevice(I->ctx->device);
  (void)hipStreamSynchronize(I->ctx->stream);
  I->ctx->insts.erase(I);
  free_inst(I);
  delete I;
}

int hobak_inst_set_pinned(hobak_inst* I, const int32_t* vids, int32_t n, const double* targets) {
  if (!I || n < 0 || (n > 0 && !vids)) return HOBAK_E_INVALID;
  std::vector<unsigned char> flag(I->nv, 0);
  std::vector<double> tgt(3 * (size_t)I->nv, 0.0);
  for (int k = 0; k < n; k++) {
    if (vids[k] < 0 || vids[k] >= I->nv) return fail(I->ctx, HOBAK_E_INVALID, "pinned vertex out of range");
    flag[vids[k]] = 1;
    for (int c = 0; c < 3; c++) tgt[3 * (size_t)vids[k] + c] = targets ? targets[3 * (size_t)k + c] : 0.0;
  }
  I->pinned_ids.assign(vids, vids + n);
  TRY(upload(I, I->pinned, flag.data(), flag.size()));
  TRY(upload(I, I->pin_target, tgt.data(), tgt.size()));
  HIPCHK(I->ctx, hipStreamSynchronize(I->ctx->stream));
  return HOBAK_OK;
}

int hobak_inst_set_state(hobak_inst* I, const double* u, const double* v) {
  if (!I) return HOBAK_E_INVALID;
  if (u) TRY(upload(I, I->u, u, (size_t)I->dofs));
  if (v) TRY(upload(I, I->v, v, (size_t)I->dofs));
  HIPCHK(I->ctx, hipStreamSynchronize(I->ctx->stream));
  I->prev_iters = -1;
  return HOBAK_OK;
}

int hobak_inst_get_state(hobak_inst* I, double* u, double* v) {
  if (!I) return HOBAK_E_INVALID;
  hipStream_t s = I->ctx->stream;
  if (u) HIPCHK(I->ctx, hipMemcpyAsync(u, I->u, sizeof(double) * I->dofs, hipMemcpyDeviceToHost, s));
  if (v) HIPCHK(I->ctx, hipMemcpyAsync(v, I->v, sizeof(double) * I->dofs, hipMemcpyDeviceToHost, s));
  HIPCHK(I->ctx, hipStreamSynchronize(s));
  return HOBAK_OK;
}

int hobak_step(hobak_inst* I, hobak_step_stats* st) {
  if (!I) return HOBAK_E_INVALID;
  (void)hipSetDevice(I->ctx->device);
  auto t0 = std::chrono::steady_clock::now();
  hobak_step_stats local;
  std::memset(&local, 0, sizeof(local));
  int rc = I->tsize == 8 ? step_impl<double>(I, &local) : step_impl<float>(I, &local);
  auto t1 = std::chrono::steady_clock::now();
  local.seconds = std::chrono::duration<double>(t1 - t0).count();
  if (st) *st = local;
  return rc;
}

int hobak_step_batch(hobak_ctx* ctx, hobak_inst* const* insts, int32_t n, hobak_step_stats* stats) {
  if (!ctx || (n > 0 && !insts)) return HOBAK_E_INVALID;
  int worst = HOBAK_OK;
  for (int i = 0; i < n; i++) {
    if (!insts[i] || insts[i]->ctx != ctx) return fail(ctx, HOBAK_E_INVALID, "instance of another context");
    int rc = hobak_step(insts[i], stats ? &stats[i] : nullptr);
    if (rc == HOBAK_E_HIP) return rc;
    if (rc != HOBAK_OK && worst == HOBAK_OK) worst = rc;
  }
  return worst;
}

int hobak_get_bsr_pattern(hobak_inst* I, int32_t* nnzb, int32_t* row_ptr, int32_t* col_idx) {
  if (!I) return HOBAK_E_INVALID;
  if (nnzb) *nnzb = I->P.nnzb;
  if (row_ptr) std::copy(I->P.row_ptr.begin(), I->P.row_ptr.end(), row_ptr);
  if (col_idx) std::copy(I->P.col_idx.begin(), I->P.col_idx.end(), col_idx);
  return HOBAK_OK;
}

}  // extern "C"

namespace {
template <typename T>
int eval_forces_hessian(hobak_inst* I, const double* u, double* forces, double* kvals) {
  hipStream_t s = I->ctx->stream;
  TRY(upload(I, I->tmp64, u, (size_t)I->dofs));
  HIPCHK(I->ctx, launch_element<T>(element_args(I, I->tmp64, kvals != nullptr), s));
  if (forces) {
    HIPCHK(I->ctx, launch_force_gather<T>(I->nv, I->vptr, (const T*)I->stage_f, I->rhs64, s));
    HIPCHK(I->ctx, hipMemcpyAsync(forces, I->rhs64, sizeof(double) * I->dofs, hipMemcpyDeviceToHost, s));
  }
  if (kvals) {
    void* tmpK = nullptr;
    HIPCHK(I->ctx, hipMalloc(&tmpK, sizeof(T) * 9 * (size_t)I->P.nslots));
    HIPCHK(I->ctx, hipMemsetAsync(tmpK, 0, sizeof(T) * 9 * (size_t)I->P.nslots, s));
    HIPCHK(I->ctx, launch_assemble<T>(asm_args(I, 0, tmpK), s));
    int rc = download_sell_as_bsr<T>(I, tmpK, kvals, -1.0);
    (void)hipFree(tmpK);
    if (rc != HOBAK_OK) return rc;
  }
  HIPCHK(I->ctx, hipStreamSynchronize(s));
  I->has_system = false;  // rhs64 was reused
  return HOBAK_OK;
}
}  // namespace

extern "C" {

int hobak_eval_forces_hessian(hobak_inst* I, const double* u, double* forces, double* kvals) {
  if (!I || !u) return HOBAK_E_INVALID;
  (void)hipSetDevice(I->ctx->device);
  return I->tsize == 8 ? eval_forces_hessian<double>(I, u, forces, kvals)
                       : eval_forces_hessian<float>(I, u, forces, kvals);
}

int hobak_get_last_system(hobak_inst* I, double* rhs, double* lhs_vals) {
  if (!I) return HOBAK_E_INVALID;
  if (!I->has_system) return fail(I->ctx, HOBAK_E_INVALID, "no system formed since the last step");
  hipStream_t s = I->ctx->stream;
  if (rhs) {
    HIPCHK(I->ctx, hipMemcpyAsync(rhs, I->rhs64, sizeof(double) * I->dofs, hipMemcpyDeviceToHost, s));
    HIPCHK(I->ctx, hipStreamSynchronize(s));
  }
  if (lhs_vals)
    TRY(I->tsize == 8 ? download_sell_as_bsr<double>(I, I->A, lhs_vals, 1.0)
                      : download_sell_as_bsr<float>(I, I->A, lhs_vals, 1.0));
  return HOBAK_OK;
}

int hobak_last_step_timing(hobak_inst* I, double* out_ms) {
  if (!I || !out_ms) return HOBAK_E_INVALID;
  std::memcpy(out_ms, I->last_ms, sizeof(I->last_ms));
  return HOBAK_OK;
}

int hobak_set_kernel_sampling(hobak_inst* I, int32_t every) {
  if (!I || every < 0) return HOBAK_E_INVALID;
  I->sample_every = every;
  I->spmv_ms_sum = 0.0;
  I->spmv_count = 0;
  if (every > 0 && I->sample_ev.empty()) {
    I->sample_ev.assign(2 * 512, nullptr);
    for (auto& e : I->sample_ev)
      HIPCHK(I->ctx, hipEventCreate(&e));
  }
  return HOBAK_OK;
}

int hobak_kernel_stats(hobak_inst* I, double* spmv_avg_ms, int64_t* spmv_count, double* spmv_bytes) {
  if (!I) return HOBAK_E_INVALID;
  if (spmv_avg_ms) *spmv_avg_ms = I->spmv_count ? I->spmv_ms_sum / (double)I->spmv_count : 0.0;
  if (spmv_count) *spmv_count = I->spmv_count;
  if (spmv_bytes) {
    const double sz = (double)I->tsize;
    *spmv_bytes = (double)I->P.nnzb * (9.0 * sz + 4.0) + 4.0 * (I->nv + 1) + 2.0 * 3.0 * sz * I->nv;
  }
  return HOBAK_OK;
}

int hobak_copy_probe(hobak_ctx* ctx, size_t bytes, int32_t reps, double* ms) {
  if (!ctx || bytes < 32 || reps <= 0) return HOBAK_E_INVALID;
  (void)hipSetDevice(ctx->device);
  bytes &= ~(size_t)31;
  void *a = nullptr, *b = nullptr;
  HIPCHK(ctx, hipMalloc(&a, bytes));
  if (hipMalloc(&b, bytes) != hipSuccess) {
    (void)hipFree(a);
    return fail(ctx, HOBAK_E_NOMEM, "hipMalloc failed");
  }
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  (void)hipMemsetAsync(a, 0, bytes, ctx->stream);
  (void)launch_copy(a, b, bytes, ctx->stream);
  (void)hipEventRecord(e0, ctx->stream);
  for (int k = 0; k < reps; k++) (void)launch_copy(a, b, bytes, ctx->stream);
  (void)hipEventRecord(e1, ctx->stream);
  hipError_t err = hipEventSynchronize(e1);
  float t = 0.f;
  (void)hipEventElapsedTime(&t, e0, e1);
  if (ms) *ms = t / reps;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  return err == hipSuccess ? HOBAK_OK : fail(ctx, HOBAK_E_HIP, "copy probe failed");
}

int hobak_ctx_set_tuning(hobak_ctx* ctx, int32_t key, int32_t value) {
  if (!ctx) return HOBAK_E_INVALID;
  switch (key) {
    case HOBAK_TUNE_SLICE_ROWS:
      if (value != 8 && value != 16 && value != 32 && value != 64)
        return fail(ctx, HOBAK_E_INVALID, "slice rows must be 8, 16, 32 or 64");
      ctx->slice_rows = value;
      return HOBAK_OK;
    case HOBAK_TUNE_POLL_CHUNK:
      if (value < 1 || value > 4096) return fail(ctx, HOBAK_E_INVALID, "poll chunk out of range");
      ctx->poll_chunk = value;
      return HOBAK_OK;
    default:
      return fail(ctx, HOBAK_E_INVALID, "unknown tuning key");
  }
}

}  // extern "C"
