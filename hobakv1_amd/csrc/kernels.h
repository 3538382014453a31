// Device records, segmented launch descriptors and launchers shared by the translation units of
// libhobak_hip.so.
//
// Device layouts (all resident in HBM, one set per simulation instance):
//   vectors        DOF-ordered AoS: entry 3*v + c (the reference's flat VECTOR order)
//   tets           int4 per tet
//   DmInv          9 planes of nt doubles (plane 3*r + c holds DmInv(r, c))       -> coalesced
//   element stage  tet-major, so that the element kernel writes whole lines.  A tet's stiffness
//                  is symmetric, so only its 10 blocks (x, y) with x <= y are staged: entry (a, b)
//                  of block pair(x, y) of tet t is stage_h[9 * (pair(x, y) * nt + t) + 3 a + b]
//                  (rows of corner x).  The upper slot (min(v_x, v_y), max(v_x, v_y)) lists its
//                  contributions in ascending tet order in cidx[cptr[slot] .. cptr[slot + 1]):
//                  each entry is pair * nt + t, with kCidxTranspose set where corner x is the
//                  slot's column (the slot sums the block transposed).  Every upper slot also
//                  writes the transpose of its sum into its mirror (lower) slot.  The force
//                  3-vector of corner x of tet t is stage_f[3 * (x * nt + t) + c]; vertex v's
//                  incident corners are listed in ascending tet order in
//                  fidx[vptr[v] .. vptr[v + 1]) (entries x * nt + t)
//   block matrix   SELL-H, k-interleaved (H = 8, 16, 32 or 64 rows per slice, L = 64 / H lanes
//                  per row): slice s owns slots [slice_off[s], slice_off[s+1]) (a multiple of 64),
//                  and the k-th block of row r sits at slot = slice_off[r / H] + H * k + (r % H).
//                  A wavefront that gives each of a slice's H rows L lanes (lane = H * (k % L) +
//                  r % H) therefore reads slice_off + 64 * (k / L) + lane: contiguous 512-byte
//                  FP64 wave loads, while the H lanes that share k gather the vector at H
//                  consecutive rows' k-th neighbours (consecutive on a lattice).
//                  9 value planes of nslots (plane 3a+b = entry (a, b) of the 3x3 block), one
//                  int32 column (vertex) and one int32 owning row (-1 for padding) per slot.
//                  Columns of a row are ascending in k.
//   preconditioner per row: 9 contiguous values (block Jacobi inverse, row-major) or per DOF
//                  (Jacobi 1/a_ii)
//
// Difference form (the reduced-precision form's operators).  Every element stiffness has block
// rows that sum to zero (a translation does not deform), so with m the lumped mass,
//   (A x)_i = dg_i x_i + sum_j A_ij (x_j - x_i),   dg_i = m_i (1/dt^2 - alpha/dt) (+ ground),
// exactly, where the sum runs over the stored blocks (the diagonal one multiplies zero).  Rounding
// A_ij to FP32 then perturbs only the response to deformation: a rigid translation is annihilated
// whatever the rounding, while in the standard form FP32 entries give it a spurious stiffness
// force of relative size ~1e-7 * stiffness / mass that grows over a trajectory.  Rows of pinned
// vertices store zero off-diagonal blocks and dg = 1; couplings to pinned columns are kept (every
// solver vector is zero at pinned vertices).  E v and C v_b in the right-hand side use the same
// form.  The plain FP64 form keeps the standard form, which is the reference's arithmetic.
//
// Batching.  Every kernel is *segmented*: one launch serves any set of instances.  Each instance
// owns a read-only InstDev record in device memory; a Batch names the records of the launch and
// a prefix table of virtual workgroups (boff[i] = first virtual workgroup of instance i).  A
// workgroup finds its instance by counting the prefix entries at or below its index, so no
// workgroup spans two instances, and
// workgroups stride over the virtual workgroups when the launch grid is capped (persistent
// form).  A single instance is a batch of one.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace hobak {

constexpr int kDefaultSliceRows = 32;
constexpr int kBlock = 256;
constexpr int kElementBlock = 128;
constexpr int kElementVerts = 4 * kElementBlock;  // distinct vertices of a workgroup's tets, at most
constexpr int kMinvStride = 9;   // block-Jacobi inverse per row; Jacobi uses 1/a_ii per DOF
constexpr int kUpdateCap = 512;  // default per-instance workgroup cap of the PCG update kernel
constexpr int kMaxSweeps = 32;   // FP64 refinement sweeps per linear solve
constexpr int kMaxShapes = 8;    // kinematic collision cubes per instance
constexpr int kAsmSlotsPerWave = 7;  // the assembly's block slots per wavefront (9 lanes each)

// Virtual-workgroup kinds: each instance contributes nb[kind] workgroups to a launch of a kernel
// of that kind.
enum Kind { K_ELEM = 0, K_SLOT, K_ROW, K_DOF, K_SPMV, K_UPD, K_ONE, K_ASM, K_COUNT };

// Scalars of one PCG solve, resident on the device so the loop runs without host round trips.
// Two copies ping-pong: an iteration of parity p reads set p and its SpMV kernel writes set
// p ^ 1 from one workgroup only, so no kernel reads a scalar that another of its own workgroups
// writes.  `first` marks the set written by the initialisation, so the iteration kernels need
// no iteration index and can be captured once and replayed.
struct CgScalars {
  double rhsNorm2;
  double threshold;
  double absNew;
  double alpha;
  double beta;
  double resNorm2;
  double maxR;        // |S b|^2, the BDF_1 Newton test value
  int iters;
  int maxIters;
  int done;           // 1 once converged / exhausted / failed, or when no solve is active
  int status;         // 0 ok, 1 not converged, 2 non-finite
  int first;          // 1 until the first SpMV of the solve has run
  int pad[3];
  double e[3];        // coarse correction of this set's z (written for the first set only)
  double pad2;
};

// FP64 iterative refinement around the reduced-precision PCG (the precision form).
struct RefineState {
  double b2;        // |S b|^2
  double res2;      // |rhs - A x64|^2 of the current iterate
  double tau;       // relative tolerance of the next inner solve
  int done;         // res2 < tol^2 b2 (or b2 == 0), sweep limit, or a non-finite value
  int status;       // 0 ok, 1 sweep limit, 2 non-finite
  int sweeps;
  int inner_iters;  // PCG iterations summed over the sweeps
  int first_iters;  // PCG iterations of the first sweep
  int pad;
};

enum Phase { PHASE_IDLE = 0, PHASE_SOLVING = 1, PHASE_FINISHED = 2 };

// Mutable per-instance control block (written by kernels with ordinary vector stores).
struct Ctl {
  CgScalars sc[2];
  RefineState rs;
  int phase;
  int pad[3];
  double ac_inv[3];  // coarse correction: 1 / (P^T A P), diagonal (see InstDev::coarse)
  unsigned spmv_work;  // SpMV launches that streamed this instance's matrix (wraps; a count)
  int pad4;
};

// What the host reads back after a group of PCG iterations, one per instance of a batch.
struct InstStatus {
  int phase;
  int iters;        // iterations of the current (or last) inner solve
  int inner_total;  // refinement: iterations summed over finished sweeps
  int sweeps;
  int cg_status;
  int ref_status;
  int ref_done;
  int first_iters;  // iterations of the first (refinement: first sweep's) inner solve
  double err;       // sqrt(res^2 / b^2) of the finished solve
  double maxR;      // |S b|^2
  unsigned spmv_work;  // Ctl::spmv_work
  int pad;
};

// Read-only per-instance record.  Pointers typed void* hold the solve's real type T.
struct InstDev {
  int nv, nt, nslots, shift;
  int nasm;           // upper slots (the assembly's work list, aslot)
  int nb[K_COUNT];
  int material, precond, bdf, ground_on, refine, max_iters;
  double mu, lam, alpha;           // SNH: lam = lambda + mu, alpha = 1 + mu / lam
  double invDt, invDt2, ra, rb;
  double g[3];
  double ground_y, ground_k;
  double tol, tau0;                // outer PCG / refinement tolerance; first inner tolerance
  const int4* tets;
  const double *rest, *dminv, *vol, *one_ring, *mass, *pin_target;
  double *u, *v, *u_old, *v_old, *tmp64, *rhs64, *x64, *res64;
  const unsigned char* pinned;
  const int *slice_off, *row_len, *col, *slot_row, *cptr, *cidx, *vptr, *fidx;
  const int* mirror;  // per upper off-diagonal slot: the slot of the transposed block (else -1)
  const int* aslot;   // the upper slots in slot order
  // element kernel: per workgroup of kElementBlock tets, its distinct vertices (ascending) in
  // evert[evptr[w] .. evptr[w + 1]); per tet, its 4 corners' positions in that list
  const int *evptr, *evert;
  const unsigned short* eloc;
  void *stage_h, *stage_f, *A, *D, *E, *C, *kneg, *minv, *r, *x, *z, *t, *p0, *p1;
  Ctl* ctl;    // == ctl_of(this) on the device
  double* dg;  // reduced-precision form: per-DOF diagonal term of the difference-form operator
  // per-workgroup partial sums, sized by the producing kernel's nb[kind]
  double *part_row_a, *part_row_b;  // K_ROW: right-hand side / inner-solve seed
  double *part_pt;                  // K_SPMV: p.t
  double *part_rr, *part_rz;        // K_UPD: r.r, r.z
  double *part_res;                 // K_SPMV: refinement residual
  // Reduced-precision form of an instance with no pinned vertex: the inner PCG's preconditioner
  // adds a coarse correction over the rigid translations to block Jacobi,
  //   z = M^-1 r + P (P^T A P)^-1 P^T r,   P = [I I ... I]^T  (one 3x3 identity per vertex).
  // The difference form annihilates a translation off the diagonal, so P^T A P = sum_i diag(dg_i)
  // and A (z + P e) = A z + dg (.) e.  z holds M^-1 r; the SpMV adds e = (P^T A P)^-1 P^T r
  // where it needs the whole z (its own rows and r.z).
  int coarse;
  double* part_row_c;  // K_ROW: 6 planes of nb[K_ROW]: sum of the seed r (3), sum of dg (3)
  double* part_upd_g;  // K_UPD: 3 planes of nb[K_UPD]: sum of r
  // Reduced-precision form of an instance with pinned vertices, held in a spatial (Hilbert-curve) vertex
  // order: local aggregates in place of the correction above.  Every update chunk (the 21 *
  // groups consecutive rows one wave of the update kernel handles together) is an aggregate I,
  // and the preconditioner is block Jacobi plus a block-diagonal coarse space,
  //   z = M^-1 r + sum_I P_I (P_I^T A P_I)^-1 P_I^T r,   P_I = translations of I's free vertices,
  // where P_I^T A P_I = sum over I's free rows i of (dg_i - sum of A_ij over the row's columns j
  // outside I's free vertices).  agg_inv (T, row-major, 9 per chunk) holds the inverses, built
  // once per solve from A; z then holds the whole correction, so the SpMV is unchanged.
  int agg;
  void* agg_inv;
  // Kinematic collision cubes and Baraff-Witkin plane constraints (contact != 0; FP64 form).
  // shape[o] = {center x, y, z, scale, inverse scale}: world = scale * local + center.
  int nshapes, contact;
  double shape[kMaxShapes][5];
  const unsigned char* surf;  // per vertex: on the surface
  int* pc_obj;                // per vertex: the constraining shape, or -1
  double *pc_cl, *pc_nl;      // per vertex: closest point and normal in the shape's frame
  unsigned char* pc_sep;      // per vertex: separating at the end of this step
  unsigned char* fmode;       // per vertex: this step's filter block, 0 I, 1 I - n n^T, 2 zero
  double *fnrm, *zc, *bunf;   // per vertex: the filter's normal, z, the unfiltered b
  double* part_sep;           // K_ROW: separating-constraint counts
  // Energy-based damping (GREEN_DAMPING) in place of Rayleigh damping: the step adds the damping
  // stress mu_d F (F^T Fdot + Fdot^T F) at the held per-tet velocity gradients (9 planes of nt,
  // row-major entry (r, c) in plane 3 r + c) to the elastic stress; ra = rb = 0 then.
  int damping;
  double dmu;
  const double* fdot;
  // BACKWARD_EULER_VELOCITY (velform != 0): the solve is for the velocity change.  The system is
  // A = M - dt^2 K (kscale = dt^2 scales the assembled stiffness; 1 in the position form, where
  // A = M/dt^2 - C/dt - K), b = dt (R + dt K v + f_ext), with K v taken from the unprojected
  // stiffness the assembly also writes into E.
  int velform;
  double dt, kscale;
};

// The control block lives in the same allocation as the record, kCtlOffset bytes after it, so a
// kernel reaches it without loading a pointer first.
constexpr size_t kCtlOffset = (sizeof(InstDev) + 255) / 256 * 256;
__host__ __device__ __forceinline__ Ctl* ctl_of(const InstDev* I) {
  return (Ctl*)((char*)const_cast<InstDev*>(I) + kCtlOffset);
}
__device__ __forceinline__ __attribute__((address_space(1))) Ctl* ctl_of(
    const __attribute__((address_space(1))) InstDev* I) {
  return (__attribute__((address_space(1))) Ctl*)((const __attribute__((address_space(1))) char*)I + kCtlOffset);
}

// One segmented launch: the instances and the virtual-workgroup prefix table of the launch's kind.
struct Batch {
  const InstDev* const* inst;  // device, n entries
  const int* boff;             // device, n + 1 entries
  const InstDev* one;          // the record itself when n == 1 (saves a dependent load)
  int n;
  int total;                   // boff[n]
};

// The virtual workgroup that launch slot ws runs.  The dispatcher hands consecutive workgroups
// to the 8 XCDs in turn and each XCD has its own L2, so the map gives XCD x (the slots with
// ws % 8 == x) one contiguous range of virtual workgroups: neighbouring rows, whose gathers touch
// the same vector entries, then share one L2.  A bijection of [0, total), so every virtual
// workgroup still runs exactly once, and results do not depend on it.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_vb(int ws, int total) {
  const int q = total / kXcds, r = total % kXcds;
  const int x = ws % kXcds, i = ws / kXcds;
  return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + i;
}

// The instance of virtual workgroup vb, its workgroup index inside the instance (lb) and the
// instance's workgroup count of this kind (lgrid).  Instance k owns virtual workgroups
// boff[k] .. boff[k+1]-1, so k is the number of entries of boff[1..n-1] that are <= vb.  The wave
// counts them with independent loads (one round for up to 256 instances) and a ballot, instead of
// a binary search whose every step waits on the previous load.  vb is uniform, so is k.
__device__ __forceinline__ const __attribute__((address_space(1))) InstDev* seg_find(const Batch& b, int vb,
                                                                                int& lb, int& lgrid) {
  typedef const __attribute__((address_space(1))) InstDev* GI;
  typedef const __attribute__((address_space(1))) int* GInt;
  if (b.n == 1) {
    lb = vb;
    lgrid = b.total;
    return (GI)b.one;
  }
  const GInt boff = (GInt)b.boff;
  const int lane = threadIdx.x & 63;
  int k = 0;
  for (int base = 1; base < b.n; base += 256) {
    bool le[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {  // unconditional loads (boff has n + 1 entries), one round
      const int i = base + 64 * q + lane;
      const int v = boff[i < b.n ? i : b.n];
      le[q] = (i < b.n) & (v <= vb);
    }
#pragma unroll
    for (int q = 0; q < 4; q++) k += __popcll(__ballot(le[q]));
  }
  const int b0 = boff[k];
  lb = vb - b0;
  lgrid = boff[k + 1] - b0;
  typedef const InstDev* const __attribute__((address_space(1)))* GTab;
  return (GI)((GTab)b.inst)[k];
}

// Pointers read out of an InstDev are generic to the compiler, which then emits flat memory
// instructions that it must wait on one by one.  Every kernel converts them with gbl() first,
// so loads are global_load instructions that can be in flight together.
#define HOBAK_GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ HOBAK_GLOBAL T* gbl(T* p) {
  return (HOBAK_GLOBAL T*)p;
}

// Workgroups per kind for an instance (host side; the same formulas size the partial arrays).
int blocks_of(int kind, int nv, int nt, int nslots, int shift, int update_cap, int update_groups);

// Launchers.  `cap` limits the launch grid (0: one workgroup per virtual workgroup).
// ELEM_EXT: a displacement the caller holds in device memory (the no-allocation entry point)
enum ElemSrc { ELEM_U = 0, ELEM_TMP = 1, ELEM_EXT = 2 };
enum AsmTarget { ASM_REST = 0, ASM_PLAIN_A = 1, ASM_SYSTEM = 2 };

// Staged stiffness blocks per tet (the upper triangle of its 4x4 block pattern), their order, and
// the flag and index bits of a contribution-list entry.
constexpr int kStagedBlocks = 10;
constexpr int kCidxTranspose = (int)0x80000000u;
constexpr int kCidxIndex = 0x7fffffff;
__host__ __device__ constexpr int staged_pair(int x, int y) { return 4 * x - (x * (x - 1)) / 2 + (y - x); }

// u_ext (with ELEM_EXT) and stage_f_ext (non-null: the force staging) replace the instance's own
// buffers for a batch of one.
template <typename T>
hipError_t launch_element(const Batch& b, int src, int with_hessian, int cap, hipStream_t s,
                          const double* u_ext = nullptr, void* stage_f_ext = nullptr);
// Adds the GREEN_DAMPING force of the instances with energy damping to the staged element forces.
template <typename T>
hipError_t launch_damping_force(const Batch& b, int src, int cap, hipStream_t s);
// Velocity gradients Fdot = [v1 - v0, v2 - v0, v3 - v0] DmInv of one instance's tets.
hipError_t launch_fdots(const InstDev& h, const double* vel, double* fdot, hipStream_t s);
// The damping stiffness blocks vol * dP/dFdot (positive semi-definite) of every tet at the
// displacement `src` selects, staged for the assembly like the element stiffness.
template <typename T>
hipError_t launch_damping_hessian(const Batch& b, int src, int cap, hipStream_t s);
hipError_t launch_begin_step(const Batch& b, int cap, hipStream_t s);
hipError_t launch_end_step(const Batch& b, int cap, hipStream_t s);
// plane constraints: new ones, the filter and z (prologue); the separation test (epilogue)
hipError_t launch_contact_begin(const Batch& b, int cap, hipStream_t s);
hipError_t launch_contact_sep(const Batch& b, int cap, hipStream_t s);
template <typename T>
hipError_t launch_apply_dx(const Batch& b, int use64, int cap, hipStream_t s);
template <typename T>
hipError_t launch_assemble(const Batch& b, int target, int cap, hipStream_t s);
template <typename T>
hipError_t launch_constant_parts(const Batch& b, int cap, hipStream_t s);
template <typename T>
hipError_t launch_rhs(const Batch& b, int init_cg, int cap, hipStream_t s);
hipError_t launch_cg_init(const Batch& b, int refine, int gate, hipStream_t s);
// pad: 1 walks every slice's padded width (uniform trip count), 0 only each row's own blocks.
// stream: 1 reads the matrix with non-temporal loads (for matrices larger than the Infinity
// Cache, whose lines would only displace the gathered vectors), 0 with ordinary loads (a matrix
// that fits stays cached from one iteration to the next).
// coarse: some instance of the batch has the reduced-precision form's coarse correction on
// (InstDev::coarse); the kernel then adds it for those instances.
template <typename T>
hipError_t launch_cg_spmv(const Batch& b, int shift, int pad, int stream, int coarse, int parity, int cap,
                          hipStream_t s, hipEvent_t start = nullptr, hipEvent_t stop = nullptr);
// coarse: bit 0, some instance has the coarse correction (InstDev::coarse); bit 1, some has
// local aggregates (InstDev::agg).
template <typename T>
hipError_t launch_cg_update(const Batch& b, int groups, int coarse, int parity, int cap, hipStream_t s);
// The local aggregates' inverses (InstDev::agg_inv) from the assembled A, once per solve, for the
// instances with agg set (a Batch of kind K_ROW; groups: the update kernel's, 21 * groups rows per
// chunk).
template <typename T>
hipError_t launch_agg_build(const Batch& b, int groups, int cap, hipStream_t s);
// The FP64 p-form of an iteration (solver_kernels.hip): p finishes the previous iteration and
// forms p_new = z + beta p_old over the update kernel's workgroups (Batch of kind K_UPD);
// spmv_p forms t = A p_new gathering p_new only.  Bitwise the same results as the two-kernel form.
hipError_t launch_cg_p(const Batch& upd, int parity, int cap, hipStream_t s);
hipError_t launch_cg_spmv_p(const Batch& b, int shift, int pad, int stream, int parity, int cap, hipStream_t s,
                            hipEvent_t start = nullptr, hipEvent_t stop = nullptr);
hipError_t launch_refine_begin(const Batch& b, int cap, hipStream_t s);
template <typename T>
hipError_t launch_refine_init(const Batch& b, int gate, int groups, int cap, hipStream_t s);
template <typename T>
hipError_t launch_refine_residual(const Batch& spmv, const Batch& dof, int shift, int cap, hipStream_t s);
hipError_t launch_finish_check(const Batch& b, int refine, hipStream_t s);
hipError_t launch_status(const Batch& b, InstStatus* out, hipStream_t s);
// The per-vertex force gather into rhs64, or from stage_f_ext into out_ext (both non-null).
template <typename T>
hipError_t launch_force_gather(const Batch& b, int cap, hipStream_t s, const void* stage_f_ext = nullptr,
                               double* out_ext = nullptr);
hipError_t launch_copy(const void* src, void* dst, size_t bytes, hipStream_t s);

}  // namespace hobak
