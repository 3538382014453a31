/*
 * hobak_c.h -- C ABI of the MI355X backward-Euler Newton step (libhobak_hip.so).
 *
 * The reference exposes this path as C++ virtual interfaces, not an FFI.  Every entry point below
 * names the reference interface it replaces (paths relative to the reference root):
 *
 *   TET_MESH(restVertices, tets)                     src/Geometry/TET_MESH.h:37-38
 *   TET_MESH_FASTER(restVertices, tets)              src/Geometry/TET_MESH_FASTER.h:28-32
 *   VOLUME::SNH(mu, lambda), VOLUME::ARAP(mu, lambda) src/Hyperelastic/Volume/SNH.h:28, ARAP.h:28
 *   BACKWARD_EULER_POSITION(mesh, material)          src/Timestepper/BACKWARD_EULER_POSITION.h:36
 *   BACKWARD_EULER_VELOCITY(mesh, material)          src/Timestepper/BACKWARD_EULER_VELOCITY.h:35
 *   TIMESTEPPER::solve(verbose)                      src/Timestepper/TIMESTEPPER.h:77
 *   position() / velocity() / externalForces()       src/Timestepper/TIMESTEPPER.h:44-58
 *   addGravity, attachKinematicConstraints, setDt,   src/Timestepper/TIMESTEPPER.h:70-100
 *   setRayeligh, scaleMass
 *   computeHyperelasticForces                        src/Geometry/TET_MESH.h:72-75
 *   computeHyperelasticClampedHessian                src/Geometry/TET_MESH.h:107-110
 *   _damping = new GREEN_DAMPING(mu), computeFdots,  src/Timestepper/BACKWARD_EULER_POSITION.cpp:71-76,
 *   computeInternalForces, computeDampingHessian     125-131; src/Geometry/TET_MESH.h:74, 114, 124
 *
 * Conventions
 *   - Every function returns HOBAK_OK (0) or a negative HOBAK_E_* code; no C++ exception crosses
 *     the ABI.  hobak_last_error() gives the message of the last failure on a context.
 *   - Host buffers are caller-owned and copied in or out; device memory is library-owned and
 *     stays resident in HBM between steps.
 *   - Vectors are DOF-ordered like the reference's VECTOR: entry 3*v + c is coordinate c of vertex
 *     v.  position/displacement is u = x - X (the reference's TIMESTEPPER::position()).
 *   - Block matrices are exported as BSR with 3x3 row-major blocks.
 *   - A context is not thread-safe without external synchronisation.  Each instance runs on its
 *     context's HIP stream.
 */
#ifndef HOBAK_C_H
#define HOBAK_C_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HOBAK_OK 0
#define HOBAK_E_INVALID (-1)     /* bad argument or shape */
#define HOBAK_E_HIP (-2)         /* a HIP runtime call failed */
#define HOBAK_E_NOMEM (-3)       /* device allocation failed or the HBM budget would be exceeded */
#define HOBAK_E_NOT_CONVERGED (-4) /* PCG hit its iteration limit (state is still advanced) */
#define HOBAK_E_NONFINITE (-5)   /* NaN or Inf appeared in the solve */
#define HOBAK_E_NODEVICE (-6)    /* no usable gfx950 device */

#define HOBAK_MAT_SNH 0          /* Stable Neo-Hookean, src/Hyperelastic/Volume/SNH.cpp */
#define HOBAK_MAT_ARAP 1         /* As-Rigid-As-Possible, src/Hyperelastic/Volume/ARAP.cpp */

#define HOBAK_PRECOND_JACOBI 0       /* the reference's DiagonalPreconditioner */
#define HOBAK_PRECOND_BLOCK_JACOBI 1 /* 3x3 block Jacobi */

#define HOBAK_PRECISION_FP64 0       /* everything in FP64 */
#define HOBAK_PRECISION_FP32_REFINE 1 /* FP32 element/assembly/PCG + FP64 iterative refinement */

#define HOBAK_MAX_NEWTON 8

/* The time integrator (hobak_step_params.integrator).
 * POSITION: BACKWARD_EULER_POSITION::solveRayleighDamped (BACKWARD_EULER_POSITION.cpp:138-327),
 *   solving for the position change; BDF_1 with newton_iters > 1.
 * VELOCITY: BACKWARD_EULER_VELOCITY::solveRayleighDamped (BACKWARD_EULER_VELOCITY.cpp:107-235),
 *   solving (M - dt^2 K) dv = dt (R + dt K v + f_ext) for the velocity change, then
 *   v += dv, u += dt v, v = S v.  The reference empties its Rayleigh matrix C before forming
 *   the system (:153 -> TIMESTEPPER.cpp:657), so the step carries no Rayleigh damping; this
 *   library computes the same.  Plane constraints take velocity targets (cl - x)/dt - v
 *   (:51-78).  FP64 form, one Newton iteration, Rayleigh damping model, no penalty ground;
 *   other combinations are refused with HOBAK_E_INVALID. */
#define HOBAK_INTEGRATOR_POSITION 0
#define HOBAK_INTEGRATOR_VELOCITY 1

typedef struct hobak_ctx hobak_ctx;
typedef struct hobak_inst hobak_inst;
typedef struct hobak_sched hobak_sched;

/* VOLUME::SNH(mu, lambda) / VOLUME::ARAP(mu, lambda): unreparameterised Lame constants. */
typedef struct {
  int32_t kind;
  double mu;
  double lambda;
} hobak_material;

/* Integrator settings.  Defaults that match BACKWARD_EULER_POSITION::initialize
 * (src/Timestepper/BACKWARD_EULER_POSITION.cpp:50-79) come from hobak_default_params(). */
typedef struct {
  double dt;              /* setDt */
  double rayleigh_alpha;  /* setRayeligh(alpha, beta) */
  double rayleigh_beta;
  double mass_scale;      /* scaleMass(rho) applied once */
  double gravity[3];      /* body force passed to addGravity every step */
  double pcg_tol;         /* Eigen CG tolerance (<= 0: machine epsilon, the Eigen default) */
  int32_t pcg_max_iters;  /* <= 0: 2 * DOFs, the Eigen default */
  int32_t precond;        /* HOBAK_PRECOND_* */
  int32_t newton_iters;   /* 1: BACKWARD_EULER_POSITION; >1: BDF_1 Newton loop (src/Timestepper/BDF_1.cpp:78-239) */
  double newton_tol;      /* BDF_1 residual test ||S(b - A z)||^2 > tol (NEWMARK.cpp:43 default 1e-7) */
  int32_t precision;      /* HOBAK_PRECISION_* */
  double refine_inner_tol; /* FP32 inner PCG relative tolerance for the refinement form */
  int32_t ground_on;      /* penalty ground plane y = ground_y, stiffness ground_k (no reference counterpart) */
  double ground_y;
  double ground_k;
  int32_t integrator;     /* HOBAK_INTEGRATOR_* (default POSITION) */
} hobak_step_params;

typedef struct {
  int32_t status;                          /* HOBAK_OK or an error code for this step */
  int32_t newton_iters;                    /* Newton iterations taken */
  int32_t pcg_iters[HOBAK_MAX_NEWTON];     /* PCG iterations per Newton iteration (Eigen count) */
  int32_t refine_iters[HOBAK_MAX_NEWTON];  /* FP64 refinement sweeps per Newton iteration */
  int32_t total_pcg_iters;
  double pcg_error[HOBAK_MAX_NEWTON];      /* sqrt(|r|^2 / |b|^2) as Eigen's error() */
  double newton_residual[HOBAK_MAX_NEWTON];/* |S b|^2 before each solve */
  double seconds;                          /* host wall time of the call */
} hobak_step_stats;

void hobak_default_params(hobak_step_params* p);

/* Context: one device, one stream, an HBM budget for admission (0 = no limit). */
int hobak_ctx_create(int device, size_t hbm_budget_bytes, hobak_ctx** out);
void hobak_ctx_destroy(hobak_ctx* ctx);
const char* hobak_last_error(hobak_ctx* ctx);
/* Bytes of device memory currently held by this context's instances. */
size_t hobak_ctx_bytes_in_use(hobak_ctx* ctx);

/* new TET_MESH_FASTER(rest, tets) + new SNH/ARAP + new BACKWARD_EULER_POSITION(mesh, material).
 * rest_xyz: nv*3 doubles; tets: nt*4 int32 (positive rest volume required, as TET_MESH.cpp:132-137). */
int hobak_inst_create(hobak_ctx* ctx, const double* rest_xyz, int32_t nv, const int32_t* tets,
                      int32_t nt, const hobak_material* material, const hobak_step_params* params,
                      hobak_inst** out);
/* Device bytes hobak_inst_create would allocate for this mesh (admission control). */
size_t hobak_inst_bytes_estimate(int32_t nv, int32_t nt, int32_t precision);
/* HOBAK_OK when a mesh of nv vertices whose padded block matrix holds nslots slots stays within
 * the 32-bit byte offsets the PCG kernels address the matrix planes and the vectors with (every
 * slot of one value plane, and every vertex's three entries of a padded vector); else
 * HOBAK_E_INVALID, which hobak_inst_create then returns.  Host-only: needs no device.  No
 * counterpart in the reference, whose Eigen types index with 64-bit integers. */
int hobak_check_offsets(int64_t nv, int64_t nslots, int32_t precision);
void hobak_inst_destroy(hobak_inst* inst);

/* attachKinematicConstraints: pinned vertex ids (ascending) and the displacement each is held at
 * (world(local(X)) - X in the reference; NULL means zero). */
int hobak_inst_set_pinned(hobak_inst* inst, const int32_t* vids, int32_t n, const double* targets);
/* addKinematicCollisionObject (src/Timestepper/TIMESTEPPER.cpp:433-441) of a
 * CUBE(center, scale) (src/Geometry/CUBE.cpp:23-30): from the next step on, surface vertices
 * that enter the cube moving into it become Baraff-Witkin plane constraints, which slide on its
 * nearest face and release when the residual force pulls away (TIMESTEPPER.cpp:252-329,
 * 446-586; BACKWARD_EULER_POSITION.cpp:85-110, 163-314).  FP64 form with one Newton iteration
 * per step; at most HOBAK_MAX_COLLISION_OBJECTS per instance, tested in the order added. */
#define HOBAK_MAX_COLLISION_OBJECTS 8
int hobak_inst_add_collision_cube(hobak_inst* inst, const double* center3, double scale);
/* The vertices under a plane constraint after the last step (planeConstraints(),
 * TIMESTEPPER.h:46): *count of them, and the first `cap` ids, ascending, into ids (may be NULL). */
int hobak_inst_plane_constraints(hobak_inst* inst, int32_t* count, int32_t* ids, int32_t cap);
/* Damping model.  HOBAK_DAMPING_RAYLEIGH (the default) is solveRayleighDamped with the params'
 * alpha and beta.  HOBAK_DAMPING_GREEN sets the integrator's damping object to
 * VOLUME::GREEN_DAMPING(mu) (src/Damping/Volume/GREEN_DAMPING.cpp), which makes solve() take
 * solveEnergyDamped (BACKWARD_EULER_POSITION.cpp:125-131, 332-473):
 *   b = (M/dt) v + R + f_ext, R the elastic force plus the damping force of the stress
 *       mu F (F^T Fdot + Fdot^T F) at the instance's held velocity gradients Fdot;
 *   A = M/dt^2 - K, with no Rayleigh terms.
 * The reference assembles the damping stiffness C as well, but computeCollisionResponse empties
 * it before A is formed (TIMESTEPPER.cpp:657), so it does not reach the system; it is available
 * through hobak_eval_damping.  The step never updates Fdot: it is zero until
 * hobak_inst_compute_fdots sets it.  One Newton iteration per step only (the reference's BDF_1
 * has no energy-damped solve, BDF_1.cpp:244-255); plane constraints work with either model. */
#define HOBAK_DAMPING_RAYLEIGH 0
#define HOBAK_DAMPING_GREEN 1
int hobak_inst_set_damping(hobak_inst* inst, int32_t model, double mu);
/* TET_MESH::computeFdots(velocity) (src/Geometry/TET_MESH.cpp:660-684): the held per-tet
 * velocity gradients Fdot = [v1 - v0, v2 - v0, v3 - v0] DmInv from a DOF-length host velocity,
 * or from the instance's current velocity when velocity is NULL.  Needs HOBAK_DAMPING_GREEN. */
int hobak_inst_compute_fdots(hobak_inst* inst, const double* velocity);
/* position() = u and velocity() = v, DOF-length host arrays; NULL leaves that vector unchanged. */
int hobak_inst_set_state(hobak_inst* inst, const double* u, const double* v);
int hobak_inst_get_state(hobak_inst* inst, double* u, double* v);

/* One SIMULATION_SCENE::stepSimulation (SIMULATION_SCENE.h:120-126): f_ext = one-ring volume *
 * gravity, then BACKWARD_EULER_POSITION::solve (or the BDF_1 loop when newton_iters > 1). */
int hobak_step(hobak_inst* inst, hobak_step_stats* stats);

/* Advance n independent instances by one step each, batched on the device: every kernel is one
 * segmented launch over all of them, and the block-diagonal PCG masks out converged instances.
 * Instances of one precision form run together (FP64 ones first, then FP32 ones); all share the
 * context's slice height.  stats: n entries.  hobak_step is the batch of one.  There is no
 * reference counterpart: the reference steps one mesh per solve() call. */
int hobak_step_batch(hobak_ctx* ctx, hobak_inst* const* insts, int32_t n, hobak_step_stats* stats);

/* Parity hooks.  BSR pattern of the stiffness (block rows by vertex, columns ascending). */
int hobak_get_bsr_pattern(hobak_inst* inst, int32_t* nnzb, int32_t* row_ptr, int32_t* col_idx);
/* computeHyperelasticForces and computeHyperelasticClampedHessian at displacement u:
 * forces (DOFs) and K (nnzb*9, row-major blocks, the reference's sign: K = -stiffness).
 * Either output may be NULL. */
int hobak_eval_forces_hessian(hobak_inst* inst, const double* u, double* forces, double* kvals);
/* No counterpart in the reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  The element forces
 * of hobak_eval_forces_hessian (computeHyperelasticForces, TET_MESH.cpp:736-770), from a
 * displacement held in device memory into device memory, allocating nothing: u_dev and
 * forces_dev are 3 nv doubles on the instance's device, in the caller's vertex numbering.  The
 * per-tet force staging lives in one static device array of HOBAK_NOALLOC_SCRATCH_BYTES per GPU,
 * so an instance may have at most HOBAK_NOALLOC_SCRATCH_BYTES / (12 * real size) tets (FP64:
 * 524,288; FP32: 1,048,576), and must not be relabelled (HOBAK_TUNE_ROW_SORT, HOBAK_TUNE_AGG,
 * which the fp32 and fp32_batch tuning entries of the Python layer set); otherwise
 * HOBAK_E_INVALID with the limit in hobak_last_error.  Every call on a GPU uses that GPU's one
 * array: calls that overlap on one GPU (from several host threads or contexts) race.  Returns
 * once the forces are written. */
#define HOBAK_NOALLOC_SCRATCH_BYTES (48u << 20)
int hobak_eval_forces_noalloc(hobak_inst* inst, const double* u_dev, double* forces_dev);
/* With energy damping: computeInternalForces(hyperelastic, damping) (elastic plus damping force,
 * TET_MESH.cpp:816-858) and TET_MESH_FASTER::computeDampingHessian (TET_MESH_FASTER.cpp:138-180)
 * at displacement u and the held Fdot: forces (DOFs) and C (nnzb*9, row-major blocks, the
 * reference's sign: C = -vol dF/dx^T dP/dFdot dF/dx).  Either output may be NULL. */
int hobak_eval_damping(hobak_inst* inst, const double* u, double* forces, double* cvals);
/* The last Newton iteration's projected system: rhs = S b (DOFs) and LHS = S A S + (I - S)
 * (nnzb*9).  Either output may be NULL. */
int hobak_get_last_system(hobak_inst* inst, double* rhs, double* lhs_vals);

/* Device timing of the last hobak_step: milliseconds spent in the element kernel, assembly,
 * right-hand side, PCG and state update (HIP events).  out: 5 doubles. */
int hobak_last_step_timing(hobak_inst* inst, double* out_ms);

/* Measurement hooks (no counterpart in the reference).
 * hobak_set_kernel_sampling: every `every`-th PCG iteration of the following steps brackets its
 * SpMV launch with HIP events on the instance's stream (0 turns sampling off and clears the
 * accumulators).  A group of PCG iterations that holds a sampled launch runs uncaptured; the
 * other groups replay their graph.  hobak_kernel_stats: average duration (ms) and count of the sampled SpMV
 * launches that did work, and the SpMV's algorithmic bytes per launch
 * nnzb * (9 s + 4) + 4 (nv + 1) + 2 * 3 s * nv  (s = bytes per real). */
int hobak_set_kernel_sampling(hobak_inst* inst, int32_t every);
int hobak_kernel_stats(hobak_inst* inst, double* spmv_avg_ms, int64_t* spmv_count,
                       double* spmv_bytes);
/* Kernel and engine tuning (no counterpart in the reference).
 *   HOBAK_TUNE_SLICE_ROWS  rows per block-matrix slice: 8, 16, 32 or 64 (64 / rows lanes per row
 *                          in the SpMV); fixed once the context holds instances
 *   HOBAK_TUNE_POLL_CHUNK  PCG iterations per queued group (rounded up to even); groups are
 *                          captured once per member set as a HIP graph
 *   HOBAK_TUNE_MAX_CHUNKS  most groups queued between two host polls
 *   HOBAK_TUNE_GRAPHS      1: replay captured graphs; 0: launch every kernel directly
 *   HOBAK_TUNE_GRID_CAP    most workgroups per launch (0: one per virtual workgroup); capped
 *                          launches stride over their virtual workgroups (persistent form)
 *   HOBAK_TUNE_UPDATE_WGS  most workgroups of the PCG update kernel per instance (each is one
 *                          partial sum the next SpMV re-reduces); applies to instances created
 *                          afterwards
 *   HOBAK_TUNE_SPMV_PAD    1: the SpMV walks every slice's padded width with a uniform trip
 *                          count; 0: each lane walks only its own row's blocks
 *   HOBAK_TUNE_SPMV_STREAM_MB  while the matrices of the instances still solving move more than
 *                          this many MiB per SpMV launch, the SpMV reads them with non-temporal
 *                          loads (default 384, 1.5x the 256 MB Infinity Cache: below it much of
 *                          the matrix is still re-read from the cache; above it streaming keeps
 *                          the gathered vectors cached instead).  Decided per group of PCG
 *                          iterations, so a batch switches to cached loads as its instances
 *                          converge; 0 streams always
 *   HOBAK_TUNE_ROW_SORT    1: instances created afterwards hold their mesh with the vertices
 *                          relabelled by descending block-row length (within windows of 2048
 *                          vertices), so that the rows of a slice have nearly equal length and
 *                          the slice padding the SpMV streams shrinks; every entry point keeps
 *                          the caller's vertex numbering.  Row sums are unchanged; the PCG dot
 *                          products are summed in another order.  0 (default): as given
 *   HOBAK_TUNE_COMPACT_PCT 0..100: a batch's PCG groups run over all of its instances (one
 *                          captured graph) while the instances still solving hold more than this
 *                          percentage of its SpMV workgroups, and over those instances only
 *                          (launched directly, no graph) at or below it, so that finished
 *                          instances' workgroups are not dispatched.  Results do not depend on
 *                          it.  Default 100: compact as soon as one instance has finished;
 *                          0 never compacts
 *   HOBAK_TUNE_UPDATE_GROUPS  groups of 21 rows that one wave of the PCG update kernel covers
 *                          (it issues all their loads before using any): 1, 2 (default) or 4;
 *                          fixed once the context holds instances.  The r.r and r.z partial
 *                          sums follow the wave layout, so FP64 results differ in the last bits
 *   HOBAK_TUNE_P_FORM      FP64 PCG iterations in three kernels: the previous iteration is
 *                          finished and p_new = z + beta p_old formed once per entry, then the
 *                          SpMV gathers p_new alone (the two-kernel form gathers z and p_old at
 *                          every block).  0: never; 1 (default): for groups whose matrices stream
 *                          (see HOBAK_TUNE_SPMV_STREAM_MB); 2: always.  Results do not depend on
 *                          it (bitwise)
 *   HOBAK_TUNE_COARSE      reduced-precision form: the inner PCG's preconditioner adds to block
 *                          (or point) Jacobi a coarse correction over the rigid translations,
 *                          z = M^-1 r + P (P^T A P)^-1 P^T r with P one 3x3 identity per vertex,
 *                          for instances with no pinned vertex, in the position form and without
 *                          energy damping (others are unaffected).  The refined solve still stops
 *                          at the same FP64 residual; fewer inner iterations get there.  No
 *                          counterpart in the reference, whose PCG preconditions with the
 *                          diagonal only.  1 (default) on, 0 off; fixed once the context holds
 *                          instances.  The FP64 form never uses it
 *   HOBAK_TUNE_AGG         reduced-precision form: 1 holds instances created afterwards in a
 *                          spatial (Hilbert-curve) vertex order, and wherever HOBAK_TUNE_COARSE does
 *                          not apply (an instance with pinned vertices, or COARSE 0) adds local
 *                          aggregates to block (or point) Jacobi: every chunk of 21 *
 *                          HOBAK_TUNE_UPDATE_GROUPS consecutive vertices of that order is an
 *                          aggregate I, and z = M^-1 r + sum_I P_I (P_I^T A P_I)^-1 P_I^T r with
 *                          P_I the translations of I's free vertices.  With HOBAK_TUNE_ROW_SORT
 *                          also, each chunk's vertices are ordered by descending row length.
 *                          Every entry point keeps the caller's numbering (as for ROW_SORT,
 *                          hobak_eval_forces_noalloc refuses a relabelled instance).  The refined
 *                          solve stops at the same FP64 residual.  No counterpart in the
 *                          reference.  0 (default) off; 2 as 1, and for instances where
 *                          HOBAK_TUNE_COARSE applies also adds the aggregates to its correction
 *                          (both at once: for bodies resting on the ground, where the rigid
 *                          translation alone no longer helps).  Fixed once the context holds
 *                          instances */
#define HOBAK_TUNE_SLICE_ROWS 0
#define HOBAK_TUNE_POLL_CHUNK 2
#define HOBAK_TUNE_MAX_CHUNKS 3
#define HOBAK_TUNE_GRAPHS 4
#define HOBAK_TUNE_GRID_CAP 5
#define HOBAK_TUNE_UPDATE_WGS 6
#define HOBAK_TUNE_SPMV_PAD 7
#define HOBAK_TUNE_SPMV_STREAM_MB 8
#define HOBAK_TUNE_ROW_SORT 9
#define HOBAK_TUNE_COMPACT_PCT 10
#define HOBAK_TUNE_UPDATE_GROUPS 11
#define HOBAK_TUNE_P_FORM 12
#define HOBAK_TUNE_COARSE 13
#define HOBAK_TUNE_AGG 14
int hobak_ctx_set_tuning(hobak_ctx* ctx, int32_t key, int32_t value);

/* Continuous-batching scheduler (no counterpart in the reference, whose callers step one scene
 * at a time through SIMULATION_SCENE::stepSimulation, SIMULATION_SCENE.h:120-126).
 * A request advances one instance by one step.  Submitting attaches the instance to the
 * scheduler (its state stays resident in HBM); every tick finishes the solves that ended, starts
 * the requests that have arrived since (their prologue kernels run as one segmented launch), runs
 * groups of PCG iterations over the instances still solving (one captured graph over every
 * attached instance while all of them solve; see HOBAK_TUNE_COMPACT_PCT) and polls.  Requests
 * that arrive while others are on the GPU thus join the running batch at the next tick.
 * Requests of one instance run in submission order. */
typedef struct {
  hobak_inst* inst;
  uint64_t tag;
  hobak_step_stats stats;   /* stats.seconds = end_s - submit_s */
  double submit_s;          /* scheduler clock: seconds since hobak_sched_create */
  double start_s;           /* the request's first kernels were queued */
  double end_s;             /* the tick that completed it returned */
} hobak_completion;

int hobak_sched_create(hobak_ctx* ctx, hobak_sched** out);
/* Drops queued and running requests and detaches every instance. */
void hobak_sched_destroy(hobak_sched* sched);
int hobak_sched_submit(hobak_sched* sched, hobak_inst* inst, uint64_t tag);
/* Detaches an idle instance (hobak_inst_destroy detaches by itself). */
int hobak_sched_detach(hobak_sched* sched, hobak_inst* inst);
/* Runs one tick (when any request is queued or running) and returns up to max_out completions;
 * *in_flight counts the requests not yet returned. */
int hobak_sched_tick(hobak_sched* sched, hobak_completion* out, int32_t max_out, int32_t* n_out,
                     int32_t* in_flight);
/* The scheduler clock (seconds since hobak_sched_create). */
double hobak_sched_clock(hobak_sched* sched);
/* Counters since creation, 11 doubles: ticks, PCG groups queued, GPU milliseconds of prologues,
 * of PCG groups and of epilogues, host milliseconds inside hobak_sched_tick, solves finished,
 * SpMV launches queued, their algorithmic bytes summed (per launch: the matrices and vectors of
 * the instances whose solve was still running in it, nnzb (9 s + 4) + 4 (nv + 1) + 2 * 3 s nv
 * each; the SpMV counts those launches per instance on the device), PCG groups
 * replayed from a captured graph, and PCG groups launched directly (over the still-solving
 * instances only, or holding a sampled launch).
 * The engine also marks its work with roctx ranges ("hobak_step", "hobak_step_batch",
 * "hobak_sched_tick" and, inside, "hobak prologue", "hobak pcg groups", "hobak epilogue",
 * "hobak poll"), which rocprofv3 --marker-trace records. */
int hobak_sched_stats(hobak_sched* sched, double* out);
/* Measurement hook of a scheduler: every `every`-th SpMV launch of its PCG groups is bracketed
 * with HIP events on the context's stream (its group runs uncaptured); 0 turns it off.  Setting
 * it clears the counters.  hobak_sched_kernel_stats returns the average sampled launch time, the
 * number of samples and the algorithmic bytes per SpMV launch averaged over every launch since
 * sampling was set (an instance's matrix and vectors count in the launches that streamed them,
 * as counted on the device; the workgroups of an instance whose solve has ended return at once).
 * Bytes over the average time is then the rate over all launches of the window, the samples
 * being spread evenly over them. */
int hobak_sched_set_kernel_sampling(hobak_sched* sched, int32_t every);
int hobak_sched_kernel_stats(hobak_sched* sched, double* spmv_avg_ms, int64_t* spmv_count,
                             double* spmv_bytes);

/* Streaming copy of `bytes` device bytes (a calibration kernel for hardware byte counters);
 * returns the kernel time in ms. */
int hobak_copy_probe(hobak_ctx* ctx, size_t bytes, int32_t reps, double* ms);

#ifdef __cplusplus
}
#endif
#endif /* HOBAK_C_H */
