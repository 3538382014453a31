// TEST INFRASTRUCTURE ONLY -- the CPU FP64 restatement of HOBAK's backward-Euler Newton step.
// Used as the checker by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg; the
// product library never links it.  All matrices crossing this ABI are column-major unless noted.
#pragma once
#include <cstdint>

extern "C" {

// ---- element level ----
void orc_svd_rv(const double* F, double* U, double* S, double* V);
double orc_psi(int material, double mu, double lambda, const double* F);
void orc_pk1(int material, double mu, double lambda, const double* F, double* P);
void orc_hessian(int material, double mu, double lambda, const double* F, double* H9);
void orc_clamped_hessian(int material, double mu, double lambda, const double* F, double* H9);
void orc_clamped_hessian_usv(int material, double mu, double lambda, const double* U,
                             const double* S, const double* V, double* H9);
// Per-tet force (12) and stiffness (144, column-major) of one tet with corner positions x (4x3
// row-major), DmInv (column-major) and rest volume.
void orc_tet_force_hessian(int material, double mu, double lambda, const double* x,
                           const double* DmInv, double restVolume, double* f12, double* H144);
// GREEN_DAMPING(mu): psi(F, Fdot), PK1(F, Fdot) and dP/dFdot (which does not depend on Fdot),
// optionally eigen-clamped.
double orc_green_psi(double mu, const double* F, const double* Fdot);
void orc_green_pk1(double mu, const double* F, const double* Fdot, double* P);
void orc_green_hessian(double mu, const double* F, int clamped, double* H9);

// ---- mesh / integrator ----
// precond: 0 = Jacobi (the reference's DiagonalPreconditioner), 1 = 3x3 block Jacobi.
void* orc_create(const double* rest_xyz, int nv, const int* tets, int nt, int material,
                 double mu, double lambda, double dt, double rayleigh_alpha, double rayleigh_beta,
                 double mass_scale, const double* gravity, const int* pinned, int npinned,
                 const double* pin_targets, double pcg_tol, int pcg_max_iters, int precond,
                 int newton_iters, double newton_tol);
void orc_destroy(void* h);
void orc_set_state(void* h, const double* u, const double* v);
void orc_get_state(void* h, double* u, double* v);
void orc_set_ground(void* h, double y, double k);
void orc_add_collision_cube(void* h, const double* center, double scale);
int orc_plane_constraints(void* h, int* vertex_ids);
int orc_surface_vertices(void* h, int* ids);
// One SIMULATION_SCENE::stepSimulation.  Returns total PCG iterations; iters_per_newton gets one
// count per Newton iteration (array of newton_iters, may be NULL).
int orc_step(void* h, double* pcg_error, double* seconds, int* iters_per_newton,
             int* newton_done);
int orc_nnzb(void* h);
void orc_bsr_pattern(void* h, int* row_ptr, int* col_idx);
// R and the assembled stiffness K (block values [nnzb][3][3], row-major blocks) at displacement u.
void orc_forces_hessian(void* h, const double* u, double* R, double* Kvals);
// The last step's right-hand side rhs = S b and system matrix (LHS blocks, row-major).
void orc_last_system(void* h, double* rhs, double* lhs_vals);
// Energy damping: GREEN_DAMPING(mu) replaces Rayleigh damping (single-iteration handles only;
// returns -1 otherwise).  compute_fdots sets the held velocity gradients from a velocity.
int orc_set_green_damping(void* h, double mu);
void orc_compute_fdots(void* h, const double* vel);
// Elastic + damping forces and the assembled damping stiffness C at displacement u.
void orc_damped_forces_hessian(void* h, const double* u, double* R, double* Cvals);
void orc_rest_volumes(void* h, double* vol, double* one_ring);
// BACKWARD_EULER_VELOCITY (BACKWARD_EULER_VELOCITY.cpp:107-235) in place of the position form:
// single-iteration Rayleigh handles without the penalty ground (returns -1 otherwise).
int orc_set_velocity_form(void* h);
void orc_set_threads(int n);
}
