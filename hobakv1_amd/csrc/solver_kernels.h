// Argument blocks and launchers of the assembly / right-hand side / PCG / state kernels.
#pragma once
#include "kernels.h"

namespace hobak {

// Row-major 3x3 block-Jacobi inverse per row (9 values) or Jacobi 1/a_ii per DOF.
constexpr int kMinvStride = 9;

struct AsmArgs {
  int nrows, nslots, nt;
  const int* slot_row;   // owning row of each slot, -1 for padding
  const int* col;
  const int* cptr;       // nslots + 1: slot s sums stage_h blocks [cptr[s], cptr[s + 1])
  const void* stage_h;   // T
  const void* D;         // T, 9 planes (mode 1)
  void* out;             // T, 9 planes
  int mode;              // 0 plain stiffness sum, 1 system matrix + preconditioner
  const unsigned char* pinned;
  void* minv;            // T
  int precond;
  int ground_on;
  double ground_y, ground_k;
  const double* rest;
  const double* u;
};

struct ConstArgs {
  int nslots;
  const int* slot_row;
  const int* col;
  const void* kneg;      // assembled -K_rest, 9 planes, T
  const double* mass;    // per DOF
  double ra, rb, invDt, invDt2;
  void* D;
  void* E;
  void* C;
};

struct RhsArgs {
  int nrows, nslots, nt;
  int slice_shift;       // log2 of rows per slice
  const int* slice_off;
  const int* row_len;
  const int* col;
  const int* vptr;       // vertex r sums stage_f 3-vectors [vptr[r], vptr[r + 1])
  const void* stage_f;   // T
  const void* Mvec;      // T: E (BEP) or C (BDF_1)
  const double* vel;
  const double* u;
  const double* uOld;
  const double* vOld;
  const double* mass;
  const double* one_ring;
  const double* rest;
  double g[3];
  double invDt, invDt2;
  int bdf;
  const unsigned char* pinned;
  int ground_on;
  double ground_y, ground_k;
  double* rhs64;
  int init_cg;           // also seed the PCG vectors (x = 0, r = rhs, z = p = M^-1 r)
  void* r;
  void* x;
  void* z;
  void* p0;              // zeroed: iteration 0 forms p = z + 0 * p0
  const void* minv;
  int precond;
  Partials out;          // per workgroup: a = |S b|^2, b = r.z
};

// Single-workgroup kernel that turns the right-hand side's partials into the PCG's first
// scalar set (ConjugateGradient.h:45-66).
struct InitArgs {
  Partials in;
  CgScalars* sc;         // sc[0] is written
  double tol;
  int max_iters;
};

struct CgArgs {
  int nrows, nslots;
  int slice_shift;       // log2 of rows per slice
  const int* slice_off;
  const int* row_len;
  const int* col;
  const void* A;
  void* x;
  void* r;
  void* z;
  void* p0;              // p double buffer: iteration it reads p[it & 1], writes p[(it + 1) & 1]
  void* p1;
  void* t;
  const void* minv;
  int precond;
  CgScalars* sc;         // two sets, see CgScalars
  int it;                // host iteration counter
  Partials pt;           // SpMV -> update: p.t
  Partials rz;           // update -> next SpMV: a = r.r, b = r.z
};

// FP64 iterative refinement around the FP32 PCG (the precision form).
struct RefineState {
  double b2;      // |S b|^2
  double res2;    // |rhs - A x64|^2 of the current iterate
  int done;       // res2 < tol^2 b2 (or b2 == 0), or a non-finite residual
  int status;     // 0 ok, 2 non-finite
  int sweeps;
  int pad;
};

struct RefineArgs {
  int nrows, nslots, slice_shift;
  const int* slice_off;
  const int* row_len;
  const int* col;
  const void* A;         // T
  const double* rhs64;
  double* x64;
  double* res64;
  void* r;               // T inner PCG vectors
  void* x;
  void* z;
  void* p0;
  const void* minv;
  int precond;
  Partials out;          // a = |.|^2, b = r.z
  RefineState* state;
  double tol;            // outer tolerance on |res| / |b|
};

hipError_t launch_begin_step(int nv, double* u, const double* v, double* uOld, double* vOld,
                             const unsigned char* pinned, const double* target, hipStream_t s);
hipError_t launch_end_step(int nv, double invDt, const double* u, const double* uOld, double* v,
                           const unsigned char* pinned, hipStream_t s);
template <typename T>
hipError_t launch_apply_dx(int n, double* u, const T* x, hipStream_t s);
template <typename T>
hipError_t launch_assemble(const AsmArgs& a, hipStream_t s);
template <typename T>
hipError_t launch_constant_parts(const ConstArgs& a, hipStream_t s);
template <typename T>
hipError_t launch_rhs_init(const RhsArgs& a, hipStream_t s);
hipError_t launch_cg_init(const InitArgs& a, hipStream_t s);
template <typename T>
hipError_t launch_cg_iteration(const CgArgs& a, hipStream_t s);
template <typename T>
hipError_t launch_cg_spmv(const CgArgs& a, hipStream_t s, hipEvent_t start = nullptr,
                          hipEvent_t stop = nullptr);
template <typename T>
hipError_t launch_cg_rest(const CgArgs& a, hipStream_t s);
hipError_t launch_refine_begin(const RefineArgs& a, const Partials& rhs_parts, hipStream_t s);
template <typename T>
hipError_t launch_refine_init(const RefineArgs& a, hipStream_t s);
template <typename T>
hipError_t launch_refine_residual(const RefineArgs& a, hipStream_t s);
hipError_t launch_refine_check(const RefineArgs& a, hipStream_t s);
template <typename T>
hipError_t launch_force_gather(int nv, const int* vptr, const T* stage_f, double* R, hipStream_t s);
hipError_t launch_copy(const void* src, void* dst, size_t bytes, hipStream_t s);
// Workgroup counts of the kernels that publish partial sums (size of each Partials array).
int rhs_grid(int nrows);
int spmv_grid(int nrows, int slice_shift);
int update_grid(int nrows);

}  // namespace hobak
