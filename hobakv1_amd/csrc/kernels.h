// Kernel argument blocks and host-side launchers shared by the translation units of
// libhobak_hip.so.  Device layouts (all resident in HBM):
//
//   vectors        DOF-ordered AoS: entry 3*v + c (the reference's flat VECTOR order)
//   tets           int4 per tet
//   DmInv          9 planes of nt doubles (plane 3*r + c holds DmInv(r, c))       -> coalesced
//   element stage  written directly in the order the assembly reads it: the 3x3 block of vertex
//                  pair (x, y) of tet t goes to stage_h[9 * hdest[16 t + 4 x + y] + 3 a + b], where
//                  hdest is the position of that contribution in its slot's ascending-tet list,
//                  so every slot sums one contiguous run; likewise force 3-vectors go to
//                  stage_f[3 * fdest[4 t + x] + c] in each vertex's ascending-tet list
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
//   scatter map    per slot, the ascending-tet list of (tet, staged block, transpose) codes that
//                  sum into it: code = tet * 32 + transpose * 16 + block
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace hobak {

constexpr int kDefaultSliceRows = 64;
constexpr int kBlock = 256;
constexpr int kStageBlocks = 16;

// Scalars of one PCG solve, resident on the device so the loop runs without host round trips.
// Two copies ping-pong: iteration `it` (counted by the host that queues the launches) reads
// set it & 1 and its last kernel writes set (it + 1) & 1 from one workgroup only, so no kernel
// ever reads a scalar that another of its own workgroups writes.
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
  int done;           // 1 once converged / exhausted / failed
  int status;         // 0 ok, 1 not converged, 2 non-finite
  int pad[4];
};

// Per-workgroup partial sums written by one kernel and reduced, in a fixed order, by every
// workgroup of the next kernel (no atomics, no in-kernel fences).
struct Partials {
  double* a;  // [grid] first value per producing workgroup
  double* b;  // [grid] second value
  int n;      // producing grid size
};

template <typename T>
struct SellView {
  int nrows;
  int nslots;
  const int* slice_off;
  const int* col;
  T* val;  // 9 planes of nslots
};

struct ElementArgs {
  int nt;
  int material;        // 0 SNH, 1 ARAP
  double mu, lam, alpha;  // SNH: lam = lambda + mu, alpha = 1 + mu / lam
  const int4* tets;
  const double* rest;
  const double* u;
  const double* dminv;  // 9 planes of nt
  const double* vol;
  void* stage_h;        // 16 * 9 * nt values, T, in assembly order
  void* stage_f;        // 4 * 3 * nt values, T, in force-gather order
  const int* hdest;     // 16 per tet
  const int* fdest;     // 4 per tet
  int with_hessian;
};


// Launchers (defined in the .hip files).  T is double or float.
template <typename T>
hipError_t launch_element(const ElementArgs& a, hipStream_t s);


}  // namespace hobak
