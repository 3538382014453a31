// Kernel argument blocks and host-side launchers shared by the translation units of
// libhobak_hip.so.  Device layouts (all resident in HBM):
//
//   vectors        DOF-ordered AoS: entry 3*v + c (the reference's flat VECTOR order)
//   tets           int4 per tet
//   DmInv          9 planes of nt doubles (plane 3*r + c holds DmInv(r, c))       -> coalesced
//   element stage  90 planes of nt (10 upper 3x3 blocks of the 12x12 stiffness) + 12 planes of
//                  forces; plane index (block * 9 + 3 * a + b)
//   block matrix   SELL-8, k-interleaved: rows (vertices) in slices of 8; slice s owns slots
//                  [slice_off[s], slice_off[s+1]) (a multiple of 64), and the k-th block of row r
//                  sits at slot = slice_off[r / 8] + 8 * k + (r % 8).  A wavefront that gives
//                  each of a slice's 8 rows 8 lanes (lane = 8 * (k % 8) + r % 8) therefore reads
//                  slot_off + 64 * (k / 8) + lane: contiguous 512-byte FP64 wave loads.
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

constexpr int kSlice = 8;     // rows per SELL slice
constexpr int kLanesPerRow = 8;
constexpr int kBlock = 256;
constexpr int kStageBlocks = 10;
constexpr int kStagePlanes = 90;

// Scalars of one PCG solve, resident on the device so the loop runs without host round trips.
struct CgScalars {
  double rhsNorm2;
  double threshold;
  double absNew;
  double alpha;
  double beta;
  double resNorm2;
  double maxR;        // |S b|^2, the BDF_1 Newton test value
  int iters;         // also selects the p double buffer: iteration i reads p[i & 1]
  int maxIters;
  int done;           // 1 once converged / exhausted / failed
  int status;         // 0 ok, 1 not converged, 2 non-finite
  int pad[4];
};

struct Reduce {
  double* partial;    // [nvals * grid]
  unsigned* counter;  // one per reduction site
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
  void* stage_h;        // 90 planes of nt, T
  void* stage_f;        // 12 planes of nt, T
  int with_hessian;
};


// Launchers (defined in the .hip files).  T is double or float.
template <typename T>
hipError_t launch_element(const ElementArgs& a, hipStream_t s);


}  // namespace hobak
