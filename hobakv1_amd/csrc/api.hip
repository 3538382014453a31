// Host side of libhobak_hip.so: mesh set-up (sparsity, scatter map), device residency, the
// batched step engine (plans, ticks, captured graphs, continuous batching) and the C ABI declared
// in include/hobak_c.h.
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
//   BACKWARD_EULER_VELOCITY::solveRayleighDamped
//                                         src/Timestepper/BACKWARD_EULER_VELOCITY.cpp:107-235
//   SIMULATION_SCENE::stepSimulation      src/Scenes/SIMULATION_SCENE.h:120-126
//
// Engine.  Every step, batched or not, runs through one tick loop:
//   1. instances whose solve finished at the last poll get x applied, then either end the step
//      (velocity update, request complete) or re-enter for the next BDF_1 Newton iteration;
//   2. newly started and continuing instances run the prologue (element, assembly, right-hand
//      side, PCG / refinement seed) as one segmented launch per kernel over just those instances;
//   3. groups of PCG iterations run over the scheduler's whole member set (a captured graph per
//      member set) while every member is solving, and over a table of the solving members only
//      (direct launches) once some have finished (compact_pct); kernels of instances that are
//      not solving return at once;
//   4. one status kernel and one small copy, then the host waits and reads the statuses.
// Requests submitted while others are on the GPU join at the next tick: continuous batching.
#include <hip/hip_runtime.h>
#include <rocprofiler-sdk-roctx/roctx.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <mutex>
#include <new>
#include <set>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/hobak_c.h"
#include "kernels.h"

using namespace hobak;

namespace {
struct Plan;
constexpr double kEps = 2.220446049250313e-16;  // Eigen's default CG tolerance (machine epsilon)
constexpr int kMaxPlans = 32;
constexpr int kMaxSamples = 512;
}  // namespace

struct hobak_ctx {
  int device = 0;
  int slice_rows = kDefaultSliceRows;  // tuning, fixed once instances exist
  int chunk_iters = 16;                // PCG iterations per queued group (even)
  int max_chunks = 64;                 // groups queued between host polls
  int use_graphs = 1;
  int grid_cap = 0;                    // max workgroups per launch (0: one per virtual workgroup)
  int update_cap = kUpdateCap;         // per-instance workgroups of the PCG update kernel
  int spmv_pad = 0;                    // SpMV loop over padded slice widths (1) or row lengths (0)
  int spmv_stream_mb = 384;            // non-temporal matrix loads above this SpMV size (MiB)
  int row_sort = 0;                    // relabel new instances' vertices by descending row length
  int compact_pct = 100;               // PCG groups launch over the solving instances only below this
  int update_groups = 2;               // 21-row groups per PCG update wave
  int p_form = 1;                      // FP64 p-form iterations: 0 never, 1 with streamed matrices, 2 always
  int coarse = 1;                      // reduced-precision form: rigid-translation coarse correction
  int agg = 0;                         // reduced-precision form: spatial order and local aggregates (2: also with the coarse correction)
  hipStream_t stream = nullptr;
  size_t budget = 0;
  size_t in_use = 0;
  std::string err;
  std::set<hobak_inst*> insts;
  std::set<hobak_sched*> scheds;
  std::vector<Plan*> plans;  // cached member plans, least recently used evicted
  uint64_t stamp = 0;
  // tables of tick-local batches; reused whenever the stream is idle
  char* arena_dev = nullptr;
  char* arena_host = nullptr;
  size_t arena_cap = 0;
  size_t arena_off = 0;
  hipEvent_t ev[8] = {};
  std::vector<hipEvent_t> sample_ev;  // pairs
  void* scratch = nullptr;  // this device's copy of `scratch` (hobak_eval_forces_noalloc)
};

// The device memory of hobak_eval_forces_noalloc, the entry point that allocates nothing: the
// per-tet force staging (12 reals per tet) of one call.  Every GPU the library runs on holds its
// own copy until the process ends; calls on one GPU share it, and the last call's contents stay
// (nothing in it needs to start at zero: a call writes every entry it reads).  Host code never
// uses this name as an address: hobak_ctx_create looks up each device's copy with
// hipGetSymbolAddress (scratch_of_device) and kernels receive that pointer.
static __device__ __align__(256) unsigned char scratch[HOBAK_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The address of `scratch` on `device` (which must be current), looked up once per device.
int scratch_of_device(int device, void** out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (device < 0 || device >= kMaxDevices) return HOBAK_E_NODEVICE;
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return HOBAK_E_HIP;
    g_scratch_ptr[device] = p;
  }
  *out = g_scratch_ptr[device];
  return HOBAK_OK;
}
}  // namespace

namespace {

struct Pattern {
  int nv = 0, nt = 0, nslices = 0, nslots = 0, nrows_pad = 0, nnzb = 0;
  int H = 16, shift = 4;  // rows per slice and its log2
  std::vector<int> row_ptr, col_idx;  // BSR
  std::vector<int> slice_off, row_len, col, slot_row, bsr_slot, diag_slot;
  std::vector<int> cptr, cidx;        // per upper slot: its contributions (pair * nt + t | flag)
  std::vector<int> mirror;            // per upper off-diagonal slot: its transposed slot
  std::vector<int> aslot;             // the upper slots in slot order (the assembly's list)
  std::vector<int> evptr, evert;      // per element workgroup: its tets' distinct vertices
  std::vector<unsigned short> eloc;   // per tet corner: its position in that list
  std::vector<int> vptr, fidx;        // per vertex: its incident corners (x * nt + t)
};

// Block sparsity (BSR rows by vertex, columns ascending), its SELL-H slot layout, and the
// contribution lists of the assembly and force gathers, built from the vertex -> tet incidence
// lists (ascending tet) with marker arrays: O(16 nt) with no searches.  With order_key (a
// relabelled mesh: the caller's id of each vertex), a row's columns are ordered by that key, so
// every row sum runs over the same blocks in the same order as for the mesh as the caller gave it.
void build_pattern(int nv, int nt, const int32_t* tets, int H, Pattern& P,
                   const std::vector<int>* order_key = nullptr) {
  P.nv = nv;
  P.nt = nt;
  P.H = H;
  P.shift = 0;
  while ((1 << P.shift) < H) P.shift++;
  const int L = 64 / H;
  // vertex -> incident tets, ascending, and the corner of each (the force-gather lists)
  P.vptr.assign(nv + 1, 0);
  for (size_t i = 0; i < 4 * (size_t)nt; i++) P.vptr[tets[i] + 1]++;
  for (int i = 0; i < nv; i++) P.vptr[i + 1] += P.vptr[i];
  std::vector<int> inc(P.vptr[nv]);
  std::vector<unsigned char> corner(P.vptr[nv]);
  P.fidx.assign(4 * (size_t)nt, 0);
  {
    std::vector<int> fill(P.vptr.begin(), P.vptr.end() - 1);
    for (int t = 0; t < nt; t++)
      for (int a = 0; a < 4; a++) {
        const int v = tets[4 * (size_t)t + a];
        const int q = fill[v]++;
        inc[q] = t;
        corner[q] = (unsigned char)a;
        P.fidx[q] = a * nt + t;
      }
  }
  // rows: distinct vertices of each vertex's incident tets, ascending
  std::vector<int> mark(nv, -1), tmp;
  tmp.reserve(128);
  P.row_ptr.assign(nv + 1, 0);
  P.col_idx.clear();
  P.col_idx.reserve((size_t)nv * 16);
  for (int i = 0; i < nv; i++) {
    tmp.clear();
    for (int q = P.vptr[i]; q < P.vptr[i + 1]; q++) {
      const int32_t* T = tets + 4 * (size_t)inc[q];
      for (int b = 0; b < 4; b++)
        if (mark[T[b]] != i) {
          mark[T[b]] = i;
          tmp.push_back(T[b]);
        }
    }
    if (order_key)
      std::sort(tmp.begin(), tmp.end(), [&](int a, int b) { return (*order_key)[a] < (*order_key)[b]; });
    else
      std::sort(tmp.begin(), tmp.end());
    P.col_idx.insert(P.col_idx.end(), tmp.begin(), tmp.end());
    P.row_ptr[i + 1] = (int)P.col_idx.size();
  }
  P.nnzb = P.row_ptr[nv];
  // SELL-H slices, width rounded up to a multiple of the lanes per row
  P.nslices = (nv + H - 1) / H;
  P.nrows_pad = P.nslices * H;
  P.slice_off.assign(P.nslices + 1, 0);
  P.row_len.assign(nv, 0);
  for (int s = 0; s < P.nslices; s++) {
    int w = 0;
    for (int r = s * H; r < std::min(nv, (s + 1) * H); r++) {
      P.row_len[r] = P.row_ptr[r + 1] - P.row_ptr[r];
      w = std::max(w, P.row_len[r]);
    }
    w = (w + L - 1) / L * L;
    P.slice_off[s + 1] = P.slice_off[s] + w * H;
  }
  P.nslots = P.slice_off[P.nslices];
  P.col.assign(P.nslots, 0);
  P.slot_row.assign(P.nslots, -1);
  P.bsr_slot.assign(P.nnzb, 0);
  P.diag_slot.assign(nv, -1);
  for (int s = 0; s < P.nslices; s++) {
    const int w = (P.slice_off[s + 1] - P.slice_off[s]) / H;
    for (int lane = 0; lane < H; lane++) {
      const int r = s * H + lane;
      for (int k = 0; k < w; k++) {
        const int slot = P.slice_off[s] + k * H + lane;
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
  // contributions: block (x, y) of tet t sums into slot (tets[t][x], tets[t][y]).  Only the upper
  // slots (row <= column) sum their contributions; a lower slot is the transpose of its mirror,
  // whose contributions are the same tets' transposed blocks.  Walking row i's incident tets in
  // ascending order visits every contribution of row i's slots in ascending tet order, which is
  // the order of the reference's gather lists (TET_MESH_FASTER.cpp:213-236); the mirror's sum
  // runs over the same tets in the same order, so it is bitwise the transpose of the lower
  // slot's own sum.
  std::vector<int> pos(nv, 0);
  std::vector<int> cnt(P.nslots + 1, 0);
  auto row_slots = [&](int i) {
    for (int k = P.row_ptr[i]; k < P.row_ptr[i + 1]; k++) pos[P.col_idx[k]] = P.bsr_slot[k];
  };
  // mirrors in one pass (the pattern is symmetric): visiting the rows in the order their columns
  // are sorted by, entry (r, c) of row r finds its transpose (c, r) at row c's next unvisited
  // entry, since row c lists its columns in that same order
  P.mirror.assign(P.nslots, -1);
  {
    std::vector<int> next(P.row_ptr.begin(), P.row_ptr.end() - 1);
    std::vector<int> order(nv);
    for (int i = 0; i < nv; i++) order[i] = i;
    if (order_key)
      std::sort(order.begin(), order.end(), [&](int a, int b) { return (*order_key)[a] < (*order_key)[b]; });
    for (int r : order)
      for (int k = P.row_ptr[r]; k < P.row_ptr[r + 1]; k++) {
        const int c = P.col_idx[k];
        const int m = next[c]++;
        if (c > r) P.mirror[P.bsr_slot[k]] = P.bsr_slot[m];  // upper slot (r, c): its mirror (c, r)
      }
  }
  for (int i = 0; i < nv; i++) {
    row_slots(i);
    for (int q = P.vptr[i]; q < P.vptr[i + 1]; q++) {
      const int32_t* T = tets + 4 * (size_t)inc[q];
      for (int b = 0; b < 4; b++)
        if (T[b] >= i) cnt[pos[T[b]] + 1]++;
    }
  }
  P.cptr.assign(P.nslots + 1, 0);
  for (int s = 0; s < P.nslots; s++) P.cptr[s + 1] = P.cptr[s] + cnt[s + 1];
  P.aslot.clear();
  for (int s = 0; s < P.nslots; s++)
    if (P.slot_row[s] >= 0 && P.col[s] >= P.slot_row[s]) P.aslot.push_back(s);
  // the element kernel's vertex lists: each workgroup of kElementBlock consecutive tets stages its
  // distinct vertices (ascending) in LDS; a corner names its vertex by its place in the list
  {
    const int nblk = (nt + kElementBlock - 1) / kElementBlock;
    P.evptr.assign(nblk + 1, 0);
    P.evert.clear();
    P.evert.reserve((size_t)nt);
    P.eloc.assign(4 * (size_t)nt, 0);
    std::vector<int> seen(nv, -1), place(nv, 0), list;
    list.reserve(kElementVerts);
    for (int w = 0; w < nblk; w++) {
      const int t1 = std::min(nt, (w + 1) * kElementBlock);
      list.clear();
      for (int t = w * kElementBlock; t < t1; t++)
        for (int a = 0; a < 4; a++) {
          const int v = tets[4 * (size_t)t + a];
          if (seen[v] != w) {
            seen[v] = w;
            list.push_back(v);
          }
        }
      std::sort(list.begin(), list.end());
      for (size_t i = 0; i < list.size(); i++) place[list[i]] = (int)i;
      for (int t = w * kElementBlock; t < t1; t++)
        for (int a = 0; a < 4; a++)
          P.eloc[4 * (size_t)t + a] = (unsigned short)place[tets[4 * (size_t)t + a]];
      P.evert.insert(P.evert.end(), list.begin(), list.end());
      P.evptr[w + 1] = (int)P.evert.size();
    }
  }
  P.cidx.assign(P.cptr[P.nslots], 0);  // 10 nt: the 10 pairs of 4 distinct corners per tet
  std::vector<int> fill(P.cptr.begin(), P.cptr.end() - 1);
  for (int i = 0; i < nv; i++) {
    row_slots(i);
    for (int q = P.vptr[i]; q < P.vptr[i + 1]; q++) {
      const int t = inc[q];
      const int32_t* T = tets + 4 * (size_t)t;
      const int x = corner[q];
      for (int y = 0; y < 4; y++) {
        if (T[y] < i) continue;
        // the element stages block (min(x, y), max(x, y)), rows of corner min(x, y): the slot
        // (row T[x]) sums it transposed when x > y
        const int w = staged_pair(std::min(x, y), std::max(x, y)) * nt + t;
        P.cidx[fill[pos[T[y]]]++] = x > y ? (w | kCidxTranspose) : w;
      }
    }
  }
}


}  // namespace

// Block-row lengths (distinct vertices of each vertex's incident tets, itself included).
static std::vector<int> row_lengths(int nv, int nt, const int32_t* tets) {
  std::vector<int> vptr(nv + 1, 0);
  for (size_t i = 0; i < 4 * (size_t)nt; i++) vptr[tets[i] + 1]++;
  for (int i = 0; i < nv; i++) vptr[i + 1] += vptr[i];
  std::vector<int> inc(vptr[nv]);
  {
    std::vector<int> fill(vptr.begin(), vptr.end() - 1);
    for (int t = 0; t < nt; t++)
      for (int a = 0; a < 4; a++) inc[fill[tets[4 * (size_t)t + a]]++] = t;
  }
  std::vector<int> mark(nv, -1), len(nv, 0);
  for (int i = 0; i < nv; i++)
    for (int q = vptr[i]; q < vptr[i + 1]; q++)
      for (int b = 0; b < 4; b++) {
        const int j = tets[4 * (size_t)inc[q] + b];
        if (mark[j] != i) {
          mark[j] = i;
          len[i]++;
        }
      }
  return len;
}

// The row_sort relabelling: within consecutive windows of kSortWindow vertices, the vertices in
// order of descending block-row length (ties keep their order).  A slice then holds rows of
// (nearly) equal length, so the SELL padding that the SpMV streams shrinks (8.7% -> 1.5% of the
// slots for the 256-instance batch at slice height 64), and the windows keep gathers local.
// perm[p] = the caller's vertex that becomes vertex p.
constexpr int kSortWindow = 2048;
static std::vector<int> row_sort_perm(int nv, int nt, const int32_t* tets) {
  const std::vector<int> len = row_lengths(nv, nt, tets);
  std::vector<int> perm(nv);
  for (int i = 0; i < nv; i++) perm[i] = i;
  for (int w0 = 0; w0 < nv; w0 += kSortWindow) {
    const int w1 = std::min(nv, w0 + kSortWindow);
    std::stable_sort(perm.begin() + w0, perm.begin() + w1, [&](int a, int b) { return len[a] > len[b]; });
  }
  return perm;
}

// Hilbert index of a point with `bits`-bit coordinates: Skilling's transform of the axes to the
// transposed index, then its bits interleaved, most significant first (J. Skilling, "Programming
// the Hilbert curve", AIP Conf. Proc. 707, 381 (2004)).
static uint64_t hilbert_key(uint32_t x[3], int bits) {
  const uint32_t top = 1u << (bits - 1);
  for (uint32_t q = top; q > 1; q >>= 1) {  // undo excess work
    const uint32_t p = q - 1;
    for (int i = 0; i < 3; i++) {
      if (x[i] & q) {
        x[0] ^= p;
      } else {
        const uint32_t t = (x[0] ^ x[i]) & p;
        x[0] ^= t;
        x[i] ^= t;
      }
    }
  }
  for (int i = 1; i < 3; i++) x[i] ^= x[i - 1];  // Gray encode
  uint32_t t = 0;
  for (uint32_t q = top; q > 1; q >>= 1)
    if (x[2] & q) t ^= q - 1;
  for (int i = 0; i < 3; i++) x[i] ^= t;
  uint64_t key = 0;
  for (int j = bits - 1; j >= 0; j--)
    for (int i = 0; i < 3; i++) key = (key << 1) | ((x[i] >> j) & 1u);
  return key;
}

// The local aggregates' relabelling: the vertices in Hilbert-curve order of their rest positions
// (21 bits per axis over the bounding box; ties keep their order), so that every chunk of
// `chunk` consecutive vertices, an aggregate, is spatially compact (more so than Z-curve chunks,
// which jump at octree boundaries).  With by_len, also by descending block-row length within each
// chunk (the slices' padding shrinks as with row_sort, and every aggregate keeps its vertices).
// perm[p] = the caller's vertex that becomes vertex p.
static std::vector<int> spatial_perm(int nv, int nt, const double* xyz, const int32_t* tets, int chunk,
                                     bool by_len) {
  double lo[3], hi[3];
  for (int c = 0; c < 3; c++) lo[c] = hi[c] = xyz[c];
  for (int v = 1; v < nv; v++)
    for (int c = 0; c < 3; c++) {
      lo[c] = std::min(lo[c], xyz[3 * (size_t)v + c]);
      hi[c] = std::max(hi[c], xyz[3 * (size_t)v + c]);
    }
  constexpr int kBits = 21;
  constexpr double kCells = (double)(1 << kBits);
  std::vector<uint64_t> key(nv, 0);
  for (int v = 0; v < nv; v++) {
    uint32_t q[3];
    for (int c = 0; c < 3; c++) {
      const double ext = hi[c] - lo[c];
      const double f = ext > 0 ? (xyz[3 * (size_t)v + c] - lo[c]) / ext * kCells : 0.0;
      q[c] = (uint32_t)std::min(kCells - 1.0, std::max(0.0, std::floor(f)));
    }
    key[v] = hilbert_key(q, kBits);
  }
  std::vector<int> perm(nv);
  for (int i = 0; i < nv; i++) perm[i] = i;
  std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return key[a] < key[b]; });
  if (by_len) {
    const std::vector<int> len = row_lengths(nv, nt, tets);
    for (int w0 = 0; w0 < nv; w0 += chunk) {
      const int w1 = std::min(nv, w0 + chunk);
      std::stable_sort(perm.begin() + w0, perm.begin() + w1, [&](int a, int b) { return len[a] > len[b]; });
    }
  }
  return perm;
}

struct hobak_inst {
  hobak_ctx* ctx = nullptr;
  int nv = 0, nt = 0, dofs = 0;
  hobak_material mat{};
  hobak_step_params prm{};
  double lam_eff = 0, alpha = 0;
  int tsize = 8;  // bytes of the element/assembly/PCG real type
  int groups = 1; // 21-row groups per PCG update wave
  bool spatial = false;  // held in the spatial order of the local aggregates (spatial_perm)
  Pattern P;
  std::vector<int> pinned_ids;
  // Vertex relabelling (row_sort): the device holds the mesh with vertex p = the caller's vertex
  // perm[p]; iperm is the inverse.  Empty: the caller's numbering.  ocol_idx / orow_ptr are the
  // block pattern in the caller's numbering and oslot the slot of each of its blocks.
  std::vector<int> perm, iperm, orow_ptr, ocol_idx, oslot;
  std::vector<std::pair<void*, size_t>> allocs;
  size_t bytes = 0;
  InstDev h{};              // host mirror of the device record
  InstDev* dev = nullptr;   // the device record
  Ctl* ctl = nullptr;
  bool has_system = false;
  double last_ms[5] = {0, 0, 0, 0, 0};
  int requests = 0;                 // queued or running requests
  hobak_sched* sched = nullptr;     // scheduler it is attached to, if any
  int pred[HOBAK_MAX_NEWTON];       // PCG iterations of the last step, per Newton iteration
  // sampled SpMV timing (measurement hook)
  int sample_every = 0;
  int first_iters = 0;              // iterations of the last step's first (inner) solve
  double spmv_ms_sum = 0.0, spmv_bytes_sum = 0.0;
  long long spmv_count = 0;
  unsigned work_seen = 0;           // Ctl::spmv_work at the last status poll
};

struct hobak_sched {
  hobak_ctx* ctx = nullptr;
  bool ephemeral = false;       // a private scheduler of one hobak_step / hobak_step_batch call
  std::vector<hobak_inst*> members;
  std::unordered_map<hobak_inst*, int> index;
  bool dirty = true;
  Plan* mplan = nullptr;
  struct Run {
    hobak_inst* I = nullptr;
    uint64_t tag = 0;
    double t_submit = 0, t_start = -1;
    int newton_k = 0;
    int stage = 0;      // 0 needs prologue, 1 solving, 2 solve finished (epilogue next)
    int begin = 1;      // prologue starts a new step (u_old, v_old, pinned targets)
    int done_iters = 0; // PCG iterations of this Newton iteration at the last poll
    hobak_step_stats st{};
  };
  std::deque<Run> queue;   // submitted, not started
  std::vector<Run> active;
  std::vector<hobak_completion> done;
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  double ms[5] = {0, 0, 0, 0, 0};
  // counters since creation: ticks, PCG groups queued, GPU ms of prologues / PCG / epilogues,
  // host wall ms inside ticks, solves finished, SpMV launches, their algorithmic bytes (those of
  // the instances solving in each launch's tick), PCG groups replayed from a captured graph and
  // PCG groups launched directly
  double stat[11] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  // sampled SpMV launches of the current batch: (iteration index since the batch's prologue, slot)
  std::vector<std::pair<int, int>> samples;
  int sample_it = 0;
  bool sampling = false;
  // a long-lived scheduler's own sampling (hobak_sched_set_kernel_sampling): every spmv_every-th
  // SpMV launch is timed and credited with the matrix bytes of the instances solving in its tick
  int spmv_every = 0;
  double spmv_ms = 0.0, spmv_bytes = 0.0;  // sampled launch time; algorithmic bytes of the window
  long long spmv_n = 0, spmv_launches = 0;  // sampled launches; all SpMV launches of the window
};

namespace {

// A profiler range (roctx) over a scope of the step engine: rocprofv3 --marker-trace shows the
// prologue, PCG groups, epilogue and poll of every tick, and every hobak_step / batch / tick call.
struct Range {
  explicit Range(const char* what) { roctxRangePushA(what); }
  ~Range() { roctxRangePop(); }
  Range(const Range&) = delete;
  Range& operator=(const Range&) = delete;
};

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

#define TRY(expr)                    \
  do {                               \
    int rc_ = (expr);                \
    if (rc_ != HOBAK_OK) return rc_; \
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

// Per-vertex vectors (3 entries per vertex) between the caller's vertex numbering and the
// device's (identical unless the instance was relabelled).
const double* to_device_order(const hobak_inst* I, const double* src, std::vector<double>& buf) {
  if (I->perm.empty() || !src) return src;
  buf.resize((size_t)I->dofs);
  for (int p = 0; p < I->nv; p++)
    for (int c = 0; c < 3; c++) buf[3 * (size_t)p + c] = src[3 * (size_t)I->perm[p] + c];
  return buf.data();
}

void to_caller_order(const hobak_inst* I, double* v) {
  if (I->perm.empty() || !v) return;
  const std::vector<double> tmp(v, v + I->dofs);
  for (int p = 0; p < I->nv; p++)
    for (int c = 0; c < 3; c++) v[3 * (size_t)I->perm[p] + c] = tmp[3 * (size_t)p + c];
}

// The block pattern in the caller's numbering (rows ascending, columns ascending) and the slot of
// each of its blocks.
void caller_pattern(hobak_inst* I) {
  const Pattern& P = I->P;
  const int nv = I->nv;
  I->orow_ptr.assign(nv + 1, 0);
  I->ocol_idx.clear();
  I->oslot.clear();
  I->ocol_idx.reserve(P.nnzb);
  I->oslot.reserve(P.nnzb);
  std::vector<std::pair<int, int>> row;
  for (int i = 0; i < nv; i++) {
    const int a = I->perm.empty() ? i : I->iperm[i];
    row.clear();
    for (int k = P.row_ptr[a]; k < P.row_ptr[a + 1]; k++)
      row.emplace_back(I->perm.empty() ? P.col_idx[k] : I->perm[P.col_idx[k]], P.bsr_slot[k]);
    std::sort(row.begin(), row.end());
    for (auto& q : row) {
      I->ocol_idx.push_back(q.first);
      I->oslot.push_back(q.second);
    }
    I->orow_ptr[i + 1] = (int)I->ocol_idx.size();
  }
}

size_t estimate_bytes(int nv, int nt, int nslots_est, int precision) {
  const size_t ts = precision == HOBAK_PRECISION_FP64 ? 8 : 4;
  size_t b = 0;
  b += (size_t)nv * 3 * 8 * 12;                 // rest, u, v, olds, target, rhs, tmp, x64, res64, mass
  b += (size_t)nt * (16 + 9 * 8 + 8);           // tets, dminv, vol
  b += (size_t)nslots_est * (4 + 4 + 9 * ts);   // col + slot_row + kneg
  b += (size_t)nslots_est * 9 * ts * 4;         // A, D, E, C
  b += (size_t)nslots_est * 12 + (size_t)nt * (kStagedBlocks + 4) * 4;  // cptr, mirror, aslot, cidx, fidx
  b += (size_t)nt * (kStagedBlocks * 9 + 12) * ts;  // staging
  b += (size_t)nv * (9 + 18) * ts;              // minv + PCG vectors
  b += (size_t)nt * (4 * 4 + 8 + 4) + (size_t)nv * 16;  // vertex lists, element vertex lists, misc
  b += 64 * 1024;                               // records, control block, partial sums
  if (precision != HOBAK_PRECISION_FP64)        // coarse-correction partial sums, aggregate inverses
    b += ((size_t)6 * (nv / kBlock + 1) + (size_t)3 * kUpdateCap) * 8 + 2 * 256 + (size_t)9 * (nv / 21 + 1) * ts;
  return b;
}

// The reduced-precision form's coarse correction over the rigid translations holds when the
// operator annihilates a translation off its diagonal and nothing pins one: no pinned vertex, the
// position form (its difference-form diagonal is dg) and no energy damping.
int coarse_on(const hobak_ctx* ctx, const hobak_inst* I) {
  const InstDev& h = I->h;
  return ctx->coarse && h.refine && !h.velform && !h.damping && I->pinned_ids.empty() ? 1 : 0;
}

// Local aggregates hold for any operator of the difference form: an instance of the
// reduced-precision form held in the spatial order, wherever the coarse correction is off, and
// with ctx->agg == 2 also added to it (both corrections, for bodies resting on the ground).
int agg_on(const hobak_ctx* ctx, const hobak_inst* I) {
  const InstDev& h = I->h;
  if (!I->spatial || !h.refine || h.contact) return 0;
  return (ctx->agg == 2 || !coarse_on(ctx, I)) ? 1 : 0;
}

// ---------------------------------------------------------------------------------------------
// plans: the device tables of one segmented batch
// ---------------------------------------------------------------------------------------------
constexpr size_t kAlign = 256;
size_t al(size_t x) { return (x + kAlign - 1) & ~(kAlign - 1); }

struct Plan {
  std::vector<hobak_inst*> members;
  int n = 0, tsize = 8, shift = 5, refine = 0, groups = 1;
  bool owned = false;
  char* dev = nullptr;
  char* host = nullptr;
  size_t bytes = 0;
  Batch b[K_COUNT];
  InstStatus* status_dev = nullptr;
  InstStatus* status_host = nullptr;
  int contact = 0;              // some member has collision objects
  int damping = 0;              // some member uses energy damping
  int coarse = 0;               // some member has the coarse correction on
  int agg = 0;                  // some member has local aggregates
  double spmv_bytes = 0.0;      // summed over the members
  hipGraphExec_t chunk_exec[2] = {nullptr, nullptr};  // PCG group with cached / streamed matrix loads
  int chunk_k[2] = {0, 0};
  hipGraphExec_t pro_exec[2] = {nullptr, nullptr};  // prologue without / with the step start
  hipGraphExec_t epi_exec = nullptr;
  uint64_t stamp = 0;
};

size_t plan_bytes(size_t n) {
  return al(sizeof(void*) * n) + K_COUNT * al(sizeof(int) * (n + 1)) + al(sizeof(InstStatus) * n);
}

double spmv_bytes_of(const hobak_inst* I) {
  const double sz = (double)I->tsize;
  return (double)I->P.nnzb * (9.0 * sz + 4.0) + 4.0 * (I->nv + 1) + 2.0 * 3.0 * sz * I->nv;
}

// Lays the tables out in the plan's host staging memory, uploads them and fills the Batch
// descriptors.  All members share the real type and the slice height.
int plan_fill(hobak_ctx* ctx, Plan& P) {
  const int n = P.n;
  size_t off = 0;
  const InstDev** hp = (const InstDev**)(P.host + off);
  const InstDev* const* dp = (const InstDev* const*)(P.dev + off);
  P.spmv_bytes = 0.0;
  P.contact = 0;
  P.damping = 0;
  P.coarse = 0;
  P.agg = 0;
  for (int i = 0; i < n; i++) {
    hp[i] = P.members[i]->dev;
    P.contact |= P.members[i]->h.contact;
    P.damping |= P.members[i]->h.damping;
    P.coarse |= P.members[i]->h.coarse;
    P.agg |= P.members[i]->h.agg;
    P.spmv_bytes += spmv_bytes_of(P.members[i]);
  }
  off += al(sizeof(void*) * n);
  for (int k = 0; k < K_COUNT; k++) {
    int* bo = (int*)(P.host + off);
    bo[0] = 0;
    for (int i = 0; i < n; i++) {
      const long next = (long)bo[i] + P.members[i]->h.nb[k];
      if (next > INT32_MAX) return fail(ctx, HOBAK_E_INVALID, "batch too large for one launch");
      bo[i + 1] = (int)next;
    }
    P.b[k] = Batch{dp, (const int*)(P.dev + off), n == 1 ? P.members[0]->dev : nullptr, n, bo[n]};
    off += al(sizeof(int) * (n + 1));
  }
  P.status_dev = (InstStatus*)(P.dev + off);
  P.status_host = (InstStatus*)(P.host + off);
  HIPCHK(ctx, hipMemcpyAsync(P.dev, P.host, off, hipMemcpyHostToDevice, ctx->stream));
  return HOBAK_OK;
}

int plan_check_members(hobak_ctx* ctx, const std::vector<hobak_inst*>& m) {
  if (m.empty()) return fail(ctx, HOBAK_E_INVALID, "empty batch");
  for (auto* I : m)
    if (I->tsize != m[0]->tsize || I->P.shift != m[0]->P.shift || I->groups != m[0]->groups)
      return fail(ctx, HOBAK_E_INVALID, "instances of one batch must share precision and slice height");
  return HOBAK_OK;
}

void plan_free(Plan* P) {
  if (!P) return;
  for (auto& g : P->chunk_exec)
    if (g) (void)hipGraphExecDestroy(g);
  for (auto& g : P->pro_exec)
    if (g) (void)hipGraphExecDestroy(g);
  if (P->epi_exec) (void)hipGraphExecDestroy(P->epi_exec);
  if (P->owned) {
    if (P->dev) (void)hipFree(P->dev);
    if (P->host) (void)hipHostFree(P->host);
  }
  delete P;
}

// Transient plan in the context's arena (valid until the stream is next idle and the arena reset).
int plan_transient(hobak_ctx* ctx, const std::vector<hobak_inst*>& m, Plan& P) {
  TRY(plan_check_members(ctx, m));
  const size_t need = plan_bytes(m.size());
  if (ctx->arena_off + need > ctx->arena_cap) return fail(ctx, HOBAK_E_INVALID, "batch table arena exhausted");
  P = Plan{};
  P.members = m;
  P.n = (int)m.size();
  P.tsize = m[0]->tsize;
  P.groups = m[0]->groups;
  P.shift = m[0]->P.shift;
  P.refine = m[0]->h.refine;
  P.dev = ctx->arena_dev + ctx->arena_off;
  P.host = ctx->arena_host + ctx->arena_off;
  P.bytes = need;
  ctx->arena_off += need;
  return plan_fill(ctx, P);
}

// Grows the arena to hold `need` bytes; only called while the stream is idle.
int arena_ensure(hobak_ctx* ctx, size_t need) {
  if (need <= ctx->arena_cap) return HOBAK_OK;
  size_t cap = std::max<size_t>(need, 2 * ctx->arena_cap);
  cap = std::max<size_t>(cap, 1 << 20);
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  if (ctx->arena_dev) (void)hipFree(ctx->arena_dev);
  if (ctx->arena_host) (void)hipHostFree(ctx->arena_host);
  ctx->arena_dev = nullptr;
  ctx->arena_host = nullptr;
  ctx->arena_cap = 0;
  HIPCHK(ctx, hipMalloc((void**)&ctx->arena_dev, cap));
  HIPCHK(ctx, hipHostMalloc((void**)&ctx->arena_host, cap, hipHostMallocDefault));
  ctx->arena_cap = cap;
  ctx->arena_off = 0;
  return HOBAK_OK;
}

// Cached member plan (own memory, graphs), keyed by the exact member list.
int plan_get(hobak_ctx* ctx, const std::vector<hobak_inst*>& m, Plan** out) {
  for (auto* P : ctx->plans)
    if (P->members == m) {
      P->stamp = ++ctx->stamp;
      *out = P;
      return HOBAK_OK;
    }
  TRY(plan_check_members(ctx, m));
  if ((int)ctx->plans.size() >= kMaxPlans) {
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    auto it = std::min_element(ctx->plans.begin(), ctx->plans.end(),
                               [](Plan* a, Plan* b) { return a->stamp < b->stamp; });
    plan_free(*it);
    ctx->plans.erase(it);
  }
  Plan* P = new (std::nothrow) Plan;
  if (!P) return fail(ctx, HOBAK_E_NOMEM, "host allocation failed");
  P->members = m;
  P->n = (int)m.size();
  P->tsize = m[0]->tsize;
  P->groups = m[0]->groups;
  P->shift = m[0]->P.shift;
  P->refine = m[0]->h.refine;
  P->owned = true;
  P->bytes = plan_bytes(m.size());
  if (hipMalloc((void**)&P->dev, P->bytes) != hipSuccess ||
      hipHostMalloc((void**)&P->host, P->bytes, hipHostMallocDefault) != hipSuccess) {
    plan_free(P);
    return fail(ctx, HOBAK_E_NOMEM, "plan allocation failed");
  }
  int rc = plan_fill(ctx, *P);
  if (rc != HOBAK_OK) {
    plan_free(P);
    return rc;
  }
  P->stamp = ++ctx->stamp;
  ctx->plans.push_back(P);
  *out = P;
  return HOBAK_OK;
}

// Drops every cached plan that names I (its records are about to go).
void plans_forget(hobak_ctx* ctx, hobak_inst* I) {
  for (size_t k = 0; k < ctx->plans.size();) {
    Plan* P = ctx->plans[k];
    if (std::find(P->members.begin(), P->members.end(), I) != P->members.end()) {
      plan_free(P);
      ctx->plans.erase(ctx->plans.begin() + k);
    } else {
      k++;
    }
  }
}

// ---------------------------------------------------------------------------------------------
// the step engine
// ---------------------------------------------------------------------------------------------
// Records (hipStreamBeginCapture .. hipStreamEndCapture) whatever `body` enqueues into an
// executable graph.
template <typename F>
int capture(hobak_ctx* ctx, hipGraphExec_t* out, F&& body) {
  hipStream_t s = ctx->stream;
  hipGraph_t g = nullptr;
  HIPCHK(ctx, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  const int rc = body();
  const hipError_t e = hipStreamEndCapture(s, &g);
  if (rc != HOBAK_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return fail(ctx, HOBAK_E_HIP, std::string("graph capture failed: ") + hipGetErrorString(e));
  const hipError_t ei = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (ei != hipSuccess)
    return fail(ctx, HOBAK_E_HIP, std::string("graph instantiation failed: ") + hipGetErrorString(ei));
  return HOBAK_OK;
}

// Element kernel, assembly, right-hand side and the PCG (or refinement) seed for the instances of
// J; S (may be null) are those that start a new step.  With `timed`, events 1 and 2 mark the ends
// of the element kernel and of the assembly.
template <typename T>
int launch_prologue(hobak_ctx* ctx, const Plan& J, const Plan* S, bool timed) {
  hipStream_t s = ctx->stream;
  const int cap = ctx->grid_cap;
  if (S) HIPCHK(ctx, launch_begin_step(S->b[K_ROW], cap, s));
  if (S && S->contact) HIPCHK(ctx, launch_contact_begin(S->b[K_ROW], cap, s));
  HIPCHK(ctx, launch_element<T>(J.b[K_ELEM], ELEM_U, 1, cap, s));
  if (J.damping) HIPCHK(ctx, launch_damping_force<T>(J.b[K_ELEM], ELEM_U, cap, s));
  if (timed) HIPCHK(ctx, hipEventRecord(ctx->ev[1], s));
  HIPCHK(ctx, launch_assemble<T>(J.b[K_ASM], ASM_SYSTEM, cap, s));
  if (timed) HIPCHK(ctx, hipEventRecord(ctx->ev[2], s));
  HIPCHK(ctx, launch_rhs<T>(J.b[K_ROW], J.refine ? 0 : 1, cap, s));
  if (J.refine) {
    HIPCHK(ctx, launch_refine_begin(J.b[K_ROW], cap, s));
    if (J.agg) HIPCHK(ctx, launch_agg_build<T>(J.b[K_ROW], J.groups, cap, s));
    HIPCHK(ctx, launch_refine_init<T>(J.b[K_ROW], 0, J.groups, cap, s));
    HIPCHK(ctx, launch_cg_init(J.b[K_ONE], 1, 0, s));
  } else {
    HIPCHK(ctx, launch_cg_init(J.b[K_ONE], 0, 0, s));
  }
  return HOBAK_OK;
}

// Whether a group's SpMV launches read the matrix with non-temporal loads: when the matrices of
// the instances still solving exceed the threshold (default 1.5x the 256 MB Infinity Cache), few
// of their lines survive to the next iteration, and streaming them keeps the gathered vectors
// cached instead.  Below it much of the matrix is re-read from the cache, which non-temporal
// loads would give up.
bool stream_matrix(const hobak_ctx* ctx, double solving_bytes) {
  return solving_bytes > (double)ctx->spmv_stream_mb * (1 << 20);
}

// k PCG iterations over the batch (k even, so the current scalar set is set 0 again at the end),
// then the gated end-of-solve kernels: the refinement residual, check and re-seed, or the plain
// form's finish flag.  With a sampling scheduler, some SpMV launches carry timing events.
template <typename T>
int launch_chunk(hobak_ctx* ctx, const Plan& P, int k, int stream, hobak_sched* S, int every) {
  hipStream_t s = ctx->stream;
  const int cap = ctx->grid_cap;
  // FP64 groups whose matrices stream from HBM form p_new once per entry and gather it alone
  // (three kernels per iteration, bitwise the same results); the extra launch does not pay for
  // a working set that stays in the caches
  const bool pform = sizeof(T) == 8 && (ctx->p_form == 2 || (ctx->p_form == 1 && stream));
  for (int j = 0; j < k; j++) {
    hipEvent_t a = nullptr, b = nullptr;
    if (S) {
      const int g = S->sample_it++;
      if (every > 0 && g % every == every - 1 && (int)S->samples.size() < kMaxSamples) {
        const int slot = (int)S->samples.size();
        a = ctx->sample_ev[2 * slot];
        b = ctx->sample_ev[2 * slot + 1];
        S->samples.emplace_back(g, slot);
      }
    }
    if (pform) {
      HIPCHK(ctx, launch_cg_p(P.b[K_UPD], j & 1, cap, s));
      HIPCHK(ctx, launch_cg_spmv_p(P.b[K_SPMV], P.shift, ctx->spmv_pad, stream, j & 1, cap, s, a, b));
    } else {
      HIPCHK(ctx, launch_cg_spmv<T>(P.b[K_SPMV], P.shift, ctx->spmv_pad, stream, P.coarse, j & 1, cap, s, a, b));
    }
    HIPCHK(ctx, launch_cg_update<T>(P.b[K_UPD], P.groups, P.coarse | (P.agg << 1), j & 1, cap, s));
  }
  if (P.refine) {
    HIPCHK(ctx, launch_refine_residual<T>(P.b[K_SPMV], P.b[K_DOF], P.shift, cap, s));
    HIPCHK(ctx, launch_finish_check(P.b[K_ONE], 1, s));
    HIPCHK(ctx, launch_refine_init<T>(P.b[K_ROW], 1, P.groups, cap, s));
    HIPCHK(ctx, launch_cg_init(P.b[K_ONE], 1, 1, s));
  } else {
    HIPCHK(ctx, launch_finish_check(P.b[K_ONE], 0, s));
  }
  return HOBAK_OK;
}

template <typename T>
int launch_epilogue(hobak_ctx* ctx, const Plan& F, const Plan* C) {
  hipStream_t s = ctx->stream;
  HIPCHK(ctx, launch_apply_dx<T>(F.b[K_DOF], F.refine, ctx->grid_cap, s));
  if (C && C->contact) HIPCHK(ctx, launch_contact_sep(C->b[K_ROW], ctx->grid_cap, s));
  if (C) HIPCHK(ctx, launch_end_step(C->b[K_ROW], ctx->grid_cap, s));
  return HOBAK_OK;
}

double clock_s(const hobak_sched* S) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - S->t0).count();
}

float elapsed(hobak_ctx* ctx, int a, int b) {
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, ctx->ev[a], ctx->ev[b]) != hipSuccess) return 0.f;
  return ms;
}

int worse(int a, int b) {
  if (a == HOBAK_E_NONFINITE || b == HOBAK_E_NONFINITE) return HOBAK_E_NONFINITE;
  if (a != HOBAK_OK) return a;
  return b;
}

// One tick of the engine (see the file comment).  Requires an idle stream on entry; leaves it idle.
template <typename T>
int tick_impl(hobak_sched* S) {
  hobak_ctx* ctx = S->ctx;
  hipStream_t s = ctx->stream;
  TRY(arena_ensure(ctx, 5 * plan_bytes(S->members.size() + 1)));
  ctx->arena_off = 0;
  TRY(plan_get(ctx, S->members, &S->mplan));
  if (S->dirty) {
    S->index.clear();
    for (size_t i = 0; i < S->members.size(); i++) S->index[S->members[i]] = (int)i;
    S->dirty = false;
  }
  Plan& M = *S->mplan;
  const bool graphs = ctx->use_graphs != 0;

  // 1. solves that finished at the last poll: x is applied, then the step ends or the next BDF_1
  //    Newton iteration re-enters (the test uses |S b|^2 of the system just solved, BDF_1.cpp:111, 172)
  std::vector<hobak_inst*> fin, comp;
  for (auto& r : S->active) {
    if (r.stage != 2) continue;
    fin.push_back(r.I);
    const hobak_step_params& p = r.I->prm;
    const bool more = p.newton_iters > 1 && r.newton_k + 1 < p.newton_iters &&
                      r.st.newton_residual[r.newton_k] > p.newton_tol;
    if (more) {
      r.newton_k++;
      r.stage = 0;
      r.begin = 0;
      r.done_iters = 0;
    } else {
      comp.push_back(r.I);
      r.stage = 3;  // complete once this tick's work has run
    }
  }
  const bool epi = !fin.empty();
  if (epi) {
    const Range r_epi("hobak epilogue");
    HIPCHK(ctx, hipEventRecord(ctx->ev[6], s));
    if (graphs && (int)fin.size() == M.n && comp.size() == fin.size()) {
      if (!M.epi_exec) TRY(capture(ctx, &M.epi_exec, [&]() { return launch_epilogue<T>(ctx, M, &M); }));
      HIPCHK(ctx, hipGraphLaunch(M.epi_exec, s));
    } else {
      Plan PF, PC;
      TRY(plan_transient(ctx, fin, PF));
      const Plan* pc = nullptr;
      if (comp.size() == fin.size()) {
        pc = &PF;
      } else if (!comp.empty()) {
        TRY(plan_transient(ctx, comp, PC));
        pc = &PC;
      }
      TRY(launch_epilogue<T>(ctx, PF, pc));
    }
    HIPCHK(ctx, hipEventRecord(ctx->ev[7], s));
  }

  // 2. queued requests whose instance is free start
  for (auto it = S->queue.begin(); it != S->queue.end();) {
    bool busy = false;
    for (auto& r : S->active)
      if (r.I == it->I) busy = true;
    if (busy) {
      ++it;
      continue;
    }
    S->active.push_back(*it);
    S->active.back().stage = 0;
    S->active.back().begin = 1;
    it = S->queue.erase(it);
  }

  // 3. prologue of every run that needs one
  std::vector<hobak_inst*> join, starters;
  for (auto& r : S->active)
    if (r.stage == 0) {
      join.push_back(r.I);
      if (r.begin) starters.push_back(r.I);
    }
  const bool pro = !join.empty();
  bool pro_split = false;
  if (pro) {
    const Range r_pro("hobak prologue");
    HIPCHK(ctx, hipEventRecord(ctx->ev[0], s));
    const bool all = (int)join.size() == M.n && (starters.empty() || starters.size() == join.size());
    if (graphs && all) {
      const int wb = starters.empty() ? 0 : 1;
      if (!M.pro_exec[wb])
        TRY(capture(ctx, &M.pro_exec[wb], [&]() { return launch_prologue<T>(ctx, M, wb ? &M : nullptr, false); }));
      HIPCHK(ctx, hipGraphLaunch(M.pro_exec[wb], s));
    } else {
      Plan PJ, PS;
      TRY(plan_transient(ctx, join, PJ));
      const Plan* ps = nullptr;
      if (starters.size() == join.size()) {
        ps = &PJ;
      } else if (!starters.empty()) {
        TRY(plan_transient(ctx, starters, PS));
        ps = &PS;
      }
      TRY(launch_prologue<T>(ctx, PJ, ps, true));
      pro_split = true;
    }
    HIPCHK(ctx, hipEventRecord(ctx->ev[3], s));
    if (S->sampling && all && !starters.empty()) {
      S->sample_it = 0;
      S->samples.clear();
    }
    const double now = clock_s(S);
    for (auto& r : S->active)
      if (r.stage == 0) {
        r.stage = 1;
        r.done_iters = 0;
        if (r.t_start < 0) r.t_start = now;
      }
  }

  // 4. groups of PCG iterations: as many as the solving instance closest to finishing is
  //    predicted to need (from its last step), at least one.  They run over the member set (one
  //    captured graph) while the solving instances hold more than compact_pct percent of its SpMV
  //    workgroups; below that, over a table of the solving instances only, launched directly, so
  //    that the workgroups of finished instances are not dispatched just to return.
  bool ran = false;
  int m = ctx->max_chunks;
  int nsolving = 0;
  double solving_bytes = 0.0;
  long solving_wgs = 0;
  std::vector<hobak_inst*> solving;
  const int k = ctx->chunk_iters;
  for (auto& r : S->active) {
    if (r.stage != 1) continue;
    nsolving++;
    solving_bytes += spmv_bytes_of(r.I);
    solving_wgs += r.I->h.nb[K_SPMV];
    solving.push_back(r.I);
    const int p = r.I->pred[r.newton_k];
    const int q = p > 0 ? (p - r.done_iters) / k : 0;
    m = std::min(m, std::max(1, q));
  }
  if (nsolving) {
    const Range r_pcg("hobak pcg groups");
    HIPCHK(ctx, hipEventRecord(ctx->ev[4], s));
    int every = 0;
    if (S->sampling && S->ephemeral) {
      for (auto* I : S->members)
        if (I->sample_every > 0) every = every ? std::min(every, I->sample_every) : I->sample_every;
    } else if (S->sampling) {
      every = S->spmv_every;
    }
    const int st = stream_matrix(ctx, solving_bytes) ? 1 : 0;
    const bool compact = nsolving < M.n && solving_wgs * 100 <= (long)ctx->compact_pct * M.b[K_SPMV].total;
    Plan PG;
    if (compact) TRY(plan_transient(ctx, solving, PG));
    if (graphs && !compact && (!M.chunk_exec[st] || M.chunk_k[st] != k)) {
      if (M.chunk_exec[st]) (void)hipGraphExecDestroy(M.chunk_exec[st]);
      M.chunk_exec[st] = nullptr;
      TRY(capture(ctx, &M.chunk_exec[st], [&]() { return launch_chunk<T>(ctx, M, k, st, nullptr, 0); }));
      M.chunk_k[st] = k;
    }
    for (int j = 0; j < m; j++) {
      // a group that holds a sampled SpMV launch (index g with g % every == every - 1) runs
      // uncaptured, with the events on that launch; every other group replays the graph
      const int g0 = S->sample_it;
      const bool sampled = every > 0 && (int)S->samples.size() < kMaxSamples &&
                           (g0 / every) * every + every - 1 < g0 + k;
      if (graphs && !compact && !sampled) {
        HIPCHK(ctx, hipGraphLaunch(M.chunk_exec[st], s));
        if (S->sampling) S->sample_it += k;
        S->stat[9] += 1;
      } else {
        TRY(launch_chunk<T>(ctx, compact ? PG : M, k, st, S->sampling ? S : nullptr, every));
        S->stat[10] += 1;
      }
    }
    HIPCHK(ctx, hipEventRecord(ctx->ev[5], s));
    ran = true;
  }

  S->stat[1] += nsolving ? m : 0;
  if (nsolving) S->stat[7] += (double)m * k;
  // 5. statuses
  {
    const Range r_poll("hobak poll");
    HIPCHK(ctx, launch_status(M.b[K_ONE], M.status_dev, s));
    HIPCHK(ctx, hipMemcpyAsync(M.status_host, M.status_dev, sizeof(InstStatus) * M.n, hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipStreamSynchronize(s));
  }
  // algorithmic SpMV bytes of this tick: each member's matrix and vectors once per launch that
  // streamed them (counted on the device; launches after a member's solve ended read nothing)
  double work_bytes = 0.0;
  for (int i = 0; i < M.n; i++) {
    hobak_inst* I = M.members[i];
    const unsigned w = M.status_host[i].spmv_work;
    work_bytes += (double)(unsigned)(w - I->work_seen) * spmv_bytes_of(I);
    I->work_seen = w;
  }
  S->stat[8] += work_bytes;
  const double now = clock_s(S);
  if (pro) {
    if (pro_split) {
      S->ms[0] += elapsed(ctx, 0, 1);
      S->ms[1] += elapsed(ctx, 1, 2);
      S->ms[2] += elapsed(ctx, 2, 3);
    } else {
      S->ms[0] += elapsed(ctx, 0, 3);
    }
  }
  if (ran) S->ms[3] += elapsed(ctx, 4, 5);
  if (epi) S->ms[4] += elapsed(ctx, 6, 7);
  if (!S->ephemeral && S->sampling) {
    // the window's algorithmic bytes and launches; the sampled launches (every `every`-th one)
    // give the average launch time, so bytes / (launches * average) is the window's rate
    S->spmv_bytes += work_bytes;
    if (ran) S->spmv_launches += (long long)m * k;
  }
  if (!S->ephemeral && S->sampling && !S->samples.empty()) {
    for (auto& sm : S->samples) {
      float ms = 0.f;
      HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->sample_ev[2 * sm.second], ctx->sample_ev[2 * sm.second + 1]));
      S->spmv_ms += ms;
      S->spmv_n += 1;
    }
    S->samples.clear();
  }
  S->stat[0] += 1;
  if (pro) S->stat[2] += elapsed(ctx, 0, 3);
  if (ran) S->stat[3] += elapsed(ctx, 4, 5);
  if (epi) S->stat[4] += elapsed(ctx, 6, 7);

  // completions: their epilogue ran before the wait above
  for (size_t i = 0; i < S->active.size();) {
    auto& r = S->active[i];
    if (r.stage != 3) {
      i++;
      continue;
    }
    hobak_completion c;
    std::memset(&c, 0, sizeof(c));
    c.inst = r.I;
    c.tag = r.tag;
    c.stats = r.st;
    c.stats.newton_iters = r.newton_k + 1;
    c.stats.seconds = now - r.t_submit;
    c.submit_s = r.t_submit;
    c.start_s = r.t_start;
    c.end_s = now;
    r.I->requests--;
    r.I->has_system = true;
    S->done.push_back(c);
    S->active.erase(S->active.begin() + i);
  }
  // solving instances: finished solves are recorded; the epilogue follows next tick
  for (auto& r : S->active) {
    if (r.stage != 1) continue;
    const InstStatus& q = M.status_host[S->index[r.I]];
    const bool refine = r.I->h.refine != 0;
    if (q.phase == PHASE_FINISHED) {
      const int nk = r.newton_k;
      const int iters = refine ? q.inner_total : q.iters;
      r.st.pcg_iters[nk] = iters;
      r.st.refine_iters[nk] = refine ? q.sweeps : 0;
      r.st.pcg_error[nk] = q.err;
      r.st.newton_residual[nk] = q.maxR;
      r.st.total_pcg_iters += iters;
      int code = HOBAK_OK;
      if (refine) {
        if (q.ref_status == 2) code = HOBAK_E_NONFINITE;
        else if (q.ref_status == 1 || !q.ref_done) code = HOBAK_E_NOT_CONVERGED;
      } else {
        if (q.cg_status == 2) code = HOBAK_E_NONFINITE;
        else if (q.cg_status == 1) code = HOBAK_E_NOT_CONVERGED;
      }
      r.st.status = worse(r.st.status, code);
      r.I->pred[nk] = iters;
      if (nk == 0) r.I->first_iters = q.first_iters;
      r.stage = 2;
      S->stat[6] += 1;
    } else {
      r.done_iters = refine ? q.inner_total + q.iters : q.iters;
    }
  }
  return HOBAK_OK;
}

int tick(hobak_sched* S) {
  if (S->members.empty()) return HOBAK_OK;
  return S->members[0]->tsize == 8 ? tick_impl<double>(S) : tick_impl<float>(S);
}

// Drops a scheduler's queued and running requests of I (after the stream is idle).
void sched_drop(hobak_sched* S, hobak_inst* I) {
  for (auto it = S->queue.begin(); it != S->queue.end();)
    if (it->I == I) {
      I->requests--;
      it = S->queue.erase(it);
    } else {
      ++it;
    }
  for (auto it = S->active.begin(); it != S->active.end();)
    if (it->I == I) {
      I->requests--;
      it = S->active.erase(it);
    } else {
      ++it;
    }
}

void sched_detach(hobak_sched* S, hobak_inst* I) {
  sched_drop(S, I);
  auto it = std::find(S->members.begin(), S->members.end(), I);
  if (it != S->members.end()) {
    S->members.erase(it);
    S->dirty = true;
  }
  if (I->sched == S) I->sched = nullptr;
}

// Advances every instance of `m` by one step through a private scheduler; stats[i] is m[i]'s.
int run_batch(hobak_ctx* ctx, const std::vector<hobak_inst*>& m, hobak_step_stats* stats) {
  hobak_sched S;
  S.ctx = ctx;
  S.ephemeral = true;
  S.members = m;
  S.sampling = true;
  for (auto* I : m) S.sampling = S.sampling && I->sample_every > 0;
  if (S.sampling && ctx->sample_ev.empty()) {
    ctx->sample_ev.assign(2 * kMaxSamples, nullptr);
    for (auto& e : ctx->sample_ev) HIPCHK(ctx, hipEventCreate(&e));
  }
  for (size_t i = 0; i < m.size(); i++) {
    hobak_sched::Run r;
    r.I = m[i];
    r.tag = i;
    r.t_submit = clock_s(&S);
    S.queue.push_back(r);
    m[i]->requests++;
  }
  int rc = HOBAK_OK;
  long guard = 0;
  while ((!S.queue.empty() || !S.active.empty()) && rc == HOBAK_OK) {
    rc = tick(&S);
    if (++guard > 100000000L) rc = fail(ctx, HOBAK_E_HIP, "step engine did not terminate");
  }
  if (rc != HOBAK_OK) {
    (void)hipStreamSynchronize(ctx->stream);
    for (auto* I : m) sched_drop(&S, I);
    return rc;
  }
  int worst = HOBAK_OK;
  for (auto& c : S.done) {
    if (stats) stats[c.tag] = c.stats;
    worst = worse(worst, c.stats.status);
  }
  // sampled SpMV launches did work for every member while their index was <= every member's
  // first-solve iteration count (launch 0 forms p = z; launch j finishes iteration j - 1); with
  // refinement, that of the first inner solve
  if (S.sampling && !S.samples.empty()) {
    int lim = INT32_MAX;
    for (auto* I : m) lim = std::min(lim, I->first_iters);
    for (auto& sm : S.samples) {
      if (sm.first > lim) continue;
      float ms = 0.f;
      HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->sample_ev[2 * sm.second], ctx->sample_ev[2 * sm.second + 1]));
      for (auto* I : m) {
        I->spmv_ms_sum += ms;
        I->spmv_count += 1;
        I->spmv_bytes_sum += S.mplan->spmv_bytes;
      }
    }
  }
  for (auto* I : m) std::memcpy(I->last_ms, S.ms, sizeof(S.ms));
  return worst;
}

template <typename T>
int setup_constant_parts(hobak_inst* I) {
  hobak_ctx* ctx = I->ctx;
  hipStream_t s = ctx->stream;
  ctx->arena_off = 0;
  Plan P;
  TRY(plan_transient(ctx, {I}, P));
  // rest-state stiffness (u = 0): TIMESTEPPER::buildRayleighDampingMatrix (TIMESTEPPER.cpp:213-236)
  HIPCHK(ctx, hipMemsetAsync(I->h.tmp64, 0, sizeof(double) * I->dofs, s));
  HIPCHK(ctx, launch_element<T>(P.b[K_ELEM], ELEM_TMP, 1, ctx->grid_cap, s));
  HIPCHK(ctx, launch_assemble<T>(P.b[K_ASM], ASM_REST, ctx->grid_cap, s));
  HIPCHK(ctx, launch_constant_parts<T>(P.b[K_SLOT], ctx->grid_cap, s));
  HIPCHK(ctx, hipStreamSynchronize(s));
  return HOBAK_OK;
}

void free_inst(hobak_inst* I) {
  for (auto& a : I->allocs) {
    (void)hipFree(a.first);
    if (I->ctx) I->ctx->in_use -= a.second;
  }
  I->allocs.clear();
}

template <typename T>
int download_sell_as_bsr(hobak_inst* I, const void* dev, double* out, double sign) {
  const size_t ns = I->P.nslots;
  std::vector<T> h(9 * ns);
  HIPCHK(I->ctx, hipMemcpyAsync(h.data(), dev, sizeof(T) * 9 * ns, hipMemcpyDeviceToHost, I->ctx->stream));
  HIPCHK(I->ctx, hipStreamSynchronize(I->ctx->stream));
  for (int k = 0; k < I->P.nnzb; k++) {
    const int slot = I->oslot[k];
    for (int e = 0; e < 9; e++) out[9 * (size_t)k + e] = sign * (double)h[e * ns + slot];
  }
  return HOBAK_OK;
}

template <typename T>
int eval_forces_hessian(hobak_inst* I, const double* u, double* forces, double* kvals) {
  hobak_ctx* ctx = I->ctx;
  hipStream_t s = ctx->stream;
  ctx->arena_off = 0;
  Plan P;
  TRY(plan_transient(ctx, {I}, P));
  std::vector<double> ubuf;
  TRY(upload(I, I->h.tmp64, to_device_order(I, u, ubuf), (size_t)I->dofs));
  HIPCHK(ctx, launch_element<T>(P.b[K_ELEM], ELEM_TMP, kvals != nullptr, ctx->grid_cap, s));
  if (forces) {
    HIPCHK(ctx, launch_force_gather<T>(P.b[K_ROW], ctx->grid_cap, s));
    HIPCHK(ctx, hipMemcpyAsync(forces, I->h.rhs64, sizeof(double) * I->dofs, hipMemcpyDeviceToHost, s));
  }
  I->has_system = false;  // rhs64 and A are reused
  if (kvals) {
    HIPCHK(ctx, launch_assemble<T>(P.b[K_ASM], ASM_PLAIN_A, ctx->grid_cap, s));
    TRY(download_sell_as_bsr<T>(I, I->h.A, kvals, -1.0));
  }
  HIPCHK(ctx, hipStreamSynchronize(s));
  to_caller_order(I, forces);
  return HOBAK_OK;
}

// computeInternalForces(hyperelastic, damping) and TET_MESH_FASTER::computeDampingHessian at u
// with the held velocity gradients.
template <typename T>
int eval_damping(hobak_inst* I, const double* u, double* forces, double* cvals) {
  hobak_ctx* ctx = I->ctx;
  hipStream_t s = ctx->stream;
  ctx->arena_off = 0;
  Plan P;
  TRY(plan_transient(ctx, {I}, P));
  std::vector<double> ubuf;
  TRY(upload(I, I->h.tmp64, to_device_order(I, u, ubuf), (size_t)I->dofs));
  I->has_system = false;  // rhs64 and A are reused
  if (forces) {
    HIPCHK(ctx, launch_element<T>(P.b[K_ELEM], ELEM_TMP, 0, ctx->grid_cap, s));
    HIPCHK(ctx, launch_damping_force<T>(P.b[K_ELEM], ELEM_TMP, ctx->grid_cap, s));
    HIPCHK(ctx, launch_force_gather<T>(P.b[K_ROW], ctx->grid_cap, s));
    HIPCHK(ctx, hipMemcpyAsync(forces, I->h.rhs64, sizeof(double) * I->dofs, hipMemcpyDeviceToHost, s));
  }
  if (cvals) {
    HIPCHK(ctx, launch_damping_hessian<T>(P.b[K_ELEM], ELEM_TMP, ctx->grid_cap, s));
    HIPCHK(ctx, launch_assemble<T>(P.b[K_ASM], ASM_PLAIN_A, ctx->grid_cap, s));
    TRY(download_sell_as_bsr<T>(I, I->h.A, cvals, -1.0));
  }
  HIPCHK(ctx, hipStreamSynchronize(s));
  to_caller_order(I, forces);
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
  bool ok = true;
  for (auto& e : ctx->ev) ok = ok && hipEventCreate(&e) == hipSuccess;
  if (!ok || arena_ensure(ctx, 1 << 20) != HOBAK_OK || scratch_of_device(device, &ctx->scratch) != HOBAK_OK) {
    hobak_ctx_destroy(ctx);
    return HOBAK_E_HIP;
  }
  *out = ctx;
  return HOBAK_OK;
}

void hobak_ctx_destroy(hobak_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  std::vector<hobak_sched*> ss(ctx->scheds.begin(), ctx->scheds.end());
  for (auto* S : ss) hobak_sched_destroy(S);
  std::vector<hobak_inst*> all(ctx->insts.begin(), ctx->insts.end());
  for (auto* I : all) hobak_inst_destroy(I);
  for (auto* P : ctx->plans) plan_free(P);
  ctx->plans.clear();
  if (ctx->arena_dev) (void)hipFree(ctx->arena_dev);
  if (ctx->arena_host) (void)hipHostFree(ctx->arena_host);
  for (auto& e : ctx->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : ctx->sample_ev)
    if (e) (void)hipEventDestroy(e);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

const char* hobak_last_error(hobak_ctx* ctx) { return ctx ? ctx->err.c_str() : "no context"; }

size_t hobak_ctx_bytes_in_use(hobak_ctx* ctx) { return ctx ? ctx->in_use : 0; }

int hobak_check_offsets(int64_t nv, int64_t nslots, int32_t precision) {
  if (nv <= 0 || nslots <= 0) return HOBAK_E_INVALID;
  const int64_t ts = precision == HOBAK_PRECISION_FP64 ? 8 : 4;
  const int64_t rows_pad = (nv + 63) / 64 * 64;  // vectors are padded to whole slices (<= 64 rows)
  const int64_t lim = (int64_t)1 << 32;
  if (nslots * ts >= lim || nslots * 4 >= lim || 3 * rows_pad * ts >= lim) return HOBAK_E_INVALID;
  return HOBAK_OK;
}

size_t hobak_inst_bytes_estimate(int32_t nv, int32_t nt, int32_t precision) {
  // a Kuhn lattice has ~15 blocks per row; padding of the slices adds a few percent
  const size_t nnzb = (size_t)nv * 16;
  return estimate_bytes(nv, nt, (int)(nnzb * 1.15), precision);
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
  if (params->integrator != HOBAK_INTEGRATOR_POSITION && params->integrator != HOBAK_INTEGRATOR_VELOCITY)
    return fail(ctx, HOBAK_E_INVALID, "unknown integrator");
  if (params->integrator == HOBAK_INTEGRATOR_VELOCITY &&
      (params->precision != HOBAK_PRECISION_FP64 || params->newton_iters > 1 || params->ground_on))
    return fail(ctx, HOBAK_E_INVALID,
                "the velocity integrator runs in the FP64 form, one Newton iteration, no penalty ground");
  for (size_t i = 0; i < 4 * (size_t)nt; i++)
    if (tets[i] < 0 || tets[i] >= nv) return fail(ctx, HOBAK_E_INVALID, "tet index out of range");
  for (size_t t = 0; t < (size_t)nt; t++) {
    const int32_t* T = tets + 4 * t;
    if (T[0] == T[1] || T[0] == T[2] || T[0] == T[3] || T[1] == T[2] || T[1] == T[3] || T[2] == T[3])
      return fail(ctx, HOBAK_E_INVALID, "a tet repeats a vertex");
  }
  // row_sort: the device holds the mesh with its vertices relabelled (row_sort_perm); the API
  // keeps the caller's numbering
  std::vector<int> perm;
  std::vector<double> rest_r;
  std::vector<int32_t> tets_r;
  const bool spatial = ctx->agg && params->precision == HOBAK_PRECISION_FP32_REFINE;
  if (spatial || ctx->row_sort) {
    perm = spatial ? spatial_perm(nv, nt, rest_xyz, tets, 21 * ctx->update_groups, ctx->row_sort != 0)
                   : row_sort_perm(nv, nt, tets);
    std::vector<int> ip(nv);
    for (int p = 0; p < nv; p++) ip[perm[p]] = p;
    rest_r.resize(3 * (size_t)nv);
    for (int p = 0; p < nv; p++)
      for (int c = 0; c < 3; c++) rest_r[3 * (size_t)p + c] = rest_xyz[3 * (size_t)perm[p] + c];
    tets_r.resize(4 * (size_t)nt);
    for (size_t i = 0; i < 4 * (size_t)nt; i++) tets_r[i] = ip[tets[i]];
    rest_xyz = rest_r.data();
    tets = tets_r.data();
  }
  // rest volumes must be positive (TET_MESH.cpp:132-137)
  std::vector<double> vol_h(nt), dminv_h(9 * (size_t)nt), one_ring_h(nv, 0.0);
  for (int t = 0; t < nt; t++) {
    const int32_t* T = tets + 4 * (size_t)t;
    double d[3][3];
    for (int k = 0; k < 3; k++)
      for (int c = 0; c < 3; c++) d[k][c] = rest_xyz[3 * (size_t)T[k + 1] + c] - rest_xyz[3 * (size_t)T[0] + c];
    const double cx = d[0][1] * d[1][2] - d[0][2] * d[1][1];
    const double cy = d[0][2] * d[1][0] - d[0][0] * d[1][2];
    const double cz = d[0][0] * d[1][1] - d[0][1] * d[1][0];
    vol_h[t] = (d[2][0] * cx + d[2][1] * cy + d[2][2] * cz) / 6.0;
    if (!(vol_h[t] > 0.0)) return fail(ctx, HOBAK_E_INVALID, "tet with non-positive rest volume");
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
      for (int c = 0; c < 3; c++) dminv_h[(3 * r + c) * (size_t)nt + t] = inv[r][c];
  }
  for (int t = 0; t < nt; t++) {  // one-ring volumes, ascending tet (TET_MESH.cpp:156-173)
    const double q = 0.25 * vol_h[t];
    for (int k = 0; k < 4; k++) one_ring_h[tets[4 * (size_t)t + k]] += q;
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
  for (auto& p : I->pred) p = -1;
  if (material->kind == HOBAK_MAT_SNH) {  // SNH.cpp:23-32 reparameterisation
    I->lam_eff = material->lambda + material->mu;
    I->alpha = 1.0 + material->mu / I->lam_eff;
  } else {
    I->lam_eff = material->lambda;
    I->alpha = 0.0;
  }
  I->tsize = params->precision == HOBAK_PRECISION_FP64 ? 8 : 4;
  I->groups = ctx->update_groups;
  I->spatial = spatial;
  build_pattern(nv, nt, tets, ctx->slice_rows, I->P, perm.empty() ? nullptr : &perm);
  I->perm = std::move(perm);
  if (!I->perm.empty()) {
    I->iperm.resize(nv);
    for (int p = 0; p < nv; p++) I->iperm[I->perm[p]] = p;
  }
  if (hobak_check_offsets(nv, I->P.nslots, params->precision) != HOBAK_OK) {
    delete I;
    return fail(ctx, HOBAK_E_INVALID, "mesh too large for the kernels' 32-bit byte offsets");
  }
  caller_pattern(I);
  const Pattern& P = I->P;
  const size_t need = estimate_bytes(nv, nt, P.nslots, params->precision);
  if (ctx->budget && ctx->in_use + need > ctx->budget) {
    delete I;
    return fail(ctx, HOBAK_E_NOMEM, "HBM budget would be exceeded");
  }
  auto bail = [&](int rc) {
    free_inst(I);
    delete I;
    return rc;
  };
  InstDev& h = I->h;
  h.nv = nv;
  h.nt = nt;
  h.nslots = P.nslots;
  h.shift = P.shift;
  for (int k = 0; k < K_COUNT; k++) h.nb[k] = blocks_of(k, nv, nt, P.nslots, P.shift, ctx->update_cap, I->groups);
  h.nb[K_ASM] = blocks_of(K_ASM, nv, nt, (int)P.aslot.size(), P.shift, ctx->update_cap, I->groups);
  h.nasm = (int)P.aslot.size();
  h.material = material->kind;
  h.precond = params->precond;
  h.bdf = I->prm.newton_iters > 1;
  h.ground_on = params->ground_on;
  h.refine = params->precision == HOBAK_PRECISION_FP32_REFINE;
  h.max_iters = params->pcg_max_iters > 0 ? params->pcg_max_iters : 2 * I->dofs;
  h.mu = material->mu;
  h.lam = I->lam_eff;
  h.alpha = I->alpha;
  h.invDt = 1.0 / params->dt;
  h.invDt2 = h.invDt * h.invDt;
  h.ra = params->rayleigh_alpha;
  h.rb = params->rayleigh_beta;
  for (int c = 0; c < 3; c++) h.g[c] = params->gravity[c];
  h.ground_y = params->ground_y;
  h.ground_k = params->ground_k;
  h.velform = params->integrator == HOBAK_INTEGRATOR_VELOCITY;
  h.dt = params->dt;
  h.kscale = h.velform ? params->dt * params->dt : 1.0;
  h.tol = params->pcg_tol > 0 ? params->pcg_tol : kEps;
  h.tau0 = params->refine_inner_tol > 0 ? params->refine_inner_tol : 1e-4;
  const size_t ts = I->tsize;
  const size_t np = P.nrows_pad;
  int rc = HOBAK_OK;
#define ALLOC(ptr, n) \
  if ((rc = dalloc(I, &ptr, (n))) != HOBAK_OK) return bail(rc)
#define ALLOCT(ptr, n) \
  if ((rc = dalloc(I, (char**)&ptr, (n), ts)) != HOBAK_OK) return bail(rc)
  double *rest, *mass, *one_ring, *pin_target, *dminv, *vol;
  int4* t4d;
  int *slice_off, *row_len, *col, *slot_row, *cptr, *cidx, *vptr, *fidx, *mirror, *aslot, *evptr, *evert;
  unsigned short* eloc;
  unsigned char* pinned;
  ALLOC(rest, 3 * (size_t)nv);
  ALLOC(h.u, 3 * (size_t)nv);
  ALLOC(h.v, 3 * (size_t)nv);
  ALLOC(h.u_old, 3 * (size_t)nv);
  ALLOC(h.v_old, 3 * (size_t)nv);
  ALLOC(mass, 3 * (size_t)nv);
  ALLOC(one_ring, (size_t)nv);
  ALLOC(pin_target, 3 * (size_t)nv);
  ALLOC(h.rhs64, 3 * (size_t)nv);
  ALLOC(h.tmp64, 3 * (size_t)nv);
  ALLOC(h.x64, 3 * (size_t)nv);
  ALLOC(h.res64, 3 * (size_t)nv);
  ALLOC(dminv, 9 * (size_t)nt);
  ALLOC(vol, (size_t)nt);
  ALLOC(t4d, (size_t)nt);
  ALLOC(slice_off, P.slice_off.size());
  ALLOC(row_len, (size_t)nv);
  ALLOC(col, (size_t)P.nslots);
  ALLOC(slot_row, (size_t)P.nslots);
  ALLOC(cptr, P.cptr.size());
  ALLOC(cidx, P.cidx.size());
  ALLOC(mirror, P.mirror.size());
  ALLOC(aslot, P.aslot.size());
  ALLOC(evptr, P.evptr.size());
  ALLOC(evert, P.evert.size());
  ALLOC(eloc, P.eloc.size());
  ALLOC(vptr, P.vptr.size());
  ALLOC(fidx, P.fidx.size());
  ALLOC(pinned, (size_t)nv);
  ALLOCT(h.kneg, 9 * (size_t)P.nslots);
  ALLOCT(h.stage_h, (size_t)9 * kStagedBlocks * nt);  // tet-major: 9 (pair nt + t) + entry
  ALLOCT(h.stage_f, (size_t)3 * 4 * nt);              // 3 (corner nt + t) + component
  ALLOCT(h.A, 9 * (size_t)P.nslots);
  ALLOCT(h.D, 9 * (size_t)P.nslots);
  ALLOCT(h.E, 9 * (size_t)P.nslots);
  ALLOCT(h.C, 9 * (size_t)P.nslots);
  ALLOCT(h.minv, 9 * np);
  ALLOCT(h.r, 3 * np);
  ALLOCT(h.x, 3 * np);
  ALLOCT(h.z, 3 * np);
  ALLOCT(h.p0, 3 * np);
  ALLOCT(h.p1, 3 * np);
  ALLOCT(h.t, 3 * np);
  if (h.refine) ALLOC(h.dg, 3 * np);
  ALLOC(h.part_row_a, (size_t)h.nb[K_ROW]);
  ALLOC(h.part_row_b, (size_t)h.nb[K_ROW]);
  ALLOC(h.part_pt, (size_t)h.nb[K_SPMV]);
  ALLOC(h.part_res, (size_t)h.nb[K_SPMV]);
  ALLOC(h.part_rr, (size_t)h.nb[K_UPD]);
  ALLOC(h.part_rz, (size_t)h.nb[K_UPD]);
  if (h.refine) {
    ALLOC(h.part_row_c, 6 * (size_t)h.nb[K_ROW]);
    ALLOC(h.part_upd_g, 3 * (size_t)h.nb[K_UPD]);
  }
  if (I->spatial) ALLOCT(h.agg_inv, 9 * (size_t)((nv + 21 * I->groups - 1) / (21 * I->groups)));
  {
    char* rec = nullptr;  // the record and, kCtlOffset bytes after it, the control block
    if ((rc = dalloc(I, &rec, kCtlOffset + sizeof(Ctl), 1)) != HOBAK_OK) return bail(rc);
    I->dev = (InstDev*)rec;
    h.ctl = ctl_of(I->dev);
  }
#undef ALLOC
#undef ALLOCT
  h.rest = rest;
  h.mass = mass;
  h.one_ring = one_ring;
  h.pin_target = pin_target;
  h.dminv = dminv;
  h.vol = vol;
  h.tets = t4d;
  h.slice_off = slice_off;
  h.row_len = row_len;
  h.col = col;
  h.slot_row = slot_row;
  h.cptr = cptr;
  h.cidx = cidx;
  h.mirror = mirror;
  h.aslot = aslot;
  h.evptr = evptr;
  h.evert = evert;
  h.eloc = eloc;
  h.vptr = vptr;
  h.fidx = fidx;
  h.pinned = pinned;
  h.coarse = coarse_on(ctx, I);
  h.agg = agg_on(ctx, I);
  I->ctl = h.ctl;
  // no solve active: both scalar sets done, so every PCG kernel passes this instance over
  Ctl ctl0;
  std::memset(&ctl0, 0, sizeof(ctl0));
  ctl0.sc[0].done = ctl0.sc[1].done = 1;
  ctl0.phase = PHASE_IDLE;
  std::vector<double> mass_h(3 * (size_t)nv);
  for (int i = 0; i < nv; i++)  // buildMassMatrix + scaleMass
    for (int c = 0; c < 3; c++) mass_h[3 * (size_t)i + c] = one_ring_h[i] * params->mass_scale;
  std::vector<int4> t4(nt);
  for (int t = 0; t < nt; t++) t4[t] = make_int4(tets[4 * t], tets[4 * t + 1], tets[4 * t + 2], tets[4 * t + 3]);
  if ((rc = upload(I, rest, rest_xyz, 3 * (size_t)nv)) ||
      (rc = upload(I, mass, mass_h.data(), mass_h.size())) ||
      (rc = upload(I, one_ring, one_ring_h.data(), one_ring_h.size())) ||
      (rc = upload(I, dminv, dminv_h.data(), dminv_h.size())) ||
      (rc = upload(I, vol, vol_h.data(), vol_h.size())) ||
      (rc = upload(I, t4d, t4.data(), t4.size())) ||
      (rc = upload(I, slice_off, P.slice_off.data(), P.slice_off.size())) ||
      (rc = upload(I, row_len, P.row_len.data(), P.row_len.size())) ||
      (rc = upload(I, col, P.col.data(), P.col.size())) ||
      (rc = upload(I, slot_row, P.slot_row.data(), P.slot_row.size())) ||
      (rc = upload(I, cptr, P.cptr.data(), P.cptr.size())) ||
      (rc = upload(I, cidx, P.cidx.data(), P.cidx.size())) ||
      (rc = upload(I, mirror, P.mirror.data(), P.mirror.size())) ||
      (rc = upload(I, aslot, P.aslot.data(), P.aslot.size())) ||
      (rc = upload(I, evptr, P.evptr.data(), P.evptr.size())) ||
      (rc = upload(I, evert, P.evert.data(), P.evert.size())) ||
      (rc = upload(I, eloc, P.eloc.data(), P.eloc.size())) ||
      (rc = upload(I, vptr, P.vptr.data(), P.vptr.size())) ||
      (rc = upload(I, fidx, P.fidx.data(), P.fidx.size())) ||
      (rc = upload(I, I->ctl, &ctl0, 1)) ||
      (rc = upload(I, I->dev, &h, 1)))
    return bail(rc);
  if (hipStreamSynchronize(ctx->stream) != hipSuccess)
    return bail(fail(ctx, HOBAK_E_HIP, "upload failed"));
  if ((rc = (I->tsize == 8 ? setup_constant_parts<double>(I) : setup_constant_parts<float>(I))) != HOBAK_OK)
    return bail(rc);
  ctx->insts.insert(I);
  *out = I;
  return HOBAK_OK;
}

void hobak_inst_destroy(hobak_inst* I) {
  if (!I) return;
  hobak_ctx* ctx = I->ctx;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  if (I->sched) sched_detach(I->sched, I);
  plans_forget(ctx, I);
  ctx->insts.erase(I);
  free_inst(I);
  delete I;
}

int hobak_inst_set_pinned(hobak_inst* I, const int32_t* vids, int32_t n, const double* targets) {
  if (!I || n < 0 || (n > 0 && !vids)) return HOBAK_E_INVALID;
  if (I->requests) return fail(I->ctx, HOBAK_E_INVALID, "instance has a step in flight");
  std::vector<unsigned char> flag(I->nv, 0);
  std::vector<double> tgt(3 * (size_t)I->nv, 0.0);
  for (int k = 0; k < n; k++) {
    if (vids[k] < 0 || vids[k] >= I->nv) return fail(I->ctx, HOBAK_E_INVALID, "pinned vertex out of range");
    const int v = I->perm.empty() ? vids[k] : I->iperm[vids[k]];
    flag[v] = 1;
    for (int c = 0; c < 3; c++) tgt[3 * (size_t)v + c] = targets ? targets[3 * (size_t)k + c] : 0.0;
  }
  I->pinned_ids.assign(vids, vids + n);
  const int coarse = coarse_on(I->ctx, I);
  const int agg = agg_on(I->ctx, I);
  if (coarse != I->h.coarse || agg != I->h.agg) {  // cached plans launch their members' variants
    I->h.coarse = coarse;
    I->h.agg = agg;
    TRY(upload(I, I->dev, &I->h, 1));
    plans_forget(I->ctx, I);
  }
  TRY(upload(I, (unsigned char*)I->h.pinned, flag.data(), flag.size()));
  TRY(upload(I, (double*)I->h.pin_target, tgt.data(), tgt.size()));
  HIPCHK(I->ctx, hipStreamSynchronize(I->ctx->stream));
  return HOBAK_OK;
}

// Surface vertices: vertices of tet faces that no other tet shares (TET_MESH.cpp:294-354, 567-593).
static std::vector<unsigned char> surface_flags(const hobak_inst* I, const std::vector<int32_t>& tets) {
  std::vector<std::array<int, 3>> faces;
  faces.reserve(4 * (size_t)I->nt);
  const int F[4][3] = {{0, 1, 3}, {0, 2, 1}, {0, 3, 2}, {1, 2, 3}};
  for (int t = 0; t < I->nt; t++)
    for (int f = 0; f < 4; f++) {
      std::array<int, 3> k = {tets[4 * (size_t)t + F[f][0]], tets[4 * (size_t)t + F[f][1]],
                              tets[4 * (size_t)t + F[f][2]]};
      std::sort(k.begin(), k.end());
      faces.push_back(k);
    }
  std::sort(faces.begin(), faces.end());
  std::vector<unsigned char> on(I->nv, 0);
  for (size_t i = 0; i < faces.size();) {
    size_t j = i + 1;
    while (j < faces.size() && faces[j] == faces[i]) j++;
    if (j - i < 2)
      for (int c = 0; c < 3; c++) on[faces[i][c]] = 1;
    i = j;
  }
  return on;
}

int hobak_inst_add_collision_cube(hobak_inst* I, const double* center, double scale) {
  if (!I || !center) return HOBAK_E_INVALID;
  hobak_ctx* ctx = I->ctx;
  if (I->requests) return fail(ctx, HOBAK_E_INVALID, "instance has a step in flight");
  if (I->tsize != 8 || I->h.refine) return fail(ctx, HOBAK_E_INVALID, "collision objects need the FP64 form");
  if (I->h.bdf) return fail(ctx, HOBAK_E_INVALID, "collision objects need one Newton iteration per step");
  if (!(scale > 0.0) || !std::isfinite(scale)) return fail(ctx, HOBAK_E_INVALID, "cube scale must be positive");
  if (I->h.nshapes >= kMaxShapes) return fail(ctx, HOBAK_E_INVALID, "too many collision objects");
  (void)hipSetDevice(ctx->device);
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  InstDev& h = I->h;
  if (!h.contact) {
    int rc;
    const size_t nv = (size_t)I->nv;
    unsigned char* surf;
    if ((rc = dalloc(I, &surf, nv)) || (rc = dalloc(I, &h.pc_obj, nv)) || (rc = dalloc(I, &h.pc_cl, 3 * nv)) ||
        (rc = dalloc(I, &h.pc_nl, 3 * nv)) || (rc = dalloc(I, &h.pc_sep, nv)) ||
        (rc = dalloc(I, &h.fmode, nv)) || (rc = dalloc(I, &h.fnrm, 3 * nv)) || (rc = dalloc(I, &h.zc, 3 * nv)) ||
        (rc = dalloc(I, &h.bunf, 3 * nv)) || (rc = dalloc(I, &h.part_sep, (size_t)h.nb[K_ROW])))
      return rc;
    std::vector<int32_t> tets(4 * (size_t)I->nt);
    HIPCHK(ctx, hipMemcpy(tets.data(), h.tets, sizeof(int32_t) * tets.size(), hipMemcpyDeviceToHost));
    const std::vector<unsigned char> on = surface_flags(I, tets);
    TRY(upload(I, surf, on.data(), nv));
    h.surf = surf;
    HIPCHK(ctx, hipMemsetAsync(h.pc_obj, 0xff, sizeof(int) * nv, ctx->stream));  // -1: none
    h.contact = 1;
  }
  double* q = h.shape[h.nshapes];
  for (int k = 0; k < 3; k++) q[k] = center[k];
  q[3] = scale;
  q[4] = (scale * scale) * (1.0 / ((scale * scale) * scale));  // Eigen's cofactor inverse (CUBE.cpp:28)
  h.nshapes++;
  TRY(upload(I, I->dev, &h, 1));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  plans_forget(ctx, I);  // cached plans were built without the contact kernels
  return HOBAK_OK;
}

int hobak_inst_plane_constraints(hobak_inst* I, int32_t* count, int32_t* ids, int32_t cap) {
  if (!I || !count) return HOBAK_E_INVALID;
  *count = 0;
  if (!I->h.contact) return HOBAK_OK;
  (void)hipSetDevice(I->ctx->device);
  std::vector<int> obj(I->nv);
  HIPCHK(I->ctx, hipMemcpyAsync(obj.data(), I->h.pc_obj, sizeof(int) * obj.size(), hipMemcpyDeviceToHost,
                                I->ctx->stream));
  HIPCHK(I->ctx, hipStreamSynchronize(I->ctx->stream));
  std::vector<int> on;
  for (int i = 0; i < I->nv; i++)
    if (obj[i] >= 0) on.push_back(I->perm.empty() ? i : I->perm[i]);
  std::sort(on.begin(), on.end());
  for (int k = 0; ids && k < (int)on.size() && k < cap; k++) ids[k] = on[k];
  *count = (int)on.size();
  return HOBAK_OK;
}

int hobak_inst_set_damping(hobak_inst* I, int32_t model, double mu) {
  if (!I) return HOBAK_E_INVALID;
  hobak_ctx* ctx = I->ctx;
  if (I->requests) return fail(ctx, HOBAK_E_INVALID, "instance has a step in flight");
  if (model != HOBAK_DAMPING_RAYLEIGH && model != HOBAK_DAMPING_GREEN)
    return fail(ctx, HOBAK_E_INVALID, "unknown damping model");
  if (model == HOBAK_DAMPING_GREEN && I->h.bdf)
    return fail(ctx, HOBAK_E_INVALID, "energy damping needs one Newton iteration per step (the "
                                      "reference's BDF_1 has no energy-damped solve)");
  if (model == HOBAK_DAMPING_GREEN && !std::isfinite(mu)) return fail(ctx, HOBAK_E_INVALID, "damping mu must be finite");
  if (model == HOBAK_DAMPING_GREEN && I->h.velform)
    return fail(ctx, HOBAK_E_INVALID, "energy damping is built for the position integrator only");
  (void)hipSetDevice(ctx->device);
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  InstDev& h = I->h;
  if (model == HOBAK_DAMPING_GREEN && !h.fdot) {
    double* fd;
    TRY(dalloc(I, &fd, 9 * (size_t)I->nt));
    HIPCHK(ctx, hipMemsetAsync(fd, 0, sizeof(double) * 9 * (size_t)I->nt, ctx->stream));  // held Fdot = 0
    h.fdot = fd;
  }
  h.damping = model == HOBAK_DAMPING_GREEN;
  h.dmu = h.damping ? mu : 0.0;
  // solveEnergyDamped has no Rayleigh terms: b = (M/dt) v + R + f, A = M/dt^2 - K
  h.ra = h.damping ? 0.0 : I->prm.rayleigh_alpha;
  h.rb = h.damping ? 0.0 : I->prm.rayleigh_beta;
  h.coarse = coarse_on(ctx, I);
  h.agg = agg_on(ctx, I);
  TRY(upload(I, I->dev, &h, 1));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  TRY(I->tsize == 8 ? setup_constant_parts<double>(I) : setup_constant_parts<float>(I));
  plans_forget(ctx, I);
  return HOBAK_OK;
}

int hobak_inst_compute_fdots(hobak_inst* I, const double* velocity) {
  if (!I) return HOBAK_E_INVALID;
  hobak_ctx* ctx = I->ctx;
  if (I->requests) return fail(ctx, HOBAK_E_INVALID, "instance has a step in flight");
  if (!I->h.fdot) return fail(ctx, HOBAK_E_INVALID, "velocity gradients need energy damping (hobak_inst_set_damping)");
  (void)hipSetDevice(ctx->device);
  const double* vel = I->h.v;
  if (velocity) {
    std::vector<double> vbuf;
    TRY(upload(I, I->h.tmp64, to_device_order(I, velocity, vbuf), (size_t)I->dofs));
    vel = I->h.tmp64;
  }
  HIPCHK(ctx, launch_fdots(I->h, vel, (double*)I->h.fdot, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return HOBAK_OK;
}

int hobak_inst_set_state(hobak_inst* I, const double* u, const double* v) {
  if (!I) return HOBAK_E_INVALID;
  if (I->requests) return fail(I->ctx, HOBAK_E_INVALID, "instance has a step in flight");
  std::vector<double> ub, vb;
  if (u) TRY(upload(I, I->h.u, to_device_order(I, u, ub), (size_t)I->dofs));
  if (v) TRY(upload(I, I->h.v, to_device_order(I, v, vb), (size_t)I->dofs));
  HIPCHK(I->ctx, hipStreamSynchronize(I->ctx->stream));
  for (auto& p : I->pred) p = -1;
  return HOBAK_OK;
}

int hobak_inst_get_state(hobak_inst* I, double* u, double* v) {
  if (!I) return HOBAK_E_INVALID;
  hipStream_t s = I->ctx->stream;
  if (u) HIPCHK(I->ctx, hipMemcpyAsync(u, I->h.u, sizeof(double) * I->dofs, hipMemcpyDeviceToHost, s));
  if (v) HIPCHK(I->ctx, hipMemcpyAsync(v, I->h.v, sizeof(double) * I->dofs, hipMemcpyDeviceToHost, s));
  HIPCHK(I->ctx, hipStreamSynchronize(s));
  to_caller_order(I, u);
  to_caller_order(I, v);
  return HOBAK_OK;
}

int hobak_step_batch(hobak_ctx* ctx, hobak_inst* const* insts, int32_t n, hobak_step_stats* stats) {
  if (!ctx || n < 0 || (n > 0 && !insts)) return HOBAK_E_INVALID;
  if (n == 0) return HOBAK_OK;
  (void)hipSetDevice(ctx->device);
  const Range r_call(n == 1 ? "hobak_step" : "hobak_step_batch");
  std::vector<hobak_inst*> g64, g32;
  std::vector<int> i64, i32;
  std::set<hobak_inst*> seen;
  for (int i = 0; i < n; i++) {
    hobak_inst* I = insts[i];
    if (!I || I->ctx != ctx) return fail(ctx, HOBAK_E_INVALID, "instance of another context");
    if (!seen.insert(I).second) return fail(ctx, HOBAK_E_INVALID, "instance listed twice in one batch");
    if (I->requests) return fail(ctx, HOBAK_E_INVALID, "instance has a step in flight");
    (I->tsize == 8 ? g64 : g32).push_back(I);
    (I->tsize == 8 ? i64 : i32).push_back(i);
  }
  auto t0 = std::chrono::steady_clock::now();
  int worst = HOBAK_OK;
  for (int pass = 0; pass < 2; pass++) {
    const auto& g = pass ? g32 : g64;
    const auto& ix = pass ? i32 : i64;
    if (g.empty()) continue;
    std::vector<hobak_step_stats> st(g.size());
    std::memset(st.data(), 0, sizeof(hobak_step_stats) * st.size());
    const int rc = run_batch(ctx, g, st.data());
    if (rc == HOBAK_E_HIP || rc == HOBAK_E_INVALID || rc == HOBAK_E_NOMEM) return rc;
    worst = worse(worst, rc);
    if (stats)
      for (size_t j = 0; j < g.size(); j++) stats[ix[j]] = st[j];
  }
  const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (stats)
    for (int i = 0; i < n; i++) stats[i].seconds = sec;
  return worst;
}

int hobak_step(hobak_inst* I, hobak_step_stats* st) {
  if (!I) return HOBAK_E_INVALID;
  return hobak_step_batch(I->ctx, &I, 1, st);
}

int hobak_get_bsr_pattern(hobak_inst* I, int32_t* nnzb, int32_t* row_ptr, int32_t* col_idx) {
  if (!I) return HOBAK_E_INVALID;
  if (nnzb) *nnzb = I->P.nnzb;
  if (row_ptr) std::copy(I->orow_ptr.begin(), I->orow_ptr.end(), row_ptr);
  if (col_idx) std::copy(I->ocol_idx.begin(), I->ocol_idx.end(), col_idx);
  return HOBAK_OK;
}

int hobak_eval_forces_hessian(hobak_inst* I, const double* u, double* forces, double* kvals) {
  if (!I || !u) return HOBAK_E_INVALID;
  if (I->requests) return fail(I->ctx, HOBAK_E_INVALID, "instance has a step in flight");
  (void)hipSetDevice(I->ctx->device);
  return I->tsize == 8 ? eval_forces_hessian<double>(I, u, forces, kvals)
                       : eval_forces_hessian<float>(I, u, forces, kvals);
}

int hobak_eval_forces_noalloc(hobak_inst* I, const double* u_dev, double* forces_dev) {
  if (!I || !u_dev || !forces_dev) return HOBAK_E_INVALID;
  hobak_ctx* ctx = I->ctx;
  if (!I->perm.empty())
    return fail(ctx, HOBAK_E_INVALID, "hobak_eval_forces_noalloc: the instance is held relabelled (row sort or aggregates); "
                                      "device vectors would need the caller's numbering");
  const size_t need = (size_t)12 * I->nt * I->tsize;
  if (need > (size_t)HOBAK_NOALLOC_SCRATCH_BYTES)
    return fail(ctx, HOBAK_E_INVALID,
                "hobak_eval_forces_noalloc: " + std::to_string(I->nt) + " tets need " + std::to_string(need) +
                    " bytes of force staging; the static scratch holds " +
                    std::to_string((size_t)HOBAK_NOALLOC_SCRATCH_BYTES) + " (at most " +
                    std::to_string((size_t)HOBAK_NOALLOC_SCRATCH_BYTES / (12 * (size_t)I->tsize)) + " tets)");
  if (I->requests) return fail(ctx, HOBAK_E_INVALID, "instance has a step in flight");
  if (!ctx->scratch) return fail(ctx, HOBAK_E_HIP, "no scratch address for this device");
  (void)hipSetDevice(ctx->device);
  hipStream_t s = ctx->stream;
  // a batch of one names its record directly: no device table is needed
  const Batch be{nullptr, nullptr, I->dev, 1, I->h.nb[K_ELEM]};
  const Batch br{nullptr, nullptr, I->dev, 1, I->h.nb[K_ROW]};
  if (I->tsize == 8) {
    HIPCHK(ctx, launch_element<double>(be, ELEM_EXT, 0, ctx->grid_cap, s, u_dev, ctx->scratch));
    HIPCHK(ctx, launch_force_gather<double>(br, ctx->grid_cap, s, ctx->scratch, forces_dev));
  } else {
    HIPCHK(ctx, launch_element<float>(be, ELEM_EXT, 0, ctx->grid_cap, s, u_dev, ctx->scratch));
    HIPCHK(ctx, launch_force_gather<float>(br, ctx->grid_cap, s, ctx->scratch, forces_dev));
  }
  HIPCHK(ctx, hipStreamSynchronize(s));
  return HOBAK_OK;
}

int hobak_eval_damping(hobak_inst* I, const double* u, double* forces, double* cvals) {
  if (!I || !u) return HOBAK_E_INVALID;
  if (I->requests) return fail(I->ctx, HOBAK_E_INVALID, "instance has a step in flight");
  if (!I->h.damping) return fail(I->ctx, HOBAK_E_INVALID, "the instance has no energy damping");
  (void)hipSetDevice(I->ctx->device);
  return I->tsize == 8 ? eval_damping<double>(I, u, forces, cvals) : eval_damping<float>(I, u, forces, cvals);
}

int hobak_get_last_system(hobak_inst* I, double* rhs, double* lhs_vals) {
  if (!I) return HOBAK_E_INVALID;
  if (!I->has_system) return fail(I->ctx, HOBAK_E_INVALID, "no system formed since the last step");
  hipStream_t s = I->ctx->stream;
  if (rhs) {
    HIPCHK(I->ctx, hipMemcpyAsync(rhs, I->h.rhs64, sizeof(double) * I->dofs, hipMemcpyDeviceToHost, s));
    HIPCHK(I->ctx, hipStreamSynchronize(s));
    to_caller_order(I, rhs);
  }
  if (lhs_vals)
    TRY(I->tsize == 8 ? download_sell_as_bsr<double>(I, I->h.A, lhs_vals, 1.0)
                      : download_sell_as_bsr<float>(I, I->h.A, lhs_vals, 1.0));
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
  I->spmv_bytes_sum = 0.0;
  I->spmv_count = 0;
  return HOBAK_OK;
}

int hobak_kernel_stats(hobak_inst* I, double* spmv_avg_ms, int64_t* spmv_count, double* spmv_bytes) {
  if (!I) return HOBAK_E_INVALID;
  const double cnt = (double)I->spmv_count;
  if (spmv_avg_ms) *spmv_avg_ms = I->spmv_count ? I->spmv_ms_sum / cnt : 0.0;
  if (spmv_count) *spmv_count = I->spmv_count;
  if (spmv_bytes) *spmv_bytes = I->spmv_count ? I->spmv_bytes_sum / cnt : spmv_bytes_of(I);
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
      if (!ctx->insts.empty() && value != ctx->slice_rows)
        return fail(ctx, HOBAK_E_INVALID, "the slice height is fixed once instances exist");
      ctx->slice_rows = value;
      return HOBAK_OK;
    case HOBAK_TUNE_POLL_CHUNK:
      if (value < 1 || value > 4096) return fail(ctx, HOBAK_E_INVALID, "poll chunk out of range");
      ctx->chunk_iters = (value + 1) & ~1;
      return HOBAK_OK;
    case HOBAK_TUNE_MAX_CHUNKS:
      if (value < 1 || value > 65536) return fail(ctx, HOBAK_E_INVALID, "max chunks out of range");
      ctx->max_chunks = value;
      return HOBAK_OK;
    case HOBAK_TUNE_GRAPHS:
      if (value != 0 && value != 1) return fail(ctx, HOBAK_E_INVALID, "graphs is 0 or 1");
      ctx->use_graphs = value;
      return HOBAK_OK;
    case HOBAK_TUNE_GRID_CAP:
      if (value < 0) return fail(ctx, HOBAK_E_INVALID, "grid cap must be >= 0");
      ctx->grid_cap = value;
      return HOBAK_OK;
    case HOBAK_TUNE_SPMV_PAD:
      if (value != 0 && value != 1) return fail(ctx, HOBAK_E_INVALID, "spmv pad is 0 or 1");
      ctx->spmv_pad = value;
      return HOBAK_OK;
    case HOBAK_TUNE_SPMV_STREAM_MB:
      if (value < 0) return fail(ctx, HOBAK_E_INVALID, "stream threshold must be >= 0");
      ctx->spmv_stream_mb = value;
      return HOBAK_OK;
    case HOBAK_TUNE_ROW_SORT:
      if (value != 0 && value != 1) return fail(ctx, HOBAK_E_INVALID, "row sort is 0 or 1");
      ctx->row_sort = value;
      return HOBAK_OK;
    case HOBAK_TUNE_COMPACT_PCT:
      if (value < 0 || value > 100) return fail(ctx, HOBAK_E_INVALID, "compact percentage must be 0..100");
      ctx->compact_pct = value;
      return HOBAK_OK;
    case HOBAK_TUNE_UPDATE_GROUPS:
      if (value != 1 && value != 2 && value != 4)
        return fail(ctx, HOBAK_E_INVALID, "update groups must be 1, 2 or 4");
      if (!ctx->insts.empty() && value != ctx->update_groups)
        return fail(ctx, HOBAK_E_INVALID, "the update groups are fixed once instances exist");
      ctx->update_groups = value;
      return HOBAK_OK;
    case HOBAK_TUNE_P_FORM:
      if (value < 0 || value > 2) return fail(ctx, HOBAK_E_INVALID, "p-form is 0, 1 or 2");
      if (value != ctx->p_form) {  // cached graphs hold the old choice
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        for (auto* P : ctx->plans) plan_free(P);
        ctx->plans.clear();
        for (auto* S : ctx->scheds) S->mplan = nullptr;
      }
      ctx->p_form = value;
      return HOBAK_OK;
    case HOBAK_TUNE_COARSE:
      if (value != 0 && value != 1) return fail(ctx, HOBAK_E_INVALID, "coarse is 0 or 1");
      if (!ctx->insts.empty() && value != ctx->coarse)
        return fail(ctx, HOBAK_E_INVALID, "the coarse correction is fixed once instances exist");
      ctx->coarse = value;
      return HOBAK_OK;
    case HOBAK_TUNE_AGG:
      if (value < 0 || value > 2) return fail(ctx, HOBAK_E_INVALID, "aggregates is 0, 1 or 2");
      if (!ctx->insts.empty() && value != ctx->agg)
        return fail(ctx, HOBAK_E_INVALID, "the local aggregates are fixed once instances exist");
      ctx->agg = value;
      return HOBAK_OK;
    case HOBAK_TUNE_UPDATE_WGS:
      if (value < 1 || value > 65536) return fail(ctx, HOBAK_E_INVALID, "update workgroups out of range");
      ctx->update_cap = value;
      return HOBAK_OK;
    default:
      return fail(ctx, HOBAK_E_INVALID, "unknown tuning key");
  }
}

// ---- continuous-batching scheduler ----
int hobak_sched_create(hobak_ctx* ctx, hobak_sched** out) {
  if (!ctx || !out) return HOBAK_E_INVALID;
  hobak_sched* S = new (std::nothrow) hobak_sched;
  if (!S) return fail(ctx, HOBAK_E_NOMEM, "host allocation failed");
  S->ctx = ctx;
  ctx->scheds.insert(S);
  *out = S;
  return HOBAK_OK;
}

void hobak_sched_destroy(hobak_sched* S) {
  if (!S) return;
  hobak_ctx* ctx = S->ctx;
  (void)hipStreamSynchronize(ctx->stream);
  std::vector<hobak_inst*> m = S->members;
  for (auto* I : m) sched_detach(S, I);
  ctx->scheds.erase(S);
  delete S;
}

int hobak_sched_submit(hobak_sched* S, hobak_inst* I, uint64_t tag) {
  if (!S || !I) return HOBAK_E_INVALID;
  hobak_ctx* ctx = S->ctx;
  if (I->ctx != ctx) return fail(ctx, HOBAK_E_INVALID, "instance of another context");
  if (I->sched && I->sched != S) return fail(ctx, HOBAK_E_INVALID, "instance is attached to another scheduler");
  if (!I->sched) {
    if (!S->members.empty() && (S->members[0]->tsize != I->tsize || S->members[0]->P.shift != I->P.shift))
      return fail(ctx, HOBAK_E_INVALID, "a scheduler's instances must share precision and slice height");
    if (I->requests) return fail(ctx, HOBAK_E_INVALID, "instance has a step in flight");
    S->members.push_back(I);
    S->dirty = true;
    I->sched = S;
  }
  hobak_sched::Run r;
  r.I = I;
  r.tag = tag;
  r.t_submit = clock_s(S);
  S->queue.push_back(r);
  I->requests++;
  return HOBAK_OK;
}

int hobak_sched_detach(hobak_sched* S, hobak_inst* I) {
  if (!S || !I) return HOBAK_E_INVALID;
  if (I->sched != S) return fail(S->ctx, HOBAK_E_INVALID, "instance is not attached to this scheduler");
  if (I->requests) return fail(S->ctx, HOBAK_E_INVALID, "instance has a step in flight");
  sched_detach(S, I);
  return HOBAK_OK;
}

int hobak_sched_tick(hobak_sched* S, hobak_completion* out, int32_t max_out, int32_t* n_out,
                     int32_t* in_flight) {
  if (!S || max_out < 0 || (max_out > 0 && !out)) return HOBAK_E_INVALID;
  (void)hipSetDevice(S->ctx->device);
  if (!S->queue.empty() || !S->active.empty()) {
    const Range r_call("hobak_sched_tick");
    const auto t0 = std::chrono::steady_clock::now();
    const int rc = tick(S);
    S->stat[5] += 1e3 * std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (rc != HOBAK_OK) return rc;
  }
  const int n = std::min<int>(max_out, (int)S->done.size());
  for (int i = 0; i < n; i++) out[i] = S->done[i];
  S->done.erase(S->done.begin(), S->done.begin() + n);
  if (n_out) *n_out = n;
  if (in_flight) *in_flight = (int32_t)(S->queue.size() + S->active.size() + S->done.size());
  return HOBAK_OK;
}

double hobak_sched_clock(hobak_sched* S) { return S ? clock_s(S) : 0.0; }

int hobak_sched_stats(hobak_sched* S, double* out) {
  if (!S || !out) return HOBAK_E_INVALID;
  std::memcpy(out, S->stat, sizeof(S->stat));
  return HOBAK_OK;
}

int hobak_sched_set_kernel_sampling(hobak_sched* S, int32_t every) {
  if (!S || every < 0) return HOBAK_E_INVALID;
  hobak_ctx* ctx = S->ctx;
  if (every > 0 && ctx->sample_ev.empty()) {
    (void)hipSetDevice(ctx->device);
    ctx->sample_ev.assign(2 * kMaxSamples, nullptr);
    for (auto& e : ctx->sample_ev) HIPCHK(ctx, hipEventCreate(&e));
  }
  S->spmv_every = every;
  S->sampling = every > 0;
  S->sample_it = 0;
  S->samples.clear();
  S->spmv_ms = S->spmv_bytes = 0.0;
  S->spmv_n = S->spmv_launches = 0;
  return HOBAK_OK;
}

int hobak_sched_kernel_stats(hobak_sched* S, double* spmv_avg_ms, int64_t* spmv_count, double* spmv_bytes) {
  if (!S) return HOBAK_E_INVALID;
  const double n = S->spmv_n ? (double)S->spmv_n : 1.0;
  if (spmv_avg_ms) *spmv_avg_ms = S->spmv_n ? S->spmv_ms / n : 0.0;
  if (spmv_count) *spmv_count = S->spmv_n;
  if (spmv_bytes) *spmv_bytes = S->spmv_launches ? S->spmv_bytes / (double)S->spmv_launches : 0.0;
  return HOBAK_OK;
}

}  // extern "C"
