// Batched contractions: many independent pairs in one persistent launch.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <memory>
#include <vector>

#include "handle.hpp"
#include "lowering.hpp"
#include "kernels.hpp"

namespace mh {

// One contraction of the batch, as the persistent kernel reads it.
struct BatchPair {
  const void* A;
  const void* B;
  void* C;
  int32_t M, N, K, NB;  // fused extents and batch (hyper) count
  int32_t tiles_m, tiles_n, tile, a_kmajor, b_kmajor;
  int32_t c_mfast;  // C's contiguous direction is M (else N)
  int32_t c_stream; // C written with nontemporal stores (kStreamStoreBytes)
  // split K: ksplit slices of kchunk elements per output tile (1 and K when
  // not split); the slices' partial tiles and the tiles' arrival counters
  int32_t ksplit, kchunk;
  void* partial;
  int32_t* tilecnt;
  OffTab rowA, kA, kB, colB, rowC, colC, bA, bB, bC;
};

// Output tile shapes: rows (M) x columns (N) per workgroup, and the 4 waves'
// arrangement wm x wn; every wave holds 4 MFMA 16x16 tiles.
struct TileShape {
  int bm, bn, wm, wn;
};
constexpr TileShape kTileShapes[MUSCLE_HIP_TILE_COUNT] = {{64, 64, 2, 2}, {32, 128, 1, 4}, {128, 32, 4, 1}};

// The persistent kernel's work list is split into one queue per XCD (the
// dispatcher deals workgroups round-robin over the XCDs, so workgroup b runs
// on XCD b % kXcds): each pair's tiles are cut into kXcds contiguous runs, one
// per queue, so the tiles that share A and B panels meet in one L2.
constexpr int kXcds = 8;
// per-XCD queues are used when the persistent items number at least this
// many per workgroup of the launch (else one queue)
constexpr int kXcdQueueItems = 4;

struct BatchArgs {
  const BatchPair* pairs;
  const int2* items;  // (pair, tile within pair), per queue largest pairs first
  int32_t* counter;   // work queue heads (kXcds of them)
  int32_t qoff[kXcds + 1];  // queue q holds items [qoff[q], qoff[q + 1])
  int32_t nq;               // queues in use: kXcds, or 1 (queue 0)
  int32_t* tilecnt;   // arrival counters of the split pairs' tiles, cleared with the queue head
  int64_t ntilecnt;
  int64_t nitems;
  int32_t complex3m;
  int32_t item_off;   // byte offset of the broadcast work item in dynamic LDS
};

// Pairs whose contracted extent K is at most this run in the shallow kernel
// (vector FMAs, one output tile per workgroup, several workgroups per CU)
// instead of the persistent MFMA launch.
constexpr int kShallowK = 16;

// Pairs whose output has a side of at most kSkinnyMaxS elements and whose
// contracted extent is at most kSkinnyMaxK run in the skinny kernel (vector
// FMAs; one workgroup covers the whole short side and a run of the long side)
// instead of the MFMA or shallow launches, whose tiles are at least 32 wide.
// A pair of fewer than kSkinnyMinTiles MFMA tiles stays on the tile routes:
// dropping its padding gains little, while a skinny item runs all of K in one
// workgroup and the launch adds to a small batch's critical path.
constexpr int kSkinnyMaxS = 16;
constexpr int kSkinnyMaxK = 1024;
constexpr int64_t kSkinnyMinTiles = 16;

inline bool skinny_pair(int64_t m, int64_t n, int64_t k) {
  return std::min(m, n) <= kSkinnyMaxS && k <= kSkinnyMaxK;
}

// Skinny item geometry for a short side of s elements and o outputs per
// thread (4 for complex FP64, else 8): st = s padded to a power of two, j =
// long positions per thread (at most o / 2, so that a thread's ring of o
// operand loads runs two or more steps of k ahead), q = short positions per
// thread (o / j), lw = threads along the long side (the 512 threads form
// 512 / lw groups of q short positions; sp = the short side they cover),
// lt = long positions per item (lw * j).
struct SkinnyGeom {
  int st, j, q, sp, lw, lt;
};
__host__ __device__ inline SkinnyGeom skinny_geom(int64_t s, int o) {
  SkinnyGeom g;
  g.st = s <= 1 ? 1 : s <= 2 ? 2 : s <= 4 ? 4 : s <= 8 ? 8 : 16;
  g.j = o / g.st < 1 ? 1 : o / g.st > o / 2 ? o / 2 : o / g.st;
  g.q = o / g.j;
  g.sp = g.st > g.q ? g.st : g.q;
  g.lw = 512 / (g.sp / g.q);
  g.lt = g.lw * g.j;
  return g;
}
__host__ __device__ inline int skinny_outputs(bool dbl, bool cplx) { return dbl && cplx ? 4 : 8; }

struct ShallowArgs {
  const BatchPair* pairs;
  const int2* items;  // the shallow (or skinny) pairs' items
  int64_t nitems;
};

// Launches the shallow kernel over a.nitems tiles of pairs whose C runs along
// M (cm) or N (nothing when there are none).
void launch_shallow(bool dbl, bool cplx, bool cm, bool cs, const ShallowArgs& a, hipStream_t s);
// Launches the skinny kernel over a.nitems items (nothing when there are none).
void launch_skinny(bool dbl, bool cplx, const ShallowArgs& a, hipStream_t s);
// Launches the persistent kernel (grid = workgroups) over a's items, then the
// split kernel over split's items (the K slices of split pairs), preceded by a
// kernel that resets the work counter and the split tiles' arrival counters
// (no memset nodes: the launch is graph captured).  lds_bytes: dynamic LDS
// for the largest tile shape present.
void launch_batch(bool dbl, bool cplx, const BatchArgs& a, const BatchArgs& split, int workgroups,
                  size_t lds_bytes, hipStream_t s);
// Dynamic LDS one workgroup needs for a tile shape.
size_t batch_lds_bytes(bool dbl, bool cplx, int tile);
// Where the work-item broadcast word goes in a launch's dynamic LDS of
// lds_bytes (grown by 16 bytes when it goes at the end), so that two
// workgroups still fit a CU's 160 KiB.
size_t batch_item_offset(bool dbl, bool cplx, size_t& lds_bytes);
// Persistent workgroups per CU the kernel is built for.
constexpr int kBatchWgPerCu = 2;

struct Batch {
  Handle* h = nullptr;
  int dtype = 0;
  // complex product form (3: Gauss, 4: four products), read from the handle
  // when the batch is built and used by every later run, so a graph captured
  // once and its replays agree with the plain loop
  int complex_mult = 3;
  std::vector<muscle_hip_tensor_t> cs, as, bs;  // for the per-pair plain loop
  std::vector<std::vector<int32_t>> modes;      // storage behind cs/as/bs
  std::vector<std::vector<int64_t>> ext, str;
  std::vector<BatchPair> hpairs;
  // persistent items (largest K first), then the split pairs' slice items,
  // then the shallow items, then the skinny items
  std::vector<int2> hitems;
  int32_t qoff[kXcds + 1] = {0};  // the persistent items' per-XCD queues (offsets into hitems)
  int32_t nq = 1;                 // queues in use (kXcds or 1)
  int64_t nskinny = 0;   // the skinny items at the end of hitems
  int64_t nshallow = 0;  // the shallow items before them
  // of them, those of pairs with C along N, along M, then the same two with
  // streaming stores for C
  int64_t nshallow_cm[4] = {0, 0, 0, 0};
  int64_t nsplit = 0;    // the slice items before them
  // split K (per pair of the work list, by position): slices, K per slice,
  // offsets of the pair's partial tiles (bytes) and arrival counters
  std::vector<int32_t> ksplit, kchunk;
  std::vector<int64_t> partial_off, tilecnt_off;
  int64_t ntilecnt = 0;
  size_t partial_bytes = 0;
  int64_t split_items = 0;
  bool allow_split = true;  // off for the autotuner's timing batches
  std::vector<int64_t> htab;
  std::vector<int> pair_class;
  // per pair, in the order given: the route its tiles take and its fused shape
  std::vector<muscle_hip_batch_pair_info_t> pinfo;
  void* dmem = nullptr;  // pairs | items | counter | tables
  BatchPair* dpairs = nullptr;
  int2* ditems = nullptr;
  int32_t* dcounter = nullptr;
  int32_t* dtilecnt = nullptr;
  int workgroups = 0;
  size_t lds = 0;
  double flops = 0, bytes = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t capture_stream = nullptr;
  // the largest pairs run as single-contraction launches (packed operands,
  // the plain GEMM tiles) inside the same graph, after the persistent launch;
  // they own their plans and one workspace
  std::vector<char> solo;                        // per pair
  std::vector<int64_t> solo_pairs;               // their indices
  std::vector<std::unique_ptr<Plan>> solo_plans;
  void* solo_ws = nullptr;
  size_t solo_ws_size = 0;
  ~Batch();
};

// Pairs of at least this many index points (M * N * K * batch) leave the
// persistent launch for a single-contraction launch of their own.
constexpr double kBatchSoloPoints = double(int64_t(1) << 29);

// Location of one two-level table inside a batch's table arena.
struct TabRef {
  size_t lo, hi;
  uint32_t P, single;
};

// One pair after host planning.
struct PlannedPair {
  ContractionDesc d;
  TabRef t[9];
  int tile = 0;
  int a_kmajor = 0, b_kmajor = 0, c_mfast = 0;
  int64_t cost = 0;
  const void* a = nullptr;  // kernel operands (after the lowering's A/B exchange)
  const void* b = nullptr;
  void* c = nullptr;
  const int64_t* dtab = nullptr;  // device tables of this pair; null: the batch's own table region
  const std::vector<int64_t>* htab = nullptr;  // the host copy its TabRefs index
  bool solo = false;              // runs as a single-contraction launch, not in the work list
};

// Device layout of a planned batch: pairs | items | counter | tables, copied
// from the host (the first `copy` bytes), then the split pairs' arrival
// counters and partial tiles, device-only.
struct BatchLayout {
  size_t pairs, items, counter, tab, tilecnt, partial, copy, total;
};
// A pair of at most kSplitMaxTiles output tiles and at least kSplitMinKTiles
// K-tiles (16 K each) runs as K slices (split kernel).
constexpr int64_t kSplitMaxTiles = 2, kSplitMinKTiles = 32;

void plan_batch(Handle& h, int64_t n, const muscle_hip_tensor_t* cs, const muscle_hip_tensor_t* as,
                const muscle_hip_tensor_t* bs, const int32_t* tiles, bool keep_descs, Batch& B,
                std::vector<PlannedPair>& lw);
// Tables and tile of one pair whose lowering is in L.d: the nine offset
// tables are appended to htab (L.t holds their positions); tile < 0 picks the
// tuned or default shape for the pair's size class.
// single_max: classes up to this many points get one-level tables.
void plan_pair(const Handle& h, PlannedPair& L, std::vector<int64_t>& htab, int tile,
               int64_t single_max = kSingleLevelMax);
// Batches planned per scheduler tick use two-level tables from this many
// points on, which keeps the tables (host time and upload) small at one more
// lookup per index.
constexpr int64_t kTickSingleMax = 256;
// Work list (the persistent items, largest K first, then the shallow pairs'
// items), workgroup count and LDS size of the pairs in lw.
void finish_batch(Handle& h, Batch& B, const std::vector<const PlannedPair*>& lw);
BatchLayout batch_layout(const Batch& B, int64_t n);
void place_batch(Batch& B, const std::vector<const PlannedPair*>& lw, char* base, char* staging, hipStream_t s);
void batch_launch(Batch& b, hipStream_t s);
std::unique_ptr<Batch> build_batch(Handle& h, int64_t n, const muscle_hip_tensor_t* cs, const muscle_hip_tensor_t* as,
                                   const muscle_hip_tensor_t* bs, const int32_t* tiles, bool allow_split = true);
void batch_run(Batch& b, int mode);
void batch_autotune(Batch& b, int samples, int reps);
int batch_tuning_class(int64_t m, int64_t n);

}  // namespace mh
