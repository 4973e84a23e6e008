// Host side of the batched contraction path: planning (lowering + compact
// offset tables + work list), graph-captured persistent launches, the plain
// per-pair loop it is compared with, and the tile-shape autotuner.
#include "batch.hpp"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <mutex>
#include <numeric>
#include <utility>

#include "lowering.hpp"

namespace mh {

Batch::~Batch() {
  // batches with their own device memory may still be read by queued work
  if (h && (dmem || exec || solo_ws)) (void)hipStreamSynchronize(h->stream);
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  if (capture_stream) (void)hipStreamDestroy(capture_stream);
  if (dmem) (void)hipFree(dmem);
  if (solo_ws) (void)hipFree(solo_ws);
}

namespace {

int bucket(int64_t x) { return x < 24 ? 0 : x < 48 ? 1 : x < 96 ? 2 : x < 192 ? 3 : 4; }

int tuning_class(int64_t m, int64_t n) { return bucket(m) * 5 + bucket(n); }

// Shape heuristic when nothing is tuned: keep most of the tile inside C.
int default_tile(int64_t m, int64_t n) {
  if (m < 48 && n >= 96) return MUSCLE_HIP_TILE_32x128;
  if (n < 48 && m >= 96) return MUSCLE_HIP_TILE_128x32;
  return MUSCLE_HIP_TILE_64x64;
}

size_t al(size_t v) { return (v + 255) & ~size_t(255); }

TabRef add_table(std::vector<int64_t>& h, const std::vector<Dim>& dims, int slot, int64_t single_max) {
  std::vector<Dim> lo, hi;
  int64_t P;
  split_dims(dims, lo, hi, P, single_max);
  TabRef r;
  const std::vector<int64_t> tl = offset_table(lo, slot), th = offset_table(hi, slot);
  r.lo = h.size();
  if (tl.empty()) h.push_back(0);
  else h.insert(h.end(), tl.begin(), tl.end());
  r.hi = h.size();
  if (th.empty()) h.push_back(0);
  else h.insert(h.end(), th.begin(), th.end());
  r.P = (uint32_t)P;
  r.single = hi.empty() ? 1u : 0u;
  return r;
}

// C is written with nontemporal stores: complex FP64 outputs of
// kStreamStoreBytes or more.
bool c_stream_of(const ContractionDesc& d) {
  const double es = (double)dtype_size(d.dt[2]);
  return es >= 16 && (double)d.M * (double)d.N * (double)d.B * es >= kStreamStoreBytes;
}

// Host value of entry i of a table (the device OffTab's lookup).
int64_t tab_at(const std::vector<int64_t>& h, const TabRef& r, int64_t i) {
  if (r.single) return h[r.lo + (size_t)i];
  return h[r.lo + (size_t)(i % r.P)] + h[r.hi + (size_t)(i / r.P)];
}

OffTab make_tab(const int64_t* base, const TabRef& r) {
  OffTab t;
  t.lo = base + r.lo;
  t.hi = base + r.hi;
  t.div = FastDiv(r.P);
  t.base = 0;
  t.single = r.single;
  return t;
}

void copy_desc(Batch& b, const muscle_hip_tensor_t& t, std::vector<muscle_hip_tensor_t>& dst) {
  b.modes.emplace_back(t.modes, t.modes + t.rank);
  b.ext.emplace_back(t.extents, t.extents + t.rank);
  b.str.emplace_back(t.strides, t.strides + t.rank);
  dst.push_back(t);
}

}  // namespace

int batch_tuning_class(int64_t m, int64_t n) { return tuning_class(m, n); }

void plan_pair(const Handle& h, PlannedPair& L, std::vector<int64_t>& htab, int tile, int64_t single_max) {
  const int T3[9][2] = {{0, 0}, {2, 0}, {2, 1}, {1, 1}, {0, 2}, {1, 2}, {3, 0}, {3, 1}, {3, 2}};
  // Order of the free modes inside M and N: a tile (or skinny item) covers the
  // lowest positions of both, so the modes along which a 128-byte line of the
  // operand or of C runs go first, in C's stride order, then the others in C's
  // stride order.  Each line of A, B and C is then read or written by one tile
  // instead of being fetched again by tiles on other XCDs.  The results do not
  // depend on this order (each output sums its K products in K's order).
  {
    const int64_t line = std::max<int64_t>(1, 128 / (int64_t)dtype_size(L.d.dt[2]));
    auto absl = [](int64_t x) { return x < 0 ? -x : x; };
    auto order = [&](std::vector<Dim>& v, int op) {
      std::stable_sort(v.begin(), v.end(), [&](const Dim& x, const Dim& y) {
        const bool px = absl(x.s[op]) < line || absl(x.s[2]) < line;
        const bool py = absl(y.s[op]) < line || absl(y.s[2]) < line;
        if (px != py) return px;
        return absl(x.s[2]) < absl(y.s[2]);
      });
    };
    order(L.d.m, 0);
    order(L.d.n, 1);
  }
  const ContractionDesc& d = L.d;
  const int64_t lim = int64_t(1) << 31;
  if (d.M >= lim || d.N >= lim || d.K >= lim || d.B >= lim)
    fail(MUSCLE_HIP_UNSUPPORTED, "batched pair with an index class of 2^31 or more elements");
  const std::vector<Dim>* cls[4] = {&d.m, &d.n, &d.k, &d.batch};
  for (int t = 0; t < 9; ++t) L.t[t] = add_table(htab, *cls[T3[t][0]], T3[t][1], single_max);
  L.htab = &htab;
  if (tile < 0 || tile >= MUSCLE_HIP_TILE_COUNT) {
    const int tc = h.tune[tuning_class(d.M, d.N)];
    tile = tc >= 0 ? tc : default_tile(d.M, d.N);
  }
  L.tile = tile;
  L.cost = (d.K + 15) / 16;
  auto minst = [](const std::vector<Dim>& v, int w) {
    int64_t b = INT64_MAX;
    for (auto& x : v)
      if (x.s[w] != 0) b = std::min(b, x.s[w] < 0 ? -x.s[w] : x.s[w]);
    return b;
  };
  L.a_kmajor = minst(d.k, 0) < minst(d.m, 0);
  L.b_kmajor = minst(d.k, 1) < minst(d.n, 1);
  L.c_mfast = minst(d.m, 2) < minst(d.n, 2);
}

void finish_batch(Handle& h, Batch& B, const std::vector<const PlannedPair*>& lw) {
  const int64_t n = (int64_t)lw.size();
  B.hitems.clear();
  B.lds = 0;
  // work list: pairs with the longest tiles (largest K) first
  std::vector<int64_t> order((size_t)n);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(),
                   [&](int64_t x, int64_t y) { return lw[(size_t)x]->cost > lw[(size_t)y]->cost; });
  bool used[MUSCLE_HIP_TILE_COUNT] = {false, false, false};
  const bool cp = is_complex(B.dtype);
  int cus = h.cus;
  if (cus <= 0) {
    MH_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, h.device));
    h.cus = cus;
  }
  B.workgroups = std::max(1, cus) * kBatchWgPerCu;
  auto tiles_of = [&](const PlannedPair& L) {
    const TileShape& ts = kTileShapes[L.tile];
    return ((L.d.M + ts.bm - 1) / ts.bm) * ((L.d.N + ts.bn - 1) / ts.bn) * L.d.B;
  };
  B.ksplit.assign((size_t)n, 1);
  B.kchunk.assign((size_t)n, 0);
  B.partial_off.assign((size_t)n, 0);
  B.tilecnt_off.assign((size_t)n, 0);
  B.ntilecnt = 0;
  B.partial_bytes = 0;
  B.split_items = 0;
  const size_t es = dtype_size(B.dtype);
  B.pinfo.assign((size_t)n, muscle_hip_batch_pair_info_t());
  for (int64_t i = 0; i < n; ++i) {
    const PlannedPair& L = *lw[(size_t)i];
    muscle_hip_batch_pair_info_t& q = B.pinfo[(size_t)i];
    std::memset(&q, 0, sizeof q);
    q.route = L.d.empty ? MUSCLE_HIP_PAIR_EMPTY : L.solo ? MUSCLE_HIP_PAIR_SOLO : MUSCLE_HIP_PAIR_TILES;
    q.tile = L.d.empty || L.solo ? -1 : L.tile;
    q.size_class = tuning_class(L.d.M, L.d.N);
    q.ksplit = 1;
    q.m = L.d.M;
    q.n = L.d.N;
    q.k = L.d.K;
    q.batch = L.d.B;
  }
  std::vector<int2> shallow[4], sliced, skinny, xq[kXcds], single;
  double xload[kXcds] = {0, 0, 0, 0, 0, 0, 0, 0};
  const int sko = skinny_outputs(is_double(B.dtype), cp);
  for (int64_t i : order) {
    const PlannedPair& L = *lw[(size_t)i];
    if (L.d.empty) continue;
    B.flops += (cp ? 8.0 : 2.0) * (double)L.d.M * (double)L.d.N * (double)L.d.K * (double)L.d.B;
    if (L.solo) continue;
    const TileShape& ts = kTileShapes[L.tile];
    // Split K for a pair of few tiles (at most kSplitMaxTiles) and a long K
    // (at least kSplitMinKTiles K-tiles): slices of at least 8 K-tiles, at
    // most 16 of them.  Such tiles would otherwise set a small launch's
    // length.  The rule reads the pair alone, so a pair's result does not
    // depend on what else is in the launch.  It takes precedence over the
    // skinny kernel, whose items run all of K in one workgroup.
    const int64_t kt = (L.d.K + 15) / 16;
    const bool split_k = B.allow_split && L.d.K > kShallowK && kt >= kSplitMinKTiles && tiles_of(L) <= kSplitMaxTiles;
    // a pair with an output side of at most kSkinnyMaxS: items of the whole
    // short side by skinny_geom's run of the long side, per batch index
    const bool sk = !split_k && skinny_pair(L.d.M, L.d.N, L.d.K) && tiles_of(L) >= kSkinnyMinTiles;
    const SkinnyGeom sg = skinny_geom(std::min(L.d.M, L.d.N), sko);
    const int64_t items = sk ? (std::max(L.d.M, L.d.N) + sg.lt - 1) / sg.lt * L.d.B : tiles_of(L);
    if (items >= (int64_t(1) << 31)) fail(MUSCLE_HIP_UNSUPPORTED, "too many tiles in one pair");
    const bool sh = !sk && L.d.K <= kShallowK;
    int64_t ks = 1, kc = L.d.K;
    if (split_k) {
      ks = std::min<int64_t>(16, kt / 8);
      kc = (kt + ks - 1) / ks * 16;
      ks = (L.d.K + kc - 1) / kc;
      const size_t pbytes = (size_t)items * (size_t)ks * (size_t)std::min<int64_t>(L.d.M, ts.bm) *
                            (size_t)std::min<int64_t>(L.d.N, ts.bn) * es;
      B.partial_off[(size_t)i] = (int64_t)B.partial_bytes;
      B.partial_bytes += (pbytes + 255) & ~size_t(255);
      B.tilecnt_off[(size_t)i] = B.ntilecnt;
      B.ntilecnt += items;
      B.split_items += items * ks;
    }
    B.ksplit[(size_t)i] = (int32_t)ks;
    B.kchunk[(size_t)i] = (int32_t)kc;
    B.pinfo[(size_t)i].ksplit = (int32_t)ks;
    if (sk) {
      B.pinfo[(size_t)i].route = MUSCLE_HIP_PAIR_SKINNY;
      B.pinfo[(size_t)i].tile = -1;
    } else if (sh) {
      B.pinfo[(size_t)i].route = MUSCLE_HIP_PAIR_SHALLOW;
      B.pinfo[(size_t)i].tile = -1;
    } else if (ks > 1) {
      B.pinfo[(size_t)i].route = MUSCLE_HIP_PAIR_SPLIT_K;
    }
    if (!sk && !sh && ks == 1) {
      // persistent tiles: kXcds contiguous runs, one per XCD queue; a pair of
      // fewer tiles goes whole to the queue with the least work so far (and
      // all of them, in order, to the single queue of a small launch)
      for (int64_t t = 0; t < items; ++t) single.push_back(make_int2((int)i, (int)t));
      const double cost = (double)L.cost;
      if (items >= kXcds) {
        for (int x = 0; x < kXcds; ++x) {
          const int64_t t0 = items * x / kXcds, t1 = items * (x + 1) / kXcds;
          for (int64_t t = t0; t < t1; ++t) xq[x].push_back(make_int2((int)i, (int)t));
          xload[x] += cost * (double)(t1 - t0);
        }
      } else {
        const int x = (int)(std::min_element(xload, xload + kXcds) - xload);
        for (int64_t t = 0; t < items; ++t) xq[x].push_back(make_int2((int)i, (int)t));
        xload[x] += cost * (double)items;
      }
    } else if (sh || sk) {
      // the vector-FMA kernels are bound by writing C: their items go in the
      // order of their corners' addresses in C, so the tiles that run at the
      // same time write neighbouring memory (with C's modes in a random order,
      // the tile order along M and N would scatter them over C)
      std::vector<int2>& dst = sk ? skinny : shallow[(L.c_mfast ? 1 : 0) + (c_stream_of(L.d) ? 2 : 0)];
      const size_t first = dst.size();
      for (int64_t t = 0; t < items; ++t) dst.push_back(make_int2((int)i, (int)t));
      if (L.htab && items > 1) {
        const std::vector<int64_t>& h = *L.htab;
        std::vector<std::pair<int64_t, int32_t>> key((size_t)items);
        const int64_t tiles_mn = sk ? (std::max(L.d.M, L.d.N) + sg.lt - 1) / sg.lt
                                    : ((L.d.M + ts.bm - 1) / ts.bm) * ((L.d.N + ts.bn - 1) / ts.bn);
        const int64_t tiles_n = (L.d.N + ts.bn - 1) / ts.bn;
        for (int64_t t = 0; t < items; ++t) {
          const int64_t bz = t / tiles_mn, r = t - bz * tiles_mn;
          int64_t off = tab_at(h, L.t[8], bz);
          if (sk) {
            const int64_t l = r * sg.lt;
            off += L.d.M <= L.d.N ? tab_at(h, L.t[5], l) : tab_at(h, L.t[4], l);
          } else {
            const int64_t tm = r / tiles_n, tn = r - tm * tiles_n;
            off += tab_at(h, L.t[4], tm * ts.bm) + tab_at(h, L.t[5], tn * ts.bn);
          }
          key[(size_t)t] = std::make_pair(off < 0 ? -off : off, (int32_t)t);
        }
        std::stable_sort(key.begin(), key.end());
        for (int64_t t = 0; t < items; ++t) dst[first + (size_t)t].y = key[(size_t)t].second;
      }
    } else {
      for (int64_t t = 0; t < items * ks; ++t) sliced.push_back(make_int2((int)i, (int)t));
    }
    if (!sh && !sk) used[L.tile] = true;
  }
  size_t nq = 0;
  for (int x = 0; x < kXcds; ++x) nq += xq[x].size();
  if (nq >= (size_t(1) << 31)) fail(MUSCLE_HIP_UNSUPPORTED, "too many tiles in one batch");
  // per-XCD queues pay off when every workgroup has several items of its own;
  // in a small launch the workgroups would mostly scan empty queues
  B.nq = (int64_t)nq >= int64_t(kXcdQueueItems) * B.workgroups ? kXcds : 1;
  B.qoff[0] = 0;
  if (B.nq == 1) {
    B.hitems.insert(B.hitems.end(), single.begin(), single.end());
    for (int x = 0; x < kXcds; ++x) B.qoff[x + 1] = (int32_t)B.hitems.size();
  } else {
    for (int x = 0; x < kXcds; ++x) {
      B.hitems.insert(B.hitems.end(), xq[x].begin(), xq[x].end());
      B.qoff[x + 1] = (int32_t)B.hitems.size();
    }
  }
  B.nsplit = (int64_t)sliced.size();
  B.nshallow = (int64_t)(shallow[0].size() + shallow[1].size() + shallow[2].size() + shallow[3].size());
  B.hitems.insert(B.hitems.end(), sliced.begin(), sliced.end());
  for (int c = 0; c < 4; ++c) {
    B.nshallow_cm[c] = (int64_t)shallow[c].size();
    B.hitems.insert(B.hitems.end(), shallow[c].begin(), shallow[c].end());
  }
  B.nskinny = (int64_t)skinny.size();
  B.hitems.insert(B.hitems.end(), skinny.begin(), skinny.end());
  const bool dbl = is_double(B.dtype);
  for (int t = 0; t < MUSCLE_HIP_TILE_COUNT; ++t)
    if (used[t]) B.lds = std::max(B.lds, batch_lds_bytes(dbl, cp, t));
}

// Host phase: lowering, compact tables, tile choice, work list.
void plan_batch(Handle& h, int64_t n, const muscle_hip_tensor_t* cs, const muscle_hip_tensor_t* as,
                const muscle_hip_tensor_t* bs, const int32_t* tiles, bool keep_descs, Batch& B,
                std::vector<PlannedPair>& lw) {
  B.h = &h;
  if (n < 0) fail(MUSCLE_HIP_INVALID_ARGUMENT, "negative batch size");
  if (n > 0 && (!cs || !as || !bs)) fail(MUSCLE_HIP_INVALID_ARGUMENT, "null tensor arrays");
  lw.assign((size_t)n, PlannedPair());
  for (int64_t i = 0; i < n; ++i) {
    check_tensor(&cs[i], "C");
    check_tensor(&as[i], "A");
    check_tensor(&bs[i], "B");
    if (i == 0) B.dtype = cs[0].dtype;
    if (cs[i].dtype != B.dtype || as[i].dtype != B.dtype || bs[i].dtype != B.dtype)
      fail(MUSCLE_HIP_UNSUPPORTED, "batched contractions need one element type for every operand");
    PlannedPair& L = lw[(size_t)i];
    L.d = lower_binary(cs[i], as[i], bs[i]);
    L.solo = keep_descs && !L.d.empty &&
             (double)L.d.M * (double)L.d.N * (double)L.d.K * (double)L.d.B >= kBatchSoloPoints;
    if (!L.solo) plan_pair(h, L, B.htab, tiles ? tiles[i] : MUSCLE_HIP_TILE_AUTO);
    L.a = L.d.swapped ? bs[i].data : as[i].data;
    L.b = L.d.swapped ? as[i].data : bs[i].data;
    L.c = cs[i].data;
    if (keep_descs) {
      copy_desc(B, cs[i], B.cs);
      copy_desc(B, as[i], B.as);
      copy_desc(B, bs[i], B.bs);
    }
    for (const muscle_hip_tensor_t* t : {&cs[i], &as[i], &bs[i]}) {
      double e = 1;
      for (int r = 0; r < t->rank; ++r) e *= (double)t->extents[r];
      B.bytes += e * (double)dtype_size(t->dtype);
    }
  }
  // point the stored descriptors at the owned copies
  for (size_t i = 0; i < B.cs.size(); ++i) {
    muscle_hip_tensor_t* ts[3] = {&B.cs[i], &B.as[i], &B.bs[i]};
    for (int j = 0; j < 3; ++j) {
      ts[j]->modes = B.modes[3 * i + j].data();
      ts[j]->extents = B.ext[3 * i + j].data();
      ts[j]->strides = B.str[3 * i + j].data();
    }
  }
  std::vector<const PlannedPair*> ptrs((size_t)n);
  for (int64_t i = 0; i < n; ++i) ptrs[(size_t)i] = &lw[(size_t)i];
  finish_batch(h, B, ptrs);
}

BatchLayout batch_layout(const Batch& B, int64_t n) {
  BatchLayout l;
  l.pairs = 0;
  l.items = al(sizeof(BatchPair) * (size_t)std::max<int64_t>(n, 1));
  l.counter = l.items + al(sizeof(int2) * std::max<size_t>(B.hitems.size(), 1));
  l.tab = l.counter + 256;
  l.copy = l.tab + al(std::max<size_t>(B.htab.size(), 1) * sizeof(int64_t));
  l.tilecnt = l.copy;
  l.partial = l.tilecnt + al(sizeof(int32_t) * (size_t)std::max<int64_t>(B.ntilecnt, 1));
  l.total = l.partial + B.partial_bytes;
  return l;
}

// Device phase: descriptors with device table pointers, copied to `base`
// (synchronously, or through `staging` pinned memory on stream s).
void place_batch(Batch& B, const std::vector<const PlannedPair*>& lw, char* base, char* staging, hipStream_t s) {
  const int64_t n = (int64_t)lw.size();
  const BatchLayout l = batch_layout(B, n);
  B.dpairs = (BatchPair*)(base + l.pairs);
  B.ditems = (int2*)(base + l.items);
  B.dcounter = (int32_t*)(base + l.counter);
  B.dtilecnt = (int32_t*)(base + l.tilecnt);
  const int64_t* dtab = (const int64_t*)(base + l.tab);
  B.hpairs.resize((size_t)n);
  for (int64_t i = 0; i < n; ++i) {
    const PlannedPair& L = *lw[(size_t)i];
    BatchPair& P = B.hpairs[(size_t)i];
    std::memset(&P, 0, sizeof P);
    if (L.solo) continue;  // no tables, never in the work list
    P.A = L.a;
    P.B = L.b;
    P.C = L.c;
    P.M = (int32_t)L.d.M;
    P.N = (int32_t)L.d.N;
    P.K = (int32_t)L.d.K;
    P.NB = (int32_t)L.d.B;
    const TileShape& ts = kTileShapes[L.tile];
    P.tiles_m = (int32_t)((L.d.M + ts.bm - 1) / ts.bm);
    P.tiles_n = (int32_t)((L.d.N + ts.bn - 1) / ts.bn);
    P.tile = L.tile;
    P.a_kmajor = L.a_kmajor;
    P.b_kmajor = L.b_kmajor;
    P.c_mfast = L.c_mfast;
    P.c_stream = c_stream_of(L.d);
    P.ksplit = B.ksplit.empty() ? 1 : B.ksplit[(size_t)i];
    P.kchunk = B.kchunk.empty() || B.kchunk[(size_t)i] == 0 ? P.K : B.kchunk[(size_t)i];
    P.partial = base + l.partial + (B.partial_off.empty() ? 0 : B.partial_off[(size_t)i]);
    P.tilecnt = (int32_t*)(base + l.tilecnt) + (B.tilecnt_off.empty() ? 0 : B.tilecnt_off[(size_t)i]);
    OffTab* tabs[9] = {&P.rowA, &P.colB, &P.kA, &P.kB, &P.rowC, &P.colC, &P.bA, &P.bB, &P.bC};
    // tables were appended in the order rowA, kA, kB, colB, rowC, colC, bA, bB, bC
    const int map[9] = {0, 3, 1, 2, 4, 5, 6, 7, 8};
    const int64_t* tb = L.dtab ? L.dtab : dtab;
    for (int t = 0; t < 9; ++t) *tabs[t] = make_tab(tb, L.t[map[t]]);
  }
  const size_t np = sizeof(BatchPair) * (size_t)n, ni = sizeof(int2) * B.hitems.size(),
               nt = B.htab.size() * sizeof(int64_t);
  if (staging) {
    std::memcpy(staging + l.pairs, B.hpairs.data(), np);
    if (ni) std::memcpy(staging + l.items, B.hitems.data(), ni);
    if (nt) std::memcpy(staging + l.tab, B.htab.data(), nt);
    MH_HIP_CHECK(hipMemcpyAsync(base, staging, l.copy, hipMemcpyHostToDevice, s));
  } else {
    if (np) MH_HIP_CHECK(hipMemcpy(base + l.pairs, B.hpairs.data(), np, hipMemcpyHostToDevice));
    if (ni) MH_HIP_CHECK(hipMemcpy(base + l.items, B.hitems.data(), ni, hipMemcpyHostToDevice));
    if (nt) MH_HIP_CHECK(hipMemcpy(base + l.tab, B.htab.data(), nt, hipMemcpyHostToDevice));
  }
}

std::unique_ptr<Batch> build_batch(Handle& h, int64_t n, const muscle_hip_tensor_t* cs, const muscle_hip_tensor_t* as,
                                   const muscle_hip_tensor_t* bs, const int32_t* tiles, bool allow_split) {
  std::unique_ptr<Batch> B(new Batch());
  B->allow_split = allow_split;
  B->complex_mult = h.complex_mult;
  std::vector<PlannedPair> lw;
  plan_batch(h, n, cs, as, bs, tiles, true, *B, lw);
  const BatchLayout l = batch_layout(*B, n);
  if (hipMalloc(&B->dmem, l.total) != hipSuccess) {
    (void)hipGetLastError();
    B->dmem = nullptr;
    fail(MUSCLE_HIP_OUT_OF_MEMORY, "batch descriptor allocation failed");
  }
  std::vector<const PlannedPair*> ptrs((size_t)n);
  for (int64_t i = 0; i < n; ++i) ptrs[(size_t)i] = &lw[(size_t)i];
  place_batch(*B, ptrs, (char*)B->dmem, nullptr, nullptr);
  // the largest pairs: plans of their own and one workspace for all of them
  B->solo.assign((size_t)n, 0);
  for (int64_t i = 0; i < n; ++i) {
    if (!lw[(size_t)i].solo) continue;
    B->solo[(size_t)i] = 1;
    B->solo_pairs.push_back(i);
    B->solo_plans.push_back(make_plan(h, 1, MUSCLE_HIP_PRECISION_NATIVE, 1e-12, &B->cs[(size_t)i], &B->as[(size_t)i],
                                      &B->bs[(size_t)i]));
    B->solo_ws_size = std::max(B->solo_ws_size, contraction_workspace_bytes(*B->solo_plans.back(), B->as[(size_t)i],
                                                                            &B->bs[(size_t)i]));
  }
  if (B->solo_ws_size) {
    if (hipMalloc(&B->solo_ws, B->solo_ws_size) != hipSuccess) {
      (void)hipGetLastError();
      B->solo_ws = nullptr;
      fail(MUSCLE_HIP_OUT_OF_MEMORY, "batch workspace allocation failed");
    }
  }
  return B;
}

void batch_launch(Batch& b, hipStream_t s) {
  const int64_t nmain = (int64_t)b.hitems.size() - b.nsplit - b.nshallow - b.nskinny;
  if (nmain > 0 || b.nsplit > 0) {
    BatchArgs a;
    a.pairs = b.dpairs;
    a.items = b.ditems;
    a.counter = b.dcounter;
    for (int x = 0; x <= kXcds; ++x) a.qoff[x] = b.qoff[x];
    a.nq = b.nq;
    a.tilecnt = b.dtilecnt;
    a.ntilecnt = b.ntilecnt;
    a.nitems = nmain;
    a.complex3m = b.complex_mult == 3;
    size_t lds = b.lds;
    a.item_off = (int32_t)batch_item_offset(is_double(b.dtype), is_complex(b.dtype), lds);
    BatchArgs sp = a;
    sp.items = b.ditems + nmain;
    sp.nitems = b.nsplit;
    launch_batch(is_double(b.dtype), is_complex(b.dtype), a, sp, b.workgroups, lds, s);
  }
  if (b.nshallow > 0) {
    // one launch per C orientation and store kind
    int64_t off = nmain + b.nsplit;
    for (int c = 0; c < 4; ++c) {
      if (b.nshallow_cm[c] > 0) {
        ShallowArgs a;
        a.pairs = b.dpairs;
        a.items = b.ditems + off;
        a.nitems = b.nshallow_cm[c];
        launch_shallow(is_double(b.dtype), is_complex(b.dtype), (c & 1) != 0, c >= 2, a, s);
      }
      off += b.nshallow_cm[c];
    }
  }
  if (b.nskinny > 0) {
    ShallowArgs a;
    a.pairs = b.dpairs;
    a.items = b.ditems + nmain + b.nsplit + b.nshallow;
    a.nitems = b.nskinny;
    launch_skinny(is_double(b.dtype), is_complex(b.dtype), a, s);
  }
}

namespace {
// Sets the handle's complex product form for the lifetime of the object
// (the batch's own form for its plain-loop and solo launches).
struct ProductForm {
  Handle& h;
  int saved;
  ProductForm(Handle& h_, int form) : h(h_), saved(h_.complex_mult) { h.complex_mult = form; }
  ~ProductForm() { h.complex_mult = saved; }
};
}  // namespace

void batch_run(Batch& b, int mode) {
  Handle& h = *b.h;
  ProductForm form(h, b.complex_mult);
  if (mode == 1) {
    for (size_t i = 0; i < b.cs.size(); ++i) {
      Plan* p = get_plan(h, 1, MUSCLE_HIP_PRECISION_NATIVE, 1e-12, &b.cs[i], &b.as[i], &b.bs[i]);
      run_contraction(h, *p, b.cs[i], b.as[i], &b.bs[i]);
    }
    return;
  }
  if (b.hitems.empty() && b.solo_plans.empty()) return;
  // the persistent launch, then the solo pairs' launches, on stream s (the
  // solo launches take the handle's stream for the duration)
  auto enqueue = [&](hipStream_t s) {
    if (!b.hitems.empty()) {
      batch_launch(b, s);
      MH_HIP_CHECK(hipGetLastError());
    }
    const hipStream_t saved = h.stream;
    h.stream = s;
    try {
      for (size_t k = 0; k < b.solo_plans.size(); ++k) {
        const size_t i = (size_t)b.solo_pairs[k];
        run_contraction(h, *b.solo_plans[k], b.cs[i], b.as[i], &b.bs[i], (char*)b.solo_ws, b.solo_ws_size);
      }
    } catch (...) {
      h.stream = saved;
      throw;
    }
    h.stream = saved;
  };
  if (mode == 2) {
    enqueue(h.stream);
    return;
  }
  if (!b.exec) {
    // capture on a private stream (the null stream cannot be captured), then
    // replay the graph on the handle's stream
    if (!b.capture_stream) MH_HIP_CHECK(hipStreamCreateWithFlags(&b.capture_stream, hipStreamNonBlocking));
    MH_HIP_CHECK(hipStreamBeginCapture(b.capture_stream, hipStreamCaptureModeThreadLocal));
    try {
      enqueue(b.capture_stream);
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(b.capture_stream, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(b.capture_stream, &g);
    if (e != hipSuccess) fail(MUSCLE_HIP_HIP_ERROR, std::string("graph capture: ") + hipGetErrorString(e));
    b.graph = g;
    MH_HIP_CHECK(hipGraphInstantiate(&b.exec, b.graph, nullptr, nullptr, 0));
  }
  MH_HIP_CHECK(hipGraphLaunch(b.exec, h.stream));
}

// Times every tile shape on up to `samples` pairs of each size class.
void batch_autotune(Batch& b, int samples, int reps) {
  Handle& h = *b.h;
  ProductForm form(h, b.complex_mult);
  std::vector<std::vector<int64_t>> by_class(Handle::kTuneClasses);
  for (size_t i = 0; i < b.hpairs.size(); ++i) {
    const BatchPair& P = b.hpairs[i];
    if (P.M * (int64_t)P.N * P.NB == 0 || (i < b.solo.size() && b.solo[i])) continue;
    if (b.pinfo[i].route == MUSCLE_HIP_PAIR_SKINNY) continue;  // no tile shape to choose
    by_class[tuning_class(P.M, P.N)].push_back((int64_t)i);
  }
  hipEvent_t e0, e1;
  MH_HIP_CHECK(hipEventCreate(&e0));
  MH_HIP_CHECK(hipEventCreate(&e1));
  for (int c = 0; c < Handle::kTuneClasses; ++c) {
    if (by_class[c].empty()) continue;
    std::vector<int64_t>& idx = by_class[c];
    if ((int)idx.size() > samples) idx.resize((size_t)samples);
    std::vector<muscle_hip_tensor_t> cs, as, bs;
    for (int64_t i : idx) {
      cs.push_back(b.cs[(size_t)i]);
      as.push_back(b.as[(size_t)i]);
      bs.push_back(b.bs[(size_t)i]);
    }
    float best = 1e30f;
    int best_t = default_tile(b.hpairs[(size_t)idx[0]].M, b.hpairs[(size_t)idx[0]].N);
    for (int t = 0; t < MUSCLE_HIP_TILE_COUNT; ++t) {
      std::vector<int32_t> tiles(idx.size(), t);
      // timed without K splits: a class's shape is chosen on its tiles alone,
      // as a sample of a few pairs would split where the whole batch need not
      std::unique_ptr<Batch> sub =
          build_batch(h, (int64_t)idx.size(), cs.data(), as.data(), bs.data(), tiles.data(), false);
      batch_run(*sub, 2);  // warm-up
      MH_HIP_CHECK(hipEventRecord(e0, h.stream));
      for (int r = 0; r < reps; ++r) batch_run(*sub, 2);
      MH_HIP_CHECK(hipEventRecord(e1, h.stream));
      MH_HIP_CHECK(hipEventSynchronize(e1));
      float ms = 0;
      MH_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
      if (ms < best) {
        best = ms;
        best_t = t;
      }
    }
    h.tune[c] = best_t;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
}

}  // namespace mh

using namespace mh;

namespace {
template <typename F>
int guarded_b(F&& f) {
  try {
    f();
    return MUSCLE_HIP_OK;
  } catch (const Error& e) {
    g_last_error = e.what();
    return e.status;
  } catch (const std::bad_alloc&) {
    g_last_error = "host allocation failed";
    return MUSCLE_HIP_OUT_OF_MEMORY;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return MUSCLE_HIP_INVALID_ARGUMENT;
  }
}
}  // namespace

extern "C" {

int muscle_hip_batch_create(muscle_hip_handle_t handle, int64_t n, const muscle_hip_tensor_t* cs,
                            const muscle_hip_tensor_t* as, const muscle_hip_tensor_t* bs, const int32_t* tiles,
                            muscle_hip_batch_t* batch) {
  return guarded_b([&] {
    Handle* h = reinterpret_cast<Handle*>(handle);
    if (!h || !batch) fail(MUSCLE_HIP_INVALID_ARGUMENT, "null handle or batch pointer");
    std::lock_guard<std::mutex> lk(h->mu);
    *batch = reinterpret_cast<muscle_hip_batch_t>(build_batch(*h, n, cs, as, bs, tiles).release());
  });
}

int muscle_hip_batch_run(muscle_hip_handle_t handle, muscle_hip_batch_t batch, int mode) {
  return guarded_b([&] {
    Handle* h = reinterpret_cast<Handle*>(handle);
    Batch* b = reinterpret_cast<Batch*>(batch);
    if (!h || !b) fail(MUSCLE_HIP_INVALID_ARGUMENT, "null handle or batch");
    if (b->h != h) fail(MUSCLE_HIP_INVALID_ARGUMENT, "batch belongs to another handle");
    if (mode < 0 || mode > 2) fail(MUSCLE_HIP_INVALID_ARGUMENT, "mode must be 0, 1 or 2");
    std::lock_guard<std::mutex> lk(h->mu);
    batch_run(*b, mode);
  });
}

int muscle_hip_batch_get_info(muscle_hip_batch_t batch, muscle_hip_batch_info_t* info) {
  return guarded_b([&] {
    Batch* b = reinterpret_cast<Batch*>(batch);
    if (!b || !info) fail(MUSCLE_HIP_INVALID_ARGUMENT, "null batch or info");
    info->pairs = (int64_t)b->cs.size();
    info->items = (int64_t)b->hitems.size();
    info->shallow_items = b->nshallow;
    info->skinny_items = b->nskinny;
    info->split_items = b->split_items;
    info->workgroups = b->workgroups;
    info->graph = b->exec ? 1 : 0;
    info->flops = b->flops;
    info->bytes = b->bytes;
    info->table_bytes = (int64_t)(b->htab.size() * 8 + b->hpairs.size() * sizeof(BatchPair) +
                                  b->hitems.size() * sizeof(int2));
  });
}

int muscle_hip_batch_get_pair_info(muscle_hip_batch_t batch, int64_t pair, muscle_hip_batch_pair_info_t* info) {
  return guarded_b([&] {
    Batch* b = reinterpret_cast<Batch*>(batch);
    if (!b || !info) fail(MUSCLE_HIP_INVALID_ARGUMENT, "null batch or info");
    if (pair < 0 || pair >= (int64_t)b->pinfo.size()) fail(MUSCLE_HIP_INVALID_ARGUMENT, "pair index out of range");
    *info = b->pinfo[(size_t)pair];
  });
}

int muscle_hip_batch_destroy(muscle_hip_batch_t batch) {
  return guarded_b([&] { delete reinterpret_cast<Batch*>(batch); });
}

int muscle_hip_batch_autotune(muscle_hip_handle_t handle, muscle_hip_batch_t batch, int samples, int reps) {
  return guarded_b([&] {
    Handle* h = reinterpret_cast<Handle*>(handle);
    Batch* b = reinterpret_cast<Batch*>(batch);
    if (!h || !b || b->h != h) fail(MUSCLE_HIP_INVALID_ARGUMENT, "null or foreign batch");
    std::lock_guard<std::mutex> lk(h->mu);
    batch_autotune(*b, std::max(1, samples), std::max(1, reps));
  });
}

int muscle_hip_tuning_save(muscle_hip_handle_t handle, const char* path) {
  return guarded_b([&] {
    Handle* h = reinterpret_cast<Handle*>(handle);
    if (!h || !path) fail(MUSCLE_HIP_INVALID_ARGUMENT, "null handle or path");
    std::ofstream f(path);
    if (!f) fail(MUSCLE_HIP_INVALID_ARGUMENT, std::string("cannot write ") + path);
    f << "# muscle_hip batched tile tuning: class tile (class = 5 * bucket(M) + bucket(N))\n";
    for (int c = 0; c < Handle::kTuneClasses; ++c)
      if (h->tune[c] >= 0) f << c << " " << h->tune[c] << "\n";
  });
}

int muscle_hip_tuning_load(muscle_hip_handle_t handle, const char* path) {
  return guarded_b([&] {
    Handle* h = reinterpret_cast<Handle*>(handle);
    if (!h || !path) fail(MUSCLE_HIP_INVALID_ARGUMENT, "null handle or path");
    std::ifstream f(path);
    if (!f) fail(MUSCLE_HIP_INVALID_ARGUMENT, std::string("cannot read ") + path);
    std::string line;
    while (std::getline(f, line)) {
      if (line.empty() || line[0] == '#') continue;
      int c = -1, t = -1;
      if (std::sscanf(line.c_str(), "%d %d", &c, &t) == 2 && c >= 0 && c < Handle::kTuneClasses && t >= 0 &&
          t < MUSCLE_HIP_TILE_COUNT)
        h->tune[c] = t;
    }
  });
}

int muscle_hip_tuning_class(int64_t m, int64_t n) { return batch_tuning_class(m, n); }

}  // extern "C"
