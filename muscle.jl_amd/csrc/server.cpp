// Contraction-tree serving with continuous batching.
//
// Jobs (trees of pairwise contractions) are compiled on planner threads as
// they arrive: every contraction is lowered once, and the batched ones get
// their offset tables and tile shape (one table arena per job).  A scheduler
// thread owns a private handle (stream, plan cache) and a device pool.
// Compiled jobs are admitted in arrival order while their footprint (all
// intermediates, the result and the table arena) fits next to the admitted
// jobs; admission uploads the job's tables.  Each tick launches every
// contraction of every admitted job whose operands have been launched
// (stream order makes them complete first): contractions big enough to fill
// the GPU alone take the single-contraction kernel, all others share one
// persistent batched launch per element type.  Up to kSlots ticks are in
// flight; when a tick's event fires, operands that have no consumer left go
// back to the pool and the owning job's reservation shrinks, so waiting jobs
// are admitted while others are still on the GPU.  Mode 1 runs the same jobs
// as the plain per-call loop (one job at a time, one launch per node).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "batch.hpp"
#include "handle.hpp"
#include "lowering.hpp"

namespace mh {
namespace {

size_t al256(size_t v) { return (v + 255) & ~size_t(255); }

// First-fit allocator over one device arena, with coalescing and a separate
// admission reservation counter.
class Pool {
 public:
  void init(char* base, size_t size) {
    base_ = base;
    size_ = size;
    free_.clear();
    free_[0] = size;
  }
  char* alloc(size_t bytes) {
    bytes = al256(std::max<size_t>(bytes, 1));
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= bytes) {
        const size_t off = it->first, left = it->second - bytes;
        free_.erase(it);
        if (left) free_[off + bytes] = left;
        used_ += bytes;
        peak_ = std::max(peak_, used_);
        return base_ + off;
      }
    }
    return nullptr;
  }
  void release(char* p, size_t bytes) {
    bytes = al256(std::max<size_t>(bytes, 1));
    size_t off = (size_t)(p - base_);
    used_ -= bytes;
    auto next = free_.lower_bound(off);
    if (next != free_.begin()) {
      auto prev = std::prev(next);
      if (prev->first + prev->second == off) {
        off = prev->first;
        bytes += prev->second;
        free_.erase(prev);
      }
    }
    next = free_.lower_bound(off);
    if (next != free_.end() && off + bytes == next->first) {
      bytes += next->second;
      free_.erase(next);
    }
    free_[off] = bytes;
  }
  size_t size() const { return size_; }
  size_t peak() const { return peak_; }
  size_t reserved = 0;

 private:
  char* base_ = nullptr;
  size_t size_ = 0, used_ = 0, peak_ = 0;
  std::map<size_t, size_t> free_;
};

struct Operand {
  muscle_hip_tensor_t t;
  std::vector<int32_t> modes;
  std::vector<int64_t> ext, str;
  void fix() {
    t.modes = modes.data();
    t.extents = ext.data();
    t.strides = str.data();
    t.rank = (int32_t)modes.size();
  }
};

struct Node {
  int lhs, rhs;
  Operand out;
  size_t bytes = 0;
  int consumers = 0;       // readers that have not finished (the root: its caller)
  std::vector<int> users;  // nodes that read this node's output
  int pending = 0;         // node operands not launched yet
  bool solo = false;       // runs as a single-contraction launch
  PlannedPair plan;        // batched nodes: lowering, tables, tile
  bool launched = false, done = false, freed = false;
};

enum CompileState { kQueued = 0, kCompiled = 1, kFailed = 2 };

struct Job {
  int64_t id = 0;
  int dtype = 0;
  std::vector<Operand> leaves;
  std::vector<Node> nodes;
  size_t reserve = 0;   // bytes of all intermediates, the result and the tables
  size_t reserved = 0;  // what is still reserved in the pool
  int status = MUSCLE_HIP_NOT_READY;
  std::string error;
  int compile = kQueued;
  std::vector<int64_t> htab;  // offset tables of the batched nodes
  char* dtab = nullptr;
  size_t dtab_bytes = 0;
  std::vector<int> ready;  // nodes whose operands have been launched
  double t_submit = 0, t_admit = 0, t_done = 0;
  int64_t ticks = 0;
  int remaining = 0;
  const Operand& operand(int k) const {
    const int nl = (int)leaves.size();
    return k < nl ? leaves[(size_t)k] : nodes[(size_t)(k - nl)].out;
  }
};

struct TickEntry {
  Job* job;
  int node;
};

struct Slot {
  char* dev = nullptr;
  char* pinned = nullptr;
  size_t size = 0;
  hipEvent_t ev = nullptr;
  std::vector<TickEntry> entries;
  std::vector<std::unique_ptr<Batch>> batches;
  std::vector<void*> extra_dev;  // oversized tick descriptors
};

constexpr int kSlots = 3;
constexpr size_t kSlotBytes = size_t(64) << 20;
// index points (M*N*K*batch) from which a contraction runs on its own
constexpr double kSoloPoints = double(int64_t(1) << 29);
constexpr int kMaxPlanners = 8;

}  // namespace

struct Server {
  Handle h;
  int mode = 0;
  int max_active = 64;
  std::mutex mu;
  std::condition_variable cv_work, cv_done, cv_plan;
  std::deque<std::shared_ptr<Job>> waiting;   // arrival order, compiled or not
  std::deque<std::shared_ptr<Job>> to_compile;
  std::vector<std::shared_ptr<Job>> active;
  std::unordered_map<int64_t, std::shared_ptr<Job>> jobs;
  Pool pool;
  void* arena = nullptr;
  Slot slots[kSlots];
  int slot_head = 0, slot_count = 0;  // in-flight ring
  std::thread worker;
  std::vector<std::thread> planners;
  bool stop = false;
  int64_t next_id = 1;
  muscle_hip_server_stats_t st{};
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  std::string worker_error;

  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }

  void loop();
  void plan_loop();
  void compile(Job& j);
  std::vector<std::shared_ptr<Job>> admit_locked();
  bool launch_tick();
  bool complete_ticks(bool block);
  void finish_node(Job& j, int n);
  void finish_job(Job& j, int status);
  void run_serial(Job& j);
};

namespace {

void column_major(Operand& o) {
  o.str.resize(o.ext.size());
  int64_t s = 1;
  for (size_t i = 0; i < o.ext.size(); ++i) {
    o.str[i] = s;
    s *= std::max<int64_t>(o.ext[i], 1);
  }
}

int64_t elems(const Operand& o) {
  int64_t n = 1;
  for (int64_t e : o.ext) n *= e;
  return n;
}

}  // namespace

// Lowers every contraction of a job; the batched ones get tables and a tile.
void Server::compile(Job& j) {
  for (Node& nd : j.nodes) {
    muscle_hip_tensor_t c = nd.out.t;
    c.data = (void*)256;  // addresses are bound at launch; the lowering reads shapes and strides only
    nd.plan.d = lower_binary(c, j.operand(nd.lhs).t, j.operand(nd.rhs).t);
    const ContractionDesc& d = nd.plan.d;
    nd.solo = (double)d.M * (double)d.N * (double)d.K * (double)d.B >= kSoloPoints;
    if (!nd.solo) plan_pair(h, nd.plan, j.htab, MUSCLE_HIP_TILE_AUTO, kTickSingleMax);
  }
  j.dtab_bytes = j.htab.empty() ? 0 : al256(j.htab.size() * sizeof(int64_t));
}

void Server::plan_loop() {
  for (;;) {
    std::shared_ptr<Job> j;
    {
      std::unique_lock<std::mutex> lk(mu);
      cv_plan.wait(lk, [&] { return stop || !to_compile.empty(); });
      if (to_compile.empty()) return;  // stopping
      j = to_compile.front();
      to_compile.pop_front();
    }
    int state = kCompiled, status = MUSCLE_HIP_OK;
    std::string msg;
    try {
      compile(*j);
    } catch (const Error& e) {
      state = kFailed;
      status = e.status;
      msg = e.what();
    } catch (const std::exception& e) {
      state = kFailed;
      status = MUSCLE_HIP_INVALID_ARGUMENT;
      msg = e.what();
    }
    {
      std::lock_guard<std::mutex> lk(mu);
      j->compile = state;
      if (state == kFailed) {
        j->status = status;
        j->error = msg;
      } else {
        j->reserve += j->dtab_bytes;
      }
    }
    cv_work.notify_one();
  }
}

// Job j ends with status; its table arena goes back to the pool.
void Server::finish_job(Job& j, int status) {
  j.status = status;
  j.t_done = now();
  if (j.dtab) {
    pool.release(j.dtab, j.dtab_bytes);
    j.dtab = nullptr;
    j.reserved -= j.dtab_bytes;
    pool.reserved -= j.dtab_bytes;
  }
  if (status == MUSCLE_HIP_OK) st.completed++;
}

// Node n of j finished: release operands that have no consumer left.
void Server::finish_node(Job& j, int n) {
  Node& nd = j.nodes[(size_t)n];
  nd.done = true;
  const int nl = (int)j.leaves.size();
  for (int k : {nd.lhs, nd.rhs}) {
    if (k >= nl) {
      Node& src = j.nodes[(size_t)(k - nl)];
      if (--src.consumers == 0 && !src.freed) {
        pool.release((char*)src.out.t.data, src.bytes);
        src.freed = true;
        const size_t r = al256(src.bytes);
        j.reserved -= r;
        pool.reserved -= r;
      }
    }
    if (nd.lhs == nd.rhs) break;
  }
  if (--j.remaining == 0) finish_job(j, MUSCLE_HIP_OK);
}

// Admits compiled jobs in arrival order while their footprint fits; returns
// the admitted jobs (their tables are uploaded by the caller).
std::vector<std::shared_ptr<Job>> Server::admit_locked() {
  std::vector<std::shared_ptr<Job>> out;
  const int cap = mode == 1 ? 1 : max_active;  // the plain loop runs one job at a time
  while (!waiting.empty() && (int)active.size() < cap) {
    std::shared_ptr<Job> j = waiting.front();
    if (j->compile == kQueued) break;
    if (j->compile == kFailed || j->reserve > pool.size()) {  // a contraction the library cannot run
      if (j->compile != kFailed) {
        j->error = "job and its tables need more device memory than the server pool holds";
        st.rejected++;
      }
      waiting.pop_front();
      j->t_admit = now();
      finish_job(*j, j->compile == kFailed ? j->status : MUSCLE_HIP_REJECTED);
      cv_done.notify_all();
      continue;
    }
    if (pool.reserved + j->reserve > pool.size()) break;
    if (j->dtab_bytes) {
      j->dtab = pool.alloc(j->dtab_bytes);
      if (!j->dtab) break;  // fragmentation: retry after frees
    }
    pool.reserved += j->reserve;
    j->reserved = j->reserve;
    j->t_admit = now();
    for (size_t n = 0; n < j->nodes.size(); ++n)
      if (j->nodes[n].pending == 0) j->ready.push_back((int)n);
    active.push_back(j);
    waiting.pop_front();
    out.push_back(j);
  }
  return out;
}

// Launches every ready contraction of the admitted jobs as one tick.
bool Server::launch_tick() {
  if (slot_count == kSlots) return false;
  std::vector<TickEntry> entries;
  {
    std::lock_guard<std::mutex> lk(mu);  // the pool and job state are shared with submit/release
    for (auto& jp : active) {
      Job& j = *jp;
      if (j.ready.empty()) continue;
      std::vector<int> cand;
      cand.swap(j.ready);
      const size_t first = entries.size();
      for (int n : cand) {
        Node& nd = j.nodes[(size_t)n];
        char* p = pool.alloc(nd.bytes);
        if (!p) {  // fragmentation: retry after frees
          j.ready.push_back(n);
          continue;
        }
        nd.out.t.data = p;
        entries.push_back({&j, n});
      }
      // consumers become ready for the next tick (not this one: a batched
      // launch has no order between its contractions)
      for (size_t e = first; e < entries.size(); ++e) {
        Node& nd = j.nodes[(size_t)entries[e].node];
        nd.launched = true;
        for (int u : nd.users)
          if (--j.nodes[(size_t)u].pending == 0) j.ready.push_back(u);
      }
      if (entries.size() > first) j.ticks++;
    }
  }
  if (entries.empty()) return false;
  Slot& sl = slots[(slot_head + slot_count) % kSlots];
  sl.entries = entries;
  sl.batches.clear();
  size_t used = 0;
  std::vector<const PlannedPair*> lw[MUSCLE_HIP_C64 + 1];
  for (const TickEntry& e : entries) {
    Job& j = *e.job;
    Node& nd = j.nodes[(size_t)e.node];
    const muscle_hip_tensor_t& a = j.operand(nd.lhs).t;
    const muscle_hip_tensor_t& b = j.operand(nd.rhs).t;
    if (nd.solo) {
      Plan* pl = get_plan(h, 1, MUSCLE_HIP_PRECISION_NATIVE, 1e-12, &nd.out.t, &a, &b);
      run_contraction(h, *pl, nd.out.t, a, &b);
      st.launched_pairs++;
      continue;
    }
    PlannedPair& L = nd.plan;
    L.a = L.d.swapped ? b.data : a.data;
    L.b = L.d.swapped ? a.data : b.data;
    L.c = nd.out.t.data;
    L.dtab = (const int64_t*)j.dtab;
    lw[j.dtype].push_back(&L);
  }
  for (int dt = MUSCLE_HIP_F32; dt <= MUSCLE_HIP_C64; ++dt) {
    if (lw[dt].empty()) continue;
    std::unique_ptr<Batch> B(new Batch());
    B->h = &h;
    B->dtype = dt;
    finish_batch(h, *B, lw[dt]);
    const BatchLayout l = batch_layout(*B, (int64_t)lw[dt].size());
    if (used + l.total <= sl.size) {
      place_batch(*B, lw[dt], sl.dev + used, sl.pinned + used, h.stream);
      used += al256(l.total);
    } else {
      void* d = nullptr;
      if (hipMalloc(&d, l.total) != hipSuccess) fail(MUSCLE_HIP_OUT_OF_MEMORY, "tick descriptors");
      sl.extra_dev.push_back(d);
      place_batch(*B, lw[dt], (char*)d, nullptr, h.stream);
    }
    if (!B->hitems.empty()) batch_launch(*B, h.stream);
    MH_HIP_CHECK(hipGetLastError());
    B->h = nullptr;  // completion is tracked by the slot event
    st.launched_pairs += (int64_t)lw[dt].size();
    sl.batches.push_back(std::move(B));
  }
  MH_HIP_CHECK(hipEventRecord(sl.ev, h.stream));
  slot_count++;
  st.ticks++;
  return true;
}

// Retires finished ticks in order; `block` waits for the oldest.
bool Server::complete_ticks(bool block) {
  bool any = false;
  while (slot_count > 0) {
    Slot& sl = slots[slot_head];
    hipError_t q = block && !any ? hipEventSynchronize(sl.ev) : hipEventQuery(sl.ev);
    if (q == hipErrorNotReady) break;
    if (q != hipSuccess) fail(MUSCLE_HIP_HIP_ERROR, std::string("tick: ") + hipGetErrorString(q));
    {
      std::lock_guard<std::mutex> lk(mu);
      for (auto& e : sl.entries) finish_node(*e.job, e.node);
      active.erase(std::remove_if(active.begin(), active.end(),
                                  [](const std::shared_ptr<Job>& j) { return j->status != MUSCLE_HIP_NOT_READY; }),
                   active.end());
    }
    cv_done.notify_all();
    for (void* d : sl.extra_dev) (void)hipFree(d);
    sl.extra_dev.clear();
    sl.batches.clear();
    sl.entries.clear();
    slot_head = (slot_head + 1) % kSlots;
    slot_count--;
    any = true;
  }
  return any;
}

// The plain per-call loop: one launch per node, one job at a time.
void Server::run_serial(Job& j) {
  for (size_t n = 0; n < j.nodes.size(); ++n) {
    Node& nd = j.nodes[n];
    char* p;
    {
      std::lock_guard<std::mutex> lk(mu);
      p = pool.alloc(nd.bytes);
    }
    if (!p) fail(MUSCLE_HIP_OUT_OF_MEMORY, "serial mode: pool exhausted");
    nd.out.t.data = p;
    const muscle_hip_tensor_t& a = j.operand(nd.lhs).t;
    const muscle_hip_tensor_t& b = j.operand(nd.rhs).t;
    Plan* pl = get_plan(h, 1, MUSCLE_HIP_PRECISION_NATIVE, 1e-12, &nd.out.t, &a, &b);
    run_contraction(h, *pl, nd.out.t, a, &b);
    nd.launched = true;
    st.ticks++;
    st.launched_pairs++;
    j.ticks++;
  }
  MH_HIP_CHECK(hipStreamSynchronize(h.stream));
  std::lock_guard<std::mutex> lk(mu);
  for (size_t n = 0; n < j.nodes.size(); ++n) finish_node(j, (int)n);
  active.erase(std::remove_if(active.begin(), active.end(),
                              [](const std::shared_ptr<Job>& x) { return x->status != MUSCLE_HIP_NOT_READY; }),
               active.end());
}

void Server::loop() {
  (void)hipSetDevice(h.device);
  try {
    for (;;) {
      bool progress = false;
      std::vector<std::shared_ptr<Job>> admitted;
      {
        std::unique_lock<std::mutex> lk(mu);
        if (stop && active.empty() && slot_count == 0) {
          for (auto& j : waiting) {  // never admitted: ended by the shutdown
            j->error = "server destroyed before the job was admitted";
            j->status = MUSCLE_HIP_HIP_ERROR;
          }
          waiting.clear();
          cv_done.notify_all();
          break;
        }
        if (!stop && (mode == 0 || active.empty())) admitted = admit_locked();
        progress = !admitted.empty();
        if (!progress && active.empty() && slot_count == 0) {
          // nothing admitted and nothing running: wait for a submission or a compiled job
          cv_work.wait_for(lk, std::chrono::milliseconds(20));
          continue;
        }
      }
      for (auto& j : admitted)
        if (j->dtab_bytes)
          MH_HIP_CHECK(hipMemcpyAsync(j->dtab, j->htab.data(), j->htab.size() * sizeof(int64_t),
                                      hipMemcpyHostToDevice, h.stream));
      if (mode == 1) {
        std::shared_ptr<Job> j;
        {
          std::lock_guard<std::mutex> lk(mu);
          if (!active.empty()) j = active.front();
        }
        if (j) run_serial(*j);
        cv_done.notify_all();
        continue;
      }
      progress |= complete_ticks(false);
      progress |= launch_tick();
      if (!progress) {
        if (slot_count > 0) {
          complete_ticks(true);  // nothing to launch: wait for the GPU
        } else {
          std::unique_lock<std::mutex> lk(mu);  // admitted jobs wait for memory: wait for a release
          cv_work.wait_for(lk, std::chrono::microseconds(200));
        }
      }
    }
  } catch (const std::exception& e) {
    std::lock_guard<std::mutex> lk(mu);
    worker_error = e.what();
    for (auto& j : active) {
      j->status = MUSCLE_HIP_HIP_ERROR;
      j->error = worker_error;
    }
    for (auto& j : waiting) {
      j->status = MUSCLE_HIP_HIP_ERROR;
      j->error = worker_error;
    }
    active.clear();
    waiting.clear();
    cv_done.notify_all();
  }
}

}  // namespace mh

using namespace mh;

namespace {
template <typename F>
int guarded_s(F&& f) {
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

void copy_operand(Operand& o, const muscle_hip_tensor_t& t) {
  o.t = t;
  o.modes.assign(t.modes, t.modes + t.rank);
  o.ext.assign(t.extents, t.extents + t.rank);
  o.str.assign(t.strides, t.strides + t.rank);
  o.fix();
}
}  // namespace

extern "C" {

int muscle_hip_server_create(int64_t pool_bytes, int max_active_jobs, int mode, muscle_hip_server_t* server) {
  return guarded_s([&] {
    if (!server) fail(MUSCLE_HIP_INVALID_ARGUMENT, "null server pointer");
    if (pool_bytes <= 0 || mode < 0 || mode > 1) fail(MUSCLE_HIP_INVALID_ARGUMENT, "bad pool size or mode");
    std::unique_ptr<Server> S(new Server());
    MH_HIP_CHECK(hipGetDevice(&S->h.device));
    MH_HIP_CHECK(hipStreamCreateWithFlags(&S->h.stream, hipStreamNonBlocking));
    S->mode = mode;
    S->max_active = std::max(1, max_active_jobs);
    if (hipMalloc(&S->arena, (size_t)pool_bytes) != hipSuccess) {
      (void)hipGetLastError();
      S->arena = nullptr;
      fail(MUSCLE_HIP_OUT_OF_MEMORY, "server pool allocation failed");
    }
    S->pool.init((char*)S->arena, (size_t)pool_bytes);
    for (Slot& sl : S->slots) {
      MH_HIP_CHECK(hipMalloc(&sl.dev, kSlotBytes));
      MH_HIP_CHECK(hipHostMalloc((void**)&sl.pinned, kSlotBytes, hipHostMallocDefault));
      MH_HIP_CHECK(hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming));
      sl.size = kSlotBytes;
    }
    S->st.pool_bytes = pool_bytes;
    Server* raw = S.get();
    S->worker = std::thread([raw] { raw->loop(); });
    const int np = std::max(1, std::min<int>(kMaxPlanners, (int)std::thread::hardware_concurrency() / 2));
    for (int i = 0; i < np; ++i) S->planners.emplace_back([raw] { raw->plan_loop(); });
    *server = reinterpret_cast<muscle_hip_server_t>(S.release());
  });
}

int muscle_hip_server_submit(muscle_hip_server_t server, const muscle_hip_job_t* job, int64_t* job_id) {
  return guarded_s([&] {
    Server* S = reinterpret_cast<Server*>(server);
    if (!S || !job || !job_id) fail(MUSCLE_HIP_INVALID_ARGUMENT, "null argument");
    if (job->n_leaves < 1 || job->n_nodes < 1 || !job->leaves || !job->lhs || !job->rhs || !job->node_off ||
        !job->node_modes || !job->label_extents)
      fail(MUSCLE_HIP_INVALID_ARGUMENT, "incomplete job description");
    std::shared_ptr<Job> j(new Job());
    j->leaves.resize((size_t)job->n_leaves);
    for (int i = 0; i < job->n_leaves; ++i) {
      check_tensor(&job->leaves[i], "leaf");
      copy_operand(j->leaves[(size_t)i], job->leaves[i]);
    }
    j->dtype = job->leaves[0].dtype;
    for (int i = 1; i < job->n_leaves; ++i)
      if (job->leaves[i].dtype != j->dtype) fail(MUSCLE_HIP_UNSUPPORTED, "a job has one element type");
    const int total = job->n_leaves + job->n_nodes;
    j->nodes.resize((size_t)job->n_nodes);
    const size_t es = dtype_size(j->dtype);
    for (int i = 0; i < job->n_nodes; ++i) {
      Node& nd = j->nodes[(size_t)i];
      nd.lhs = job->lhs[i];
      nd.rhs = job->rhs[i];
      if (nd.lhs < 0 || nd.rhs < 0 || nd.lhs >= job->n_leaves + i || nd.rhs >= job->n_leaves + i)
        fail(MUSCLE_HIP_INVALID_ARGUMENT, "node operands must precede the node");
      const int o0 = job->node_off[i], o1 = job->node_off[i + 1];
      if (o0 < 0 || o1 < o0) fail(MUSCLE_HIP_INVALID_ARGUMENT, "bad node_off");
      nd.out.modes.assign(job->node_modes + o0, job->node_modes + o1);
      for (int32_t m : nd.out.modes) {
        if (m < 0 || m >= job->n_labels) fail(MUSCLE_HIP_INVALID_ARGUMENT, "mode label out of range");
        nd.out.ext.push_back(job->label_extents[m]);
      }
      column_major(nd.out);
      nd.out.t.dtype = j->dtype;
      nd.out.t.data = nullptr;
      nd.out.fix();
      nd.bytes = (size_t)elems(nd.out) * es;
      j->reserve += al256(std::max<size_t>(nd.bytes, 1));
      (void)total;
    }
    for (int i = 0; i < job->n_nodes; ++i) {
      Node& nd = j->nodes[(size_t)i];
      for (int k : {nd.lhs, nd.rhs}) {
        if (k >= job->n_leaves) {
          Node& src = j->nodes[(size_t)(k - job->n_leaves)];
          src.consumers++;
          src.users.push_back(i);
          nd.pending++;
        }
        if (nd.lhs == nd.rhs) break;
      }
    }
    // the root is consumed by the caller (muscle_hip_server_release)
    j->nodes.back().consumers++;
    j->remaining = job->n_nodes;
    {
      std::lock_guard<std::mutex> lk(S->mu);
      S->st.submitted++;
      if (j->reserve > S->pool.size()) {
        S->st.rejected++;
        fail(MUSCLE_HIP_REJECTED, "job needs more device memory than the server pool holds");
      }
      j->id = S->next_id++;
      j->t_submit = S->now();
      S->jobs[j->id] = j;
      S->waiting.push_back(j);
      S->to_compile.push_back(j);
      *job_id = j->id;
    }
    // contractions are lowered and planned on the planner threads; one the
    // library cannot run ends the job with that status (seen by wait)
    S->cv_plan.notify_one();
  });
}

int muscle_hip_server_wait(muscle_hip_server_t server, int64_t job_id, double timeout_s,
                           muscle_hip_job_result_t* result) {
  Server* S = reinterpret_cast<Server*>(server);
  if (!S) return MUSCLE_HIP_INVALID_ARGUMENT;
  std::unique_lock<std::mutex> lk(S->mu);
  auto it = S->jobs.find(job_id);
  if (it == S->jobs.end()) {
    g_last_error = "unknown job";
    return MUSCLE_HIP_INVALID_ARGUMENT;
  }
  std::shared_ptr<Job> j = it->second;
  auto done = [&] { return j->status != MUSCLE_HIP_NOT_READY; };
  if (timeout_s < 0) S->cv_done.wait(lk, done);
  else S->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), done);
  if (result) {
    std::memset(result, 0, sizeof *result);
    result->status = j->status;
    result->dtype = j->dtype;
    result->result = j->nodes.back().out.t.data;
    result->result_elems = elems(j->nodes.back().out);
    result->submit_s = j->t_submit;
    result->admit_s = j->t_admit;
    result->done_s = j->t_done;
    result->ticks = j->ticks;
  }
  if (j->status == MUSCLE_HIP_NOT_READY) return MUSCLE_HIP_NOT_READY;
  if (j->status != MUSCLE_HIP_OK) g_last_error = j->error;
  return j->status;
}

int muscle_hip_server_copy_result(muscle_hip_server_t server, int64_t job_id, void* dst) {
  return guarded_s([&] {
    Server* S = reinterpret_cast<Server*>(server);
    if (!S || !dst) fail(MUSCLE_HIP_INVALID_ARGUMENT, "null argument");
    const void* src = nullptr;
    size_t bytes = 0;
    {
      std::lock_guard<std::mutex> lk(S->mu);
      auto it = S->jobs.find(job_id);
      if (it == S->jobs.end()) fail(MUSCLE_HIP_INVALID_ARGUMENT, "unknown job");
      Job& j = *it->second;
      if (j.status != MUSCLE_HIP_OK) fail(MUSCLE_HIP_NOT_READY, "job has no result");
      src = j.nodes.back().out.t.data;
      bytes = j.nodes.back().bytes;
    }
    MH_HIP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyDefault));
  });
}

int muscle_hip_server_release(muscle_hip_server_t server, int64_t job_id) {
  return guarded_s([&] {
    Server* S = reinterpret_cast<Server*>(server);
    if (!S) fail(MUSCLE_HIP_INVALID_ARGUMENT, "null server");
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->jobs.find(job_id);
    if (it == S->jobs.end()) fail(MUSCLE_HIP_INVALID_ARGUMENT, "unknown job");
    Job& j = *it->second;
    if (j.status == MUSCLE_HIP_NOT_READY) fail(MUSCLE_HIP_NOT_READY, "job still running");
    Node& root = j.nodes.back();
    if (j.status == MUSCLE_HIP_OK && !root.freed && root.out.t.data) {
      S->pool.release((char*)root.out.t.data, root.bytes);
      root.freed = true;
      const size_t r = al256(std::max<size_t>(root.bytes, 1));
      j.reserved -= r;
      S->pool.reserved -= r;
    }
    S->jobs.erase(it);
    S->cv_work.notify_one();
  });
}

int muscle_hip_server_get_stats(muscle_hip_server_t server, muscle_hip_server_stats_t* stats) {
  return guarded_s([&] {
    Server* S = reinterpret_cast<Server*>(server);
    if (!S || !stats) fail(MUSCLE_HIP_INVALID_ARGUMENT, "null argument");
    std::lock_guard<std::mutex> lk(S->mu);
    *stats = S->st;
    stats->pool_peak_bytes = (int64_t)S->pool.peak();
    stats->active_jobs = (int64_t)S->active.size();
    stats->waiting_jobs = (int64_t)S->waiting.size();
  });
}

double muscle_hip_server_clock(muscle_hip_server_t server) {
  Server* S = reinterpret_cast<Server*>(server);
  return S ? S->now() : 0.0;
}

int muscle_hip_server_destroy(muscle_hip_server_t server) {
  return guarded_s([&] {
    Server* S = reinterpret_cast<Server*>(server);
    if (!S) return;
    {
      std::lock_guard<std::mutex> lk(S->mu);
      S->stop = true;
    }
    S->cv_work.notify_all();
    S->cv_plan.notify_all();
    for (auto& t : S->planners) t.join();
    if (S->worker.joinable()) S->worker.join();
    (void)hipStreamSynchronize(S->h.stream);
    for (Slot& sl : S->slots) {
      if (sl.dev) (void)hipFree(sl.dev);
      if (sl.pinned) (void)hipHostFree(sl.pinned);
      if (sl.ev) (void)hipEventDestroy(sl.ev);
    }
    if (S->arena) (void)hipFree(S->arena);
    S->h.cache.clear();
    (void)hipStreamDestroy(S->h.stream);
    S->h.stream = nullptr;
    delete S;
  });
}

}  // extern "C"
