// Prediction serving: resident forests and continuous batching.
//
// The reference re-uploads the whole forest and dataset on every Predict
// call and runs one launch + sync per 65,536 (tree, row) pairs
// (gpudte_algorithm.cc:269-331, 598-705).  Here a server keeps forests
// resident in HBM and runs a scheduler thread over a small set of in-flight
// batch slots, each with its own stream, pinned staging buffers and device
// buffers sized at load time under the server's HBM budget:
//
//   submit   -> the request (row-major rows, copied) joins the pending FIFO;
//   schedule -> whenever a slot is free, pending requests are packed into it
//               in arrival order, whatever forest they are for (a request
//               larger than the slot is split across batches); the batch's
//               rows are grouped by forest, staged, copied in one transfer,
//               traversed by one launch per forest and the results copied
//               back in one transfer on the slot's stream;
//   finish   -> when the slot's last event completes, every request's rows
//               are scattered into its result; a request whose last rows
//               landed is done.
//
// Requests that arrive while batches are on the GPU are admitted into the
// next free slot, so the copies and traversals of consecutive batches
// overlap on the device.  Admission: forests and slot buffers must fit the
// HBM budget (hipMemGetInfo when no budget is given); the pending backlog is
// bounded, a submit beyond it is refused with DTE_ERR_ADMISSION.  Each row's
// result depends only on its own features and the forest, so a result is
// the same whatever batch it rides in.
//
// A request's rows are copied once, at submit, into a pinned block from the
// server's pool; a batch copies large pieces straight from those blocks to
// the device and stages only small pieces through the slot's pinned buffer,
// so most bytes cross the host once.  A batch records four events (start,
// rows copied, traversals done, results copied) and the scheduler times its
// host staging and scatter, so the server reports where a batch's time goes
// (dte_server_stats).
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

#include "dte_internal.h"

namespace dte {
namespace {

constexpr int kSlots = 4;
// Pieces of at least this many bytes go to the device straight from their
// request's pinned block; smaller ones are staged into the slot buffer.
constexpr size_t kDirectBytes = (size_t)256 << 10;

// Pinned host blocks for request rows, reused by size class (powers of two
// from 64 KiB); at most kCacheCap bytes are kept free.
struct PinnedPool {
  static constexpr size_t kMinBlock = (size_t)64 << 10;
  static constexpr size_t kCacheCap = (size_t)1 << 30;
  std::mutex mu;
  std::multimap<size_t, void*> free_blocks;
  size_t cached = 0;
  void* acquire(size_t bytes, size_t* got) {
    size_t want = kMinBlock;
    while (want < bytes) want <<= 1;
    {
      std::lock_guard<std::mutex> lk(mu);
      auto it = free_blocks.find(want);
      if (it != free_blocks.end()) {
        void* p = it->second;
        free_blocks.erase(it);
        cached -= want;
        *got = want;
        return p;
      }
    }
    void* p = nullptr;
    DTE_HIP(hipHostMalloc(&p, want, hipHostMallocDefault));
    *got = want;
    return p;
  }
  void release(void* p, size_t bytes) {
    {
      std::lock_guard<std::mutex> lk(mu);
      if (cached + bytes <= kCacheCap) {
        free_blocks.emplace(bytes, p);
        cached += bytes;
        return;
      }
    }
    (void)hipHostFree(p);
  }
  ~PinnedPool() {
    for (auto& kv : free_blocks) (void)hipHostFree(kv.second);
  }
};

// A request's rows in a pinned block of the pool (returned on destruction).
struct PinnedRows {
  PinnedPool* pool = nullptr;
  float* p = nullptr;
  size_t bytes = 0;
  PinnedRows() = default;
  PinnedRows(const PinnedRows&) = delete;
  PinnedRows& operator=(const PinnedRows&) = delete;
  PinnedRows& operator=(PinnedRows&& o) noexcept {
    if (this != &o) {
      reset();
      pool = o.pool;
      p = o.p;
      bytes = o.bytes;
      o.p = nullptr;
    }
    return *this;
  }
  ~PinnedRows() { reset(); }
  void reset() {
    if (p) pool->release(p, bytes);
    p = nullptr;
  }
};

struct Resident {
  Forest f;  // pnodes + tree_start only
  int64_t bytes = 0;
};

struct Request {
  uint64_t ticket = 0;
  const Resident* forest = nullptr;  // heap object: never moves while the server lives
  int64_t n_rows = 0;
  int n_features = 0;
  int width = 1;  // result floats per row
  PinnedRows rows;  // n_rows x n_features, row-major
  std::vector<float> out;
  int64_t next_row = 0;  // rows handed to batches
  int64_t rows_done = 0;
  bool done = false;
};

struct Piece {
  Request* req;
  int64_t row0;  // first row of the request
  int64_t n;
  int64_t off;   // first row in the batch (rows of one forest are contiguous)
};

// Rows of one forest inside a batch: rows [row0, row0 + n) of the batch,
// whose features start at x_off floats and whose results start at o_off
// result words of the slot buffers.
struct Group {
  const Resident* forest;
  int64_t row0, n, x_off, o_off;
};

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev_in = nullptr, ev_k = nullptr, ev1 = nullptr;
  float* h_rows = nullptr;   // pinned
  void* h_out = nullptr;     // pinned, int32 votes or fp32 sums per group
  DevBuf<float> d_rows;
  DevBuf<int32_t> d_out;     // int32 votes or fp32 sums (same 4-byte words)
  bool busy = false;
  int64_t n = 0;
  int64_t x_floats = 0, o_words = 0;
  double stage_ms = 0.0;
  std::vector<Piece> pieces;
  std::vector<Group> groups;
};

}  // namespace
}  // namespace dte

struct dte_server {
  int device = 0;
  int64_t budget = 0;
  int64_t max_rows = 0;
  int f_cap = 0;  // feature capacity of the slot buffers
  int c_cap = 0;  // result width capacity
  int64_t slot_bytes = 0;
  int64_t backlog_rows_cap = 0;
  dte::PinnedPool pinned;  // before `requests`: outlives the requests' blocks
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::vector<std::unique_ptr<dte::Resident>> forests;
  std::deque<dte::Request*> pending;
  int64_t pending_rows = 0;
  std::unordered_map<uint64_t, std::unique_ptr<dte::Request>> requests;
  dte::Slot slots[dte::kSlots];
  bool stop = false;
  bool pause = false;  // slot buffers being resized
  std::string fatal;
  uint64_t next_ticket = 1;
  dte_server_stats st{};
  std::thread th;
};

namespace dte {
namespace {

using Clock = std::chrono::steady_clock;

double ms_since(Clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
}

int64_t forest_bytes(const Forest& f) { return (int64_t)(f.pnodes.bytes() + f.tree_start.bytes()); }

void free_slot_buffers(dte_server& s) {
  for (auto& sl : s.slots) {
    if (sl.h_rows) (void)hipHostFree(sl.h_rows);
    if (sl.h_out) (void)hipHostFree(sl.h_out);
    sl.h_rows = nullptr;
    sl.h_out = nullptr;
    sl.d_rows.release();
    sl.d_out.release();
  }
  s.st.hbm_in_use -= s.slot_bytes;
  s.slot_bytes = 0;
}

// (Re)sizes every slot for f features and c result words per row.
void size_slots(dte_server& s, int f, int c) {
  const int64_t need = (int64_t)kSlots * s.max_rows * ((int64_t)f * 4 + (int64_t)c * 4);
  const int64_t others = s.st.hbm_in_use - s.slot_bytes;
  if (others + need > s.budget) fail(DTE_ERR_ADMISSION, "slot buffers do not fit the server's HBM budget");
  free_slot_buffers(s);
  for (auto& sl : s.slots) {
    sl.d_rows.alloc((size_t)s.max_rows * f);
    sl.d_out.alloc((size_t)s.max_rows * c);
    DTE_HIP(hipHostMalloc((void**)&sl.h_rows, sizeof(float) * s.max_rows * f, hipHostMallocDefault));
    DTE_HIP(hipHostMalloc(&sl.h_out, sizeof(int32_t) * s.max_rows * c, hipHostMallocDefault));
  }
  s.f_cap = f;
  s.c_cap = c;
  s.slot_bytes = need;
  s.st.hbm_in_use += need;
}

// Packs pending requests into slot sl in arrival order, for any forest, up
// to max_rows rows; then lays the pieces out grouped by forest.  Called with
// the lock held; returns false when nothing was taken.
bool take_batch(dte_server& s, Slot& sl) {
  if (s.pending.empty()) return false;
  sl.pieces.clear();
  sl.groups.clear();
  sl.n = 0;
  while (!s.pending.empty() && sl.n < s.max_rows) {
    Request* r = s.pending.front();
    const int64_t take = std::min(r->n_rows - r->next_row, s.max_rows - sl.n);
    sl.pieces.push_back(Piece{r, r->next_row, take, 0});
    r->next_row += take;
    sl.n += take;
    s.pending_rows -= take;
    if (r->next_row == r->n_rows) s.pending.pop_front();
  }
  // Group by forest, in the order of each forest's first piece; the pieces
  // of a forest keep their order.  Every forest's F <= f_cap and width <=
  // c_cap, so the groups fit the slot buffers.
  std::vector<const Resident*> order;
  for (const Piece& p : sl.pieces)
    if (std::find(order.begin(), order.end(), p.req->forest) == order.end()) order.push_back(p.req->forest);
  auto rank = [&](const Piece& p) { return std::find(order.begin(), order.end(), p.req->forest) - order.begin(); };
  std::stable_sort(sl.pieces.begin(), sl.pieces.end(), [&](const Piece& a, const Piece& b) { return rank(a) < rank(b); });
  int64_t row = 0, x = 0, o = 0;
  for (Piece& p : sl.pieces) {
    const Resident* fr = p.req->forest;
    if (sl.groups.empty() || sl.groups.back().forest != fr) sl.groups.push_back(Group{fr, row, 0, x, o});
    Group& g = sl.groups.back();
    p.off = row;
    g.n += p.n;
    row += p.n;
    x += p.n * p.req->n_features;
    o += p.n * p.req->width;
  }
  sl.x_floats = x;
  sl.o_words = o;
  return sl.n > 0;
}

// Stages and launches the batch in sl (lock not held: only the scheduler
// touches a busy slot, requests with pieces in flight stay alive and
// resident forests are never freed while the server runs).
void launch_batch(Slot& sl) {
  DTE_HIP(hipEventRecord(sl.ev0, sl.stream));
  // Pieces lie back to back in the batch's row space, grouped by forest:
  // small ones are staged into h_rows at their own offsets and go over in
  // runs, large ones straight from their request's pinned block.
  double stage = 0.0;
  int64_t run_lo = -1, run_hi = -1;  // staged run, in floats of the row space
  auto flush = [&] {
    if (run_lo >= 0)
      DTE_HIP(hipMemcpyAsync(sl.d_rows.p + run_lo, sl.h_rows + run_lo, sizeof(float) * (run_hi - run_lo),
                             hipMemcpyHostToDevice, sl.stream));
    run_lo = run_hi = -1;
  };
  size_t gi = 0;
  for (const Piece& p : sl.pieces) {
    while (p.off >= sl.groups[gi].row0 + sl.groups[gi].n) ++gi;
    const Group& g = sl.groups[gi];
    const int64_t F = g.forest->f.n_features;
    const int64_t x = g.x_off + (p.off - g.row0) * F;
    const size_t bytes = sizeof(float) * (size_t)(p.n * F);
    const float* src = p.req->rows.p + p.row0 * F;
    if (bytes >= kDirectBytes) {
      flush();
      DTE_HIP(hipMemcpyAsync(sl.d_rows.p + x, src, bytes, hipMemcpyHostToDevice, sl.stream));
    } else {
      const Clock::time_point t0 = Clock::now();
      std::memcpy(sl.h_rows + x, src, bytes);
      stage += ms_since(t0);
      if (run_lo < 0) run_lo = x;
      run_hi = x + p.n * F;
    }
  }
  flush();
  sl.stage_ms = stage;
  DTE_HIP(hipEventRecord(sl.ev_in, sl.stream));
  for (const Group& g : sl.groups) {
    const Forest& f = g.forest->f;
    const float* x = sl.d_rows.p + g.x_off;
    if (f.task == 0)
      predict_forest(f, x, g.n, f.n_features, 1, sl.d_out.p + g.o_off, nullptr, sl.stream);
    else
      predict_forest(f, x, g.n, f.n_features, 1, nullptr, reinterpret_cast<float*>(sl.d_out.p + g.o_off),
                     sl.stream);
  }
  DTE_HIP(hipEventRecord(sl.ev_k, sl.stream));
  DTE_HIP(hipMemcpyAsync(sl.h_out, sl.d_out.p, sizeof(int32_t) * sl.o_words, hipMemcpyDeviceToHost, sl.stream));
  DTE_HIP(hipEventRecord(sl.ev1, sl.stream));
}

// Scatters a finished batch into its requests (lock held).
void finish_batch(dte_server& s, Slot& sl) {
  const Clock::time_point t0 = Clock::now();
  float a = 0.f, b = 0.f, c = 0.f;
  if (hipEventElapsedTime(&a, sl.ev0, sl.ev_in) == hipSuccess &&
      hipEventElapsedTime(&b, sl.ev_in, sl.ev_k) == hipSuccess &&
      hipEventElapsedTime(&c, sl.ev_k, sl.ev1) == hipSuccess) {
    s.st.h2d_ms += a;
    s.st.kernel_ms += b;
    s.st.d2h_ms += c;
    s.st.busy_ms += (double)a + b + c;
  }
  s.st.stage_ms += sl.stage_ms;
  size_t gi = 0;
  for (const Piece& p : sl.pieces) {
    while (p.off >= sl.groups[gi].row0 + sl.groups[gi].n) ++gi;
    const Group& g = sl.groups[gi];
    Request* r = p.req;
    const Forest& f = g.forest->f;
    if (f.task == 0) {
      const int C = f.n_classes;
      const int32_t* v = static_cast<const int32_t*>(sl.h_out) + g.o_off + (p.off - g.row0) * C;
      float* o = r->out.data() + p.row0 * C;
      for (int64_t i = 0; i < p.n * C; ++i) o[i] = (float)v[i];
    } else {
      std::memcpy(r->out.data() + p.row0, static_cast<const float*>(sl.h_out) + g.o_off + (p.off - g.row0),
                  sizeof(float) * p.n);
    }
    r->rows_done += p.n;
    if (r->rows_done == r->n_rows) {
      r->done = true;
      r->rows.reset();  // no piece of it is in flight any more
      s.st.requests_done += 1;
      s.st.rows_done += r->n_rows;
    }
  }
  s.st.groups_launched += (int64_t)sl.groups.size();
  sl.pieces.clear();
  sl.groups.clear();
  sl.busy = false;
  s.st.scatter_ms += ms_since(t0);
}

void scheduler(dte_server* s) {
  (void)hipSetDevice(s->device);
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    bool progressed = false;
    bool any_busy = false;
    for (auto& sl : s->slots) {
      if (!sl.busy) continue;
      const hipError_t q = hipEventQuery(sl.ev1);
      if (q == hipSuccess) {
        finish_batch(*s, sl);
        progressed = true;
      } else if (q != hipErrorNotReady) {
        s->fatal = std::string("HIP error in a serving batch: ") + hipGetErrorString(q);
        sl.busy = false;
      } else {
        any_busy = true;
      }
    }
    if (progressed) s->cv_done.notify_all();
    if (s->fatal.empty() && !s->pause) {
      for (auto& sl : s->slots) {
        if (sl.busy || !take_batch(*s, sl)) continue;
        sl.busy = true;
        any_busy = true;
        progressed = true;
        s->st.batches_launched += 1;
        lk.unlock();
        std::string err;
        try {
          launch_batch(sl);
        } catch (const std::exception& e) {
          err = e.what();
        }
        lk.lock();
        if (!err.empty()) {
          s->fatal = "serving batch launch failed: " + err;
          sl.busy = false;
        }
      }
    }
    if (!s->fatal.empty()) s->cv_done.notify_all();
    if (s->stop && !any_busy && (s->pending.empty() || !s->fatal.empty())) break;
    if (!any_busy) s->cv_done.notify_all();  // idle: wakes a resize waiting for the slots
    if (!progressed) {
      if (any_busy)
        s->cv_work.wait_for(lk, std::chrono::microseconds(20));
      else
        s->cv_work.wait(lk, [&] { return s->stop || (!s->pending.empty() && !s->pause); });
    }
  }
}

Request* find(dte_server* s, uint64_t ticket) {
  auto it = s->requests.find(ticket);
  if (it == s->requests.end()) fail(DTE_ERR_NOT_FOUND, "unknown ticket");
  return it->second.get();
}

void deliver(dte_server* s, uint64_t ticket, Request* r, float* out) {
  if (out) std::memcpy(out, r->out.data(), sizeof(float) * r->out.size());
  s->requests.erase(ticket);
}

}  // namespace
}  // namespace dte

using namespace dte;

extern "C" {

int dte_server_create(int32_t device_id, int64_t hbm_budget_bytes, int32_t max_batch_rows, dte_server** out) {
  return run_guarded([&] {
    if (!out) fail(DTE_ERR_INVALID_ARGUMENT, "out is NULL");
    if (max_batch_rows <= 0) fail(DTE_ERR_INVALID_ARGUMENT, "max_batch_rows must be positive");
    if (hbm_budget_bytes < 0) fail(DTE_ERR_INVALID_ARGUMENT, "negative HBM budget");
    select_device(device_id);
    auto s = std::make_unique<dte_server>();
    s->device = device_id;
    s->max_rows = max_batch_rows;
    size_t free_b = 0, total_b = 0;
    DTE_HIP(hipMemGetInfo(&free_b, &total_b));
    const int64_t usable = (int64_t)(free_b / 10 * 9);
    s->budget = hbm_budget_bytes > 0 ? std::min<int64_t>(hbm_budget_bytes, usable) : usable;
    s->backlog_rows_cap = std::max<int64_t>((int64_t)64 * max_batch_rows, 1 << 22);
    for (auto& sl : s->slots) {
      DTE_HIP(hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking));
      DTE_HIP(hipEventCreate(&sl.ev0));
      DTE_HIP(hipEventCreate(&sl.ev_in));
      DTE_HIP(hipEventCreate(&sl.ev_k));
      DTE_HIP(hipEventCreate(&sl.ev1));
    }
    s->th = std::thread(scheduler, s.get());
    *out = s.release();
  });
}

int dte_server_load_forest(dte_server* s, const dte_forest* forest, int32_t* slot) {
  return run_guarded([&] {
    if (!s || !forest || !slot) fail(DTE_ERR_INVALID_ARGUMENT, "NULL argument");
    const Forest& src = forest->f;
    if (src.device != s->device) fail(DTE_ERR_INVALID_ARGUMENT, "forest lives on another device");
    select_device(s->device);
    std::unique_lock<std::mutex> lk(s->mu);
    const int width = src.task == 0 ? src.n_classes : 1;
    const int64_t fb = forest_bytes(src);
    if (s->st.hbm_in_use + fb > s->budget) {
      s->st.requests_rejected += 1;
      fail(DTE_ERR_ADMISSION, "forest does not fit the server's HBM budget");
    }
    if (src.n_features > s->f_cap || width > s->c_cap) {
      // Resize the slot buffers once every in-flight batch has finished.
      s->pause = true;
      s->cv_done.wait(lk, [&] {
        return std::none_of(std::begin(s->slots), std::end(s->slots), [](const Slot& q) { return q.busy; });
      });
      try {
        size_slots(*s, std::max(src.n_features, s->f_cap), std::max(width, s->c_cap));
      } catch (...) {
        s->pause = false;
        s->cv_work.notify_all();
        s->st.requests_rejected += 1;
        throw;
      }
      s->pause = false;
      s->cv_work.notify_all();
      if (s->st.hbm_in_use + fb > s->budget) {
        s->st.requests_rejected += 1;
        fail(DTE_ERR_ADMISSION, "forest does not fit the server's HBM budget");
      }
    }
    auto r = std::make_unique<Resident>();
    Forest& f = r->f;
    f.device = src.device;
    f.task = src.task;
    f.algorithm = src.algorithm;
    f.n_classes = src.n_classes;
    f.n_features = src.n_features;
    f.n_trees = src.n_trees;
    f.fixed_shift = src.fixed_shift;
    f.n_nodes = src.n_nodes;
    f.tree_start.alloc(src.n_trees + 1);
    f.pnodes.alloc(src.n_nodes);
    hipStream_t ls = library_stream(s->device);
    DTE_HIP(hipMemcpyAsync(f.tree_start.p, src.tree_start.p, src.tree_start.bytes(), hipMemcpyDeviceToDevice, ls));
    DTE_HIP(hipMemcpyAsync(f.pnodes.p, src.pnodes.p, sizeof(PNode) * src.n_nodes, hipMemcpyDeviceToDevice, ls));
    DTE_HIP(hipStreamSynchronize(ls));
    r->bytes = forest_bytes(f);
    s->st.hbm_in_use += r->bytes;
    s->forests.push_back(std::move(r));
    *slot = (int32_t)s->forests.size() - 1;
  });
}

int dte_server_submit(dte_server* s, int32_t slot, const float* rows, int64_t n_rows, uint64_t* ticket) {
  return run_guarded([&] {
    if (!s || !ticket || (!rows && n_rows > 0)) fail(DTE_ERR_INVALID_ARGUMENT, "NULL argument");
    if (n_rows <= 0) fail(DTE_ERR_INVALID_ARGUMENT, "a request needs at least one row");
    std::unique_lock<std::mutex> lk(s->mu);
    if (!s->fatal.empty()) fail(DTE_ERR_HIP, s->fatal);
    if (slot < 0 || slot >= (int32_t)s->forests.size()) fail(DTE_ERR_NOT_FOUND, "no forest in that slot");
    if (s->pending_rows + n_rows > s->backlog_rows_cap) {
      s->st.requests_rejected += 1;
      fail(DTE_ERR_ADMISSION, "request backlog full");
    }
    const Resident* fr = s->forests[slot].get();
    auto r = std::make_unique<Request>();
    r->ticket = s->next_ticket++;
    r->forest = fr;
    r->n_rows = n_rows;
    r->n_features = fr->f.n_features;
    r->width = fr->f.task == 0 ? fr->f.n_classes : 1;
    lk.unlock();
    const size_t bytes = sizeof(float) * (size_t)n_rows * r->n_features;
    select_device(s->device);  // pinned blocks of this server's device
    r->rows.pool = &s->pinned;
    r->rows.p = static_cast<float*>(s->pinned.acquire(bytes, &r->rows.bytes));
    std::memcpy(r->rows.p, rows, bytes);
    r->out.assign((size_t)n_rows * r->width, 0.0f);
    lk.lock();
    Request* raw = r.get();
    *ticket = r->ticket;
    s->requests.emplace(r->ticket, std::move(r));
    s->pending.push_back(raw);
    s->pending_rows += n_rows;
    s->cv_work.notify_all();
  });
}

int dte_server_poll(dte_server* s, uint64_t ticket, float* out, int32_t* done) {
  return run_guarded([&] {
    if (!s || !done) fail(DTE_ERR_INVALID_ARGUMENT, "NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    Request* r = find(s, ticket);
    if (!r->done) {
      if (!s->fatal.empty()) fail(DTE_ERR_HIP, s->fatal);
      *done = 0;
      return;
    }
    *done = 1;
    deliver(s, ticket, r, out);
  });
}

int dte_server_wait(dte_server* s, uint64_t ticket, float* out) {
  return run_guarded([&] {
    if (!s) fail(DTE_ERR_INVALID_ARGUMENT, "NULL argument");
    std::unique_lock<std::mutex> lk(s->mu);
    (void)find(s, ticket);
    // The request is looked up again after every wake-up: another caller
    // that polled or waited on the same ticket may have taken the result
    // (and freed the request) in the meantime.
    Request* r = nullptr;
    s->cv_done.wait(lk, [&] {
      auto it = s->requests.find(ticket);
      r = it == s->requests.end() ? nullptr : it->second.get();
      return !r || r->done || !s->fatal.empty();
    });
    if (!r) fail(DTE_ERR_NOT_FOUND, "the ticket's result was taken by another caller");
    if (!r->done) fail(DTE_ERR_HIP, s->fatal);
    deliver(s, ticket, r, out);
  });
}

int dte_server_get_stats(dte_server* s, dte_server_stats* out) {
  return run_guarded([&] {
    if (!s || !out) fail(DTE_ERR_INVALID_ARGUMENT, "NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->st;
  });
}

int dte_server_destroy(dte_server* s) {
  return run_guarded([&] {
    if (!s) return;
    {
      std::lock_guard<std::mutex> lk(s->mu);
      s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->th.joinable()) s->th.join();
    (void)hipSetDevice(s->device);
    for (auto& sl : s->slots) {
      if (sl.stream) (void)hipStreamSynchronize(sl.stream);
    }
    free_slot_buffers(*s);
    for (auto& sl : s->slots) {
      for (hipEvent_t e : {sl.ev0, sl.ev_in, sl.ev_k, sl.ev1})
        if (e) (void)hipEventDestroy(e);
      if (sl.stream) (void)hipStreamDestroy(sl.stream);
    }
    delete s;
  });
}

}  // extern "C"
