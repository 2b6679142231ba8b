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
//   schedule -> whenever a slot is free, the pending requests of the head
//               request's forest are packed into it (a request larger than
//               the slot is split across batches), rows staged, copied,
//               traversed and the votes copied back on the slot's stream;
//   finish   -> when the slot's event completes, every request's rows are
//               scattered into its result; a request whose last rows landed
//               is done.
//
// Requests that arrive while batches are on the GPU are admitted into the
// next free slot, so the copies and traversals of consecutive batches
// overlap on the device.  Admission: forests and slot buffers must fit the
// HBM budget (hipMemGetInfo when no budget is given); the pending backlog is
// bounded, a submit beyond it is refused with DTE_ERR_ADMISSION.  Each row's
// result depends only on its own features and the forest, so a result is
// the same whatever batch it rides in.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "dte_internal.h"

namespace dte {
namespace {

constexpr int kSlots = 4;

struct Resident {
  Forest f;  // pnodes + tree_start only
  int64_t bytes = 0;
};

struct Request {
  uint64_t ticket = 0;
  int slot = 0;
  int64_t n_rows = 0;
  int width = 1;  // result floats per row
  std::vector<float> rows;
  std::vector<float> out;
  int64_t next_row = 0;  // rows handed to batches
  int64_t rows_done = 0;
  bool done = false;
};

struct Piece {
  Request* req;
  int64_t row0;  // first row of the request
  int64_t n;
  int64_t off;   // first row in the batch
};

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  float* h_rows = nullptr;   // pinned
  void* h_out = nullptr;     // pinned, int32 votes or fp32 sums
  DevBuf<float> d_rows;
  DevBuf<int32_t> d_votes;
  DevBuf<float> d_sums;
  bool busy = false;
  int forest = -1;
  int64_t n = 0;
  std::vector<Piece> pieces;
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

int64_t forest_bytes(const Forest& f) { return (int64_t)(f.pnodes.bytes() + f.tree_start.bytes()); }

void free_slot_buffers(dte_server& s) {
  for (auto& sl : s.slots) {
    if (sl.h_rows) (void)hipHostFree(sl.h_rows);
    if (sl.h_out) (void)hipHostFree(sl.h_out);
    sl.h_rows = nullptr;
    sl.h_out = nullptr;
    sl.d_rows.release();
    sl.d_votes.release();
    sl.d_sums.release();
  }
  s.st.hbm_in_use -= s.slot_bytes;
  s.slot_bytes = 0;
}

// (Re)sizes every slot for f features and c result floats per row.
void size_slots(dte_server& s, int f, int c) {
  const int64_t need = (int64_t)kSlots * s.max_rows * ((int64_t)f * 4 + (int64_t)c * 4 + 4);
  const int64_t others = s.st.hbm_in_use - s.slot_bytes;
  if (others + need > s.budget) fail(DTE_ERR_ADMISSION, "slot buffers do not fit the server's HBM budget");
  free_slot_buffers(s);
  for (auto& sl : s.slots) {
    sl.d_rows.alloc((size_t)s.max_rows * f);
    sl.d_votes.alloc((size_t)s.max_rows * c);
    sl.d_sums.alloc((size_t)s.max_rows);
    DTE_HIP(hipHostMalloc((void**)&sl.h_rows, sizeof(float) * s.max_rows * f, hipHostMallocDefault));
    DTE_HIP(hipHostMalloc(&sl.h_out, sizeof(float) * s.max_rows * c, hipHostMallocDefault));
  }
  s.f_cap = f;
  s.c_cap = c;
  s.slot_bytes = need;
  s.st.hbm_in_use += need;
}

// Packs pending requests of the head request's forest into slot sl.
// Called with the lock held; returns false when nothing was taken.
bool take_batch(dte_server& s, Slot& sl) {
  if (s.pending.empty()) return false;
  const int fslot = s.pending.front()->slot;
  sl.pieces.clear();
  sl.n = 0;
  sl.forest = fslot;
  for (auto it = s.pending.begin(); it != s.pending.end() && sl.n < s.max_rows;) {
    Request* r = *it;
    if (r->slot != fslot) {
      ++it;
      continue;
    }
    const int64_t take = std::min(r->n_rows - r->next_row, s.max_rows - sl.n);
    sl.pieces.push_back(Piece{r, r->next_row, take, sl.n});
    r->next_row += take;
    sl.n += take;
    s.pending_rows -= take;
    if (r->next_row == r->n_rows)
      it = s.pending.erase(it);
    else
      ++it;
  }
  return sl.n > 0;
}

// Stages and launches the batch in sl (lock not held: only the scheduler
// touches a busy slot, and requests with pieces in flight stay alive).
void launch_batch(dte_server& s, Slot& sl) {
  const Forest& f = s.forests[sl.forest]->f;
  const int F = f.n_features;
  for (const Piece& p : sl.pieces)
    std::memcpy(sl.h_rows + p.off * F, p.req->rows.data() + p.row0 * F, sizeof(float) * p.n * F);
  DTE_HIP(hipEventRecord(sl.ev0, sl.stream));
  DTE_HIP(hipMemcpyAsync(sl.d_rows.p, sl.h_rows, sizeof(float) * sl.n * F, hipMemcpyHostToDevice, sl.stream));
  if (f.task == 0) {
    predict_forest(f, sl.d_rows.p, sl.n, F, 1, sl.d_votes.p, nullptr, sl.stream);
    DTE_HIP(hipMemcpyAsync(sl.h_out, sl.d_votes.p, sizeof(int32_t) * sl.n * f.n_classes, hipMemcpyDeviceToHost,
                           sl.stream));
  } else {
    predict_forest(f, sl.d_rows.p, sl.n, F, 1, nullptr, sl.d_sums.p, sl.stream);
    DTE_HIP(hipMemcpyAsync(sl.h_out, sl.d_sums.p, sizeof(float) * sl.n, hipMemcpyDeviceToHost, sl.stream));
  }
  DTE_HIP(hipEventRecord(sl.ev1, sl.stream));
}

// Scatters a finished batch into its requests (lock held).
void finish_batch(dte_server& s, Slot& sl) {
  const Forest& f = s.forests[sl.forest]->f;
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, sl.ev0, sl.ev1) == hipSuccess) s.st.busy_ms += ms;
  for (const Piece& p : sl.pieces) {
    Request* r = p.req;
    if (f.task == 0) {
      const int C = f.n_classes;
      const int32_t* v = static_cast<const int32_t*>(sl.h_out) + p.off * C;
      float* o = r->out.data() + p.row0 * C;
      for (int64_t i = 0; i < p.n * C; ++i) o[i] = (float)v[i];
    } else {
      std::memcpy(r->out.data() + p.row0, static_cast<const float*>(sl.h_out) + p.off, sizeof(float) * p.n);
    }
    r->rows_done += p.n;
    if (r->rows_done == r->n_rows) {
      r->done = true;
      std::vector<float>().swap(r->rows);
      s.st.requests_done += 1;
      s.st.rows_done += r->n_rows;
    }
  }
  sl.pieces.clear();
  sl.busy = false;
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
          launch_batch(*s, sl);
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
    const Forest& f = s->forests[slot]->f;
    auto r = std::make_unique<Request>();
    r->ticket = s->next_ticket++;
    r->slot = slot;
    r->n_rows = n_rows;
    r->width = f.task == 0 ? f.n_classes : 1;
    lk.unlock();
    r->rows.assign(rows, rows + n_rows * f.n_features);
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
    Request* r = find(s, ticket);
    s->cv_done.wait(lk, [&] { return r->done || !s->fatal.empty(); });
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
      if (sl.ev0) (void)hipEventDestroy(sl.ev0);
      if (sl.ev1) (void)hipEventDestroy(sl.ev1);
      if (sl.stream) (void)hipStreamDestroy(sl.stream);
    }
    delete s;
  });
}

}  // extern "C"
