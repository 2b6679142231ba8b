// C-ABI glue of libdte: status codes, handles, host<->device staging,
// kernel-time profiling.  Exceptions never cross this boundary.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <thread>
#include <atomic>
#include <cstdio>
#include <cstring>
#include <map>
#include <functional>
#include <memory>
#include <mutex>
#include <new>
#include <queue>
#include <stdexcept>
#include <string>
#include <vector>

#include "dte_internal.h"

namespace dte {

namespace {
thread_local std::string t_err;
thread_local int t_code = 0;

struct DteException : std::exception {
  int code;
  std::string msg;
  DteException(int c, std::string m) : code(c), msg(std::move(m)) {}
  const char* what() const noexcept override { return msg.c_str(); }
};

std::mutex g_dev_mu;
std::map<int, hipStream_t> g_streams;
std::map<int, std::vector<hipStream_t>> g_workers;  // streams of concurrent fits

std::mutex g_prof_mu;
bool g_prof_on = false;
struct Pending {
  std::string name;
  hipEvent_t a, b;
  double bytes;
};
std::vector<Pending> g_pending;
std::vector<KernelStat> g_stats;

void drain_pending() {
  for (auto& p : g_pending) {
    float ms = 0.f;
    (void)hipEventSynchronize(p.b);
    (void)hipEventElapsedTime(&ms, p.a, p.b);
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
    auto it = std::find_if(g_stats.begin(), g_stats.end(), [&](const KernelStat& k) { return k.name == p.name; });
    if (it == g_stats.end()) {
      g_stats.push_back(KernelStat{p.name, 0, 0.0, 0.0});
      it = g_stats.end() - 1;
    }
    it->launches += 1;
    it->total_ms += ms;
    it->bytes += p.bytes;
  }
  g_pending.clear();
}

template <typename F>
int guarded(F&& fn) {
  try {
    fn();
    t_code = 0;
    return DTE_OK;
  } catch (const DteException& e) {
    t_code = e.code;
    t_err = e.msg;
    return e.code;
  } catch (const std::bad_alloc&) {
    t_code = DTE_ERR_OUT_OF_MEMORY;
    t_err = "host allocation failed";
    return t_code;
  } catch (const std::exception& e) {
    t_code = DTE_ERR_INTERNAL;
    t_err = e.what();
    return t_code;
  }
}

void check_params(const dte_params* p) {
  if (!p) fail(DTE_ERR_INVALID_ARGUMENT, "params is NULL");
  if (p->algorithm != DTE_RANDOM_FOREST && p->algorithm != DTE_EXTRA_TREES)
    fail(DTE_ERR_INVALID_ARGUMENT, "unknown algorithm");
  if (p->task != DTE_CLASSIFICATION && p->task != DTE_REGRESSION) fail(DTE_ERR_INVALID_ARGUMENT, "unknown task");
  if (p->n_trees <= 0) fail(DTE_ERR_INVALID_ARGUMENT, "n_trees must be positive");
  if (p->n_features_per_node < 0) fail(DTE_ERR_INVALID_ARGUMENT, "n_features_per_node must be >= 0");
  if (p->max_depth < 0) fail(DTE_ERR_INVALID_ARGUMENT, "max_depth must be >= 0");
  if (p->min_node_size < 0) fail(DTE_ERR_INVALID_ARGUMENT, "min_node_size must be >= 0");
  if (p->task == DTE_CLASSIFICATION && (p->n_classes <= 0 || p->n_classes > 32))
    fail(DTE_ERR_INVALID_ARGUMENT, "n_classes must be in [1, 32]");
  if (p->build_mode != DTE_BUILD_BATCHED && p->build_mode != DTE_BUILD_PLAIN &&
      p->build_mode != DTE_BUILD_BATCHED_HOST)
    fail(DTE_ERR_INVALID_ARGUMENT, "unknown build_mode");
  if (p->tree_offset < 0 || (int64_t)p->tree_offset + p->n_trees > 0x7FFFFFFF)
    fail(DTE_ERR_INVALID_ARGUMENT, "tree_offset out of range");
  auto in = [](int32_t v, int32_t lo, int32_t hi) { return v == 0 || (v >= lo && v <= hi); };
  if (!in(p->cut_team, 64, kMaxCutTeam)) fail(DTE_ERR_INVALID_ARGUMENT, "cut_team must be 0 or in [64, 512]");
  if (!in(p->cut_block, 64, 2048)) fail(DTE_ERR_INVALID_ARGUMENT, "cut_block must be 0 or in [64, 2048]");
  if (p->cut_team && p->cut_block && p->cut_block < p->cut_team)
    fail(DTE_ERR_INVALID_ARGUMENT, "cut_block must not be below cut_team");
  if (!in(p->bucket_target, 64, 512)) fail(DTE_ERR_INVALID_ARGUMENT, "bucket_target must be 0 or in [64, 512]");
  if (!in(p->cut_ert, 64, 1 << 20)) fail(DTE_ERR_INVALID_ARGUMENT, "cut_ert must be 0 or in [64, 2^20]");
  for (int i = 0; i < 2; ++i)
    if (p->reserved[i] != 0) fail(DTE_ERR_INVALID_ARGUMENT, "reserved fields must be zero");
}

void set_device(int device) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) fail(DTE_ERR_NO_DEVICE, "no HIP device available");
  if (device < 0 || device >= n) fail(DTE_ERR_NO_DEVICE, "device id out of range");
  DTE_HIP(hipSetDevice(device));
  device_init(device);
}

}  // namespace

void set_error(int code, const std::string& msg) {
  t_code = code;
  t_err = msg;
}

int run_guarded(const std::function<void()>& fn) { return guarded(fn); }

void select_device(int device) { set_device(device); }

void fail(int code, const std::string& msg) { throw DteException(code, msg); }

// ------------------------------------------------------- scratch pool --
namespace {
std::mutex g_pool_mu;
// (device, stream) -> cached blocks by size
std::map<std::pair<int, hipStream_t>, std::multimap<size_t, void*>> g_pool;
thread_local bool t_pool_on = false;
thread_local hipStream_t t_pool_s = nullptr;
constexpr size_t kPoolGrain = (size_t)64 << 10;
}  // namespace

bool pool_on() { return t_pool_on; }

PoolScope::PoolScope(hipStream_t s) : prev_on(t_pool_on), prev_s(t_pool_s) {
  t_pool_on = true;
  t_pool_s = s;
}
PoolScope::~PoolScope() {
  t_pool_on = prev_on;
  t_pool_s = prev_s;
}
PoolSuspend::PoolSuspend() : prev_on(t_pool_on) { t_pool_on = false; }
PoolSuspend::~PoolSuspend() { t_pool_on = prev_on; }

void* pool_alloc(size_t bytes, size_t* got, int* dev, hipStream_t* owner) {
  int d = 0;
  DTE_HIP(hipGetDevice(&d));
  const size_t want = (bytes + kPoolGrain - 1) / kPoolGrain * kPoolGrain;
  *dev = d;
  *owner = t_pool_s;
  {
    std::lock_guard<std::mutex> lk(g_pool_mu);
    auto& m = g_pool[{d, t_pool_s}];
    auto it = m.lower_bound(want);
    // a cached block serves requests down to two thirds of its size
    if (it != m.end() && it->first <= want + want / 2) {
      void* p = it->second;
      *got = it->first;
      m.erase(it);
      return p;
    }
  }
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, want);
  if (e == hipErrorOutOfMemory) {
    (void)hipGetLastError();
    pool_trim();
    e = hipMalloc(&p, want);
  }
  if (e != hipSuccess)
    fail(e == hipErrorOutOfMemory ? DTE_ERR_OUT_OF_MEMORY : DTE_ERR_HIP,
         std::string("hipMalloc: ") + hipGetErrorString(e));
  *got = want;
  return p;
}

void pool_free(void* p, size_t bytes, int dev, hipStream_t owner) {
  std::lock_guard<std::mutex> lk(g_pool_mu);
  g_pool[{dev, owner}].emplace(bytes, p);
}

size_t pool_trim() {
  std::lock_guard<std::mutex> lk(g_pool_mu);
  size_t freed = 0;
  int cur = 0;
  (void)hipGetDevice(&cur);
  for (auto& kv : g_pool) {
    if (kv.second.empty()) continue;
    // blocks may still be read by work queued on their stream
    (void)hipSetDevice(kv.first.first);
    (void)hipDeviceSynchronize();
    for (auto& b : kv.second) {
      (void)hipFree(b.second);
      freed += b.first;
    }
    kv.second.clear();
  }
  (void)hipSetDevice(cur);
  return freed;
}

bool profiling_enabled() { return g_prof_on; }

void profile_record(const char* name, hipEvent_t a, hipEvent_t b, double bytes) {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  g_pending.push_back(Pending{name, a, b, bytes});
}

void upload_ln_table(int device);

// ------------------------------------------- no-allocation predict ----
// Device scratch of dte_predict_noalloc: the staged samples and the votes /
// sums of one call.  A file-scope device array: every GPU the library runs
// on holds its own copy, at its own address, until the process ends.  In
// host code the name `scratch` is only a stand-in for the symbol, so its
// device address is taken once per device at initialisation
// (hipGetSymbolAddress, with that device current) and that address is what
// kernels and copies receive.
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
namespace {
std::map<int, void*> g_scratch;                // device -> address of its scratch (under g_dev_mu)
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu[kMaxScratchDevices];   // one call at a time per device
}  // namespace

void device_init(int device) {
  std::lock_guard<std::mutex> lk(g_dev_mu);
  if (g_streams.count(device)) return;
  upload_ln_table(device);
  void* sp = nullptr;
  DTE_HIP(hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch)));
  g_scratch[device] = sp;
  hipStream_t s;
  DTE_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  g_streams[device] = s;
}

void* noalloc_scratch(int device) {
  std::lock_guard<std::mutex> lk(g_dev_mu);
  auto it = g_scratch.find(device);
  if (it == g_scratch.end() || !it->second) fail(DTE_ERR_INTERNAL, "device not initialised");
  return it->second;
}

// k-th stream of the concurrent-fit workers of a device (created on first
// use, kept for the life of the process so that their scratch pools are
// reused).  The device must be current.
hipStream_t worker_stream(int device, int k) {
  std::lock_guard<std::mutex> lk(g_dev_mu);
  auto& v = g_workers[device];
  while ((int)v.size() <= k) {
    hipStream_t s;
    DTE_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    v.push_back(s);
  }
  return v[k];
}

hipStream_t library_stream(int device) {
  std::lock_guard<std::mutex> lk(g_dev_mu);
  auto it = g_streams.find(device);
  if (it == g_streams.end()) fail(DTE_ERR_INTERNAL, "device not initialised");
  return it->second;
}

// Add a host-measured interval (ms) to a family, as one launch.
void profile_host(const char* name, double ms) {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  drain_pending();
  auto it = std::find_if(g_stats.begin(), g_stats.end(), [&](const KernelStat& k) { return k.name == name; });
  if (it == g_stats.end()) {
    g_stats.push_back(KernelStat{name, 0, 0.0, 0.0});
    it = g_stats.end() - 1;
  }
  it->launches += 1;
  it->total_ms += ms;
}

// Add a count to a family's launches without a time (e.g. graph replays).
void profile_count(const char* name, int64_t n) {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  drain_pending();
  auto it = std::find_if(g_stats.begin(), g_stats.end(), [&](const KernelStat& k) { return k.name == name; });
  if (it == g_stats.end()) {
    g_stats.push_back(KernelStat{name, 0, 0.0, 0.0});
    it = g_stats.end() - 1;
  }
  it->launches += n;
}

// Add algorithmic bytes to a kernel family (known only after the launches
// it covers have finished).
void profile_add_bytes(const char* name, double bytes) {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  drain_pending();
  auto it = std::find_if(g_stats.begin(), g_stats.end(), [&](const KernelStat& k) { return k.name == name; });
  if (it == g_stats.end()) {
    g_stats.push_back(KernelStat{name, 0, 0.0, 0.0});
    it = g_stats.end() - 1;
  }
  it->bytes += bytes;
}

}  // namespace dte

using namespace dte;

namespace {
struct HostModel {
  std::vector<int64_t> ts;
  std::vector<int32_t> feature, left;
  std::vector<float> thr, value;
  std::vector<uint32_t> cls;
};

HostModel pull_model(const Forest& f) {
  hipStream_t s = library_stream(f.device);
  HostModel m;
  const int64_t n = f.n_nodes;
  m.ts.resize(f.n_trees + 1);
  DTE_HIP(hipMemcpyAsync(m.ts.data(), f.tree_start.p, 8 * (f.n_trees + 1), hipMemcpyDeviceToHost, s));
  m.feature.resize(n);
  m.left.resize(n);
  m.thr.resize(n);
  m.value.resize(n);
  DTE_HIP(hipMemcpyAsync(m.feature.data(), f.feature.p, 4 * n, hipMemcpyDeviceToHost, s));
  DTE_HIP(hipMemcpyAsync(m.left.data(), f.left.p, 4 * n, hipMemcpyDeviceToHost, s));
  DTE_HIP(hipMemcpyAsync(m.thr.data(), f.threshold.p, 4 * n, hipMemcpyDeviceToHost, s));
  DTE_HIP(hipMemcpyAsync(m.value.data(), f.value.p, 4 * n, hipMemcpyDeviceToHost, s));
  if (f.task == 0) {
    m.cls.resize((size_t)n * f.n_classes);
    DTE_HIP(hipMemcpyAsync(m.cls.data(), f.class_counts.p, 4 * m.cls.size(), hipMemcpyDeviceToHost, s));
  }
  DTE_HIP(hipStreamSynchronize(s));
  return m;
}

int64_t count_leaves(const std::vector<int32_t>& feature) {
  int64_t n = 0;
  for (auto v : feature) n += v < 0;
  return n;
}
}  // namespace

extern "C" {

int dte_last_error(char* buf, int64_t cap) {
  if (buf && cap > 0) {
    const size_t n = std::min<size_t>((size_t)cap - 1, t_err.size());
    std::memcpy(buf, t_err.data(), n);
    buf[n] = 0;
  }
  return t_code;
}

int dte_device_count(int32_t* count) {
  return guarded([&] {
    if (!count) fail(DTE_ERR_INVALID_ARGUMENT, "count is NULL");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    *count = n;
  });
}

int dte_params_default(int32_t algorithm, dte_params* out) {
  return guarded([&] {
    if (!out) fail(DTE_ERR_INVALID_ARGUMENT, "out is NULL");
    if (algorithm != DTE_RANDOM_FOREST && algorithm != DTE_EXTRA_TREES)
      fail(DTE_ERR_INVALID_ARGUMENT, "unknown algorithm");
    std::memset(out, 0, sizeof(*out));
    out->algorithm = algorithm;
    out->task = DTE_CLASSIFICATION;
    out->n_trees = 100;
    out->n_features_per_node = 0;
    out->max_depth = 1000;
    out->min_node_size = 5;
    out->n_classes = 2;
    out->device_id = 0;
    out->seed = algorithm == DTE_RANDOM_FOREST ? 324123u : 123512u;
    out->build_mode = DTE_BUILD_BATCHED;
  });
}

static void fit_common(const float* samples, bool on_device, const float* targets, int64_t n_rows,
                       int32_t n_features, const dte_params* params, dte_forest** out) {
  check_params(params);
  if (!samples || !targets || !out) fail(DTE_ERR_INVALID_ARGUMENT, "NULL pointer argument");
  if (n_rows <= 0 || n_features <= 0) fail(DTE_ERR_INVALID_ARGUMENT, "empty data set");
  if (n_rows > 0x7FFFFFFF) fail(DTE_ERR_INVALID_ARGUMENT, "n_rows above 2^31-1");
  set_device(params->device_id);
  hipStream_t s = library_stream(params->device_id);
  DevBuf<float> X;
  const float* Xd = samples;
  if (!on_device) {
    X.alloc((size_t)n_rows * n_features);
    DTE_HIP(hipMemcpyAsync(X.p, samples, sizeof(float) * n_rows * n_features, hipMemcpyHostToDevice, s));
    Xd = X.p;
  }
  std::unique_ptr<Forest> f(fit_forest(Xd, targets, n_rows, n_features, *params, s));
  auto* h = new dte_forest();
  h->f = std::move(*f);
  *out = h;
}

int dte_fit(const float* samples, const float* targets, int64_t n_rows, int32_t n_features,
            const dte_params* params, dte_forest** out) {
  return guarded([&] { fit_common(samples, false, targets, n_rows, n_features, params, out); });
}

int dte_fit_device(const float* samples_dev, const float* targets, int64_t n_rows, int32_t n_features,
                   const dte_params* params, dte_forest** out) {
  return guarded([&] { fit_common(samples_dev, true, targets, n_rows, n_features, params, out); });
}

int dte_autotune(const float* samples_dev, const float* targets, int64_t n_rows, int32_t n_features,
                 const dte_params* params, int32_t trial_trees, dte_params* tuned, double* trial_seconds) {
  return guarded([&] {
    check_params(params);
    if (!samples_dev || !targets || !tuned) fail(DTE_ERR_INVALID_ARGUMENT, "NULL pointer argument");
    if (n_rows <= 0 || n_features <= 0) fail(DTE_ERR_INVALID_ARGUMENT, "empty data set");
    if (n_rows > 0x7FFFFFFF) fail(DTE_ERR_INVALID_ARGUMENT, "n_rows above 2^31-1");
    if (trial_trees < 0) fail(DTE_ERR_INVALID_ARGUMENT, "trial_trees < 0");
    set_device(params->device_id);
    hipStream_t s = library_stream(params->device_id);
    const auto t_start = std::chrono::steady_clock::now();
    dte_params best = *params;
    best.build_mode = DTE_BUILD_BATCHED;
    if (!best.cut_team) best.cut_team = kDefCutTeam;
    if (!best.cut_block) best.cut_block = std::max(kDefCutBlock, best.cut_team);
    if (!best.bucket_target) best.bucket_target = kDefBucketTarget;
    if (!best.cut_ert) best.cut_ert = kDefCutErt;
    best.n_trees = std::min(params->n_trees, trial_trees > 0 ? trial_trees : 16);
    auto once = [&](const dte_params& q) {
      const auto t0 = std::chrono::steady_clock::now();
      std::unique_ptr<Forest> f(fit_forest(samples_dev, targets, n_rows, n_features, q, s));
      DTE_HIP(hipStreamSynchronize(s));
      return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    };
    auto timed = [&](const dte_params& q) { return std::min(once(q), once(q)); };  // best of two
    (void)once(best);  // warm-up: device scratch, code objects
    double bt = timed(best);
    struct Knob {
      int32_t dte_params::*field;
      std::vector<int32_t> values;
    };
    std::vector<Knob> knobs;
    if (params->algorithm == DTE_RANDOM_FOREST) {
      knobs.push_back({&dte_params::bucket_target, {256, 320, 384, 448}});
      knobs.push_back({&dte_params::cut_team, {256, 384, 512}});
      knobs.push_back({&dte_params::cut_block, {1536, 2048}});
    } else {
      knobs.push_back({&dte_params::cut_ert, {1024, 4096, 16384, 65536}});
    }
    for (const Knob& k : knobs) {
      const int32_t cur = best.*k.field;
      for (int32_t v : k.values) {
        if (v == cur) continue;
        dte_params q = best;
        q.*k.field = v;
        if (q.cut_block < q.cut_team) continue;
        const double t = timed(q);
        if (t < 0.97 * bt) {
          best = q;
          bt = t;
        }
      }
    }
    *tuned = *params;
    tuned->cut_team = best.cut_team;
    tuned->cut_block = best.cut_block;
    tuned->bucket_target = best.bucket_target;
    tuned->cut_ert = best.cut_ert;
    if (trial_seconds)
      *trial_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
  });
}

int dte_fit_many(int32_t n, const float* const* samples_dev, const float* const* targets, const int64_t* n_rows,
                 const int32_t* n_features, const dte_params* params, int32_t max_concurrent, dte_forest** out) {
  return guarded([&] {
    if (n < 0) fail(DTE_ERR_INVALID_ARGUMENT, "n < 0");
    if (n == 0) return;
    if (!samples_dev || !targets || !n_rows || !n_features || !params || !out)
      fail(DTE_ERR_INVALID_ARGUMENT, "NULL pointer argument");
    for (int32_t i = 0; i < n; ++i) {
      check_params(&params[i]);
      if (!samples_dev[i] || !targets[i]) fail(DTE_ERR_INVALID_ARGUMENT, "NULL pointer argument");
      if (n_rows[i] <= 0 || n_features[i] <= 0) fail(DTE_ERR_INVALID_ARGUMENT, "empty data set");
      if (n_rows[i] > 0x7FFFFFFF) fail(DTE_ERR_INVALID_ARGUMENT, "n_rows above 2^31-1");
      if (params[i].device_id != params[0].device_id)
        fail(DTE_ERR_INVALID_ARGUMENT, "all fits of one call run on one device");
      out[i] = nullptr;
    }
    const int dev = params[0].device_id;
    set_device(dev);
    // Extra-trees forests of one task, batched: one joint build whose level
    // steps serve the open nodes of every forest at once.
    bool joint = n >= 2;
    for (int32_t i = 0; i < n; ++i)
      joint = joint && params[i].algorithm == DTE_EXTRA_TREES && params[i].build_mode == DTE_BUILD_BATCHED &&
              params[i].task == params[0].task;
    if (joint) {
      std::vector<Forest*> fs = fit_joint(samples_dev, targets, n_rows, n_features, params, n, library_stream(dev));
      for (int32_t i = 0; i < n; ++i) {
        auto* h = new dte_forest();
        h->f = std::move(*fs[i]);
        delete fs[i];
        out[i] = h;
      }
      return;
    }
    const int W = std::max(1, std::min<int>(max_concurrent > 0 ? max_concurrent : 8, n));
    std::vector<hipStream_t> streams(W);
    for (int k = 0; k < W; ++k) streams[k] = worker_stream(dev, k);
    std::vector<std::unique_ptr<Forest>> res(n);
    std::atomic<int> next{0};
    std::atomic<bool> failed{false};
    std::mutex emu;
    int ecode = 0;
    std::string emsg;
    auto work = [&](int k) {
      try {
        DTE_HIP(hipSetDevice(dev));
        for (;;) {
          const int i = next.fetch_add(1);
          if (i >= n || failed.load()) break;
          res[i].reset(fit_forest(samples_dev[i], targets[i], n_rows[i], n_features[i], params[i], streams[k]));
        }
      } catch (const DteException& e) {
        std::lock_guard<std::mutex> lk(emu);
        if (!ecode) {
          ecode = e.code;
          emsg = e.msg;
        }
        failed = true;
      } catch (const std::exception& e) {
        std::lock_guard<std::mutex> lk(emu);
        if (!ecode) {
          ecode = DTE_ERR_INTERNAL;
          emsg = e.what();
        }
        failed = true;
      }
    };
    std::vector<std::thread> th;
    for (int k = 1; k < W; ++k) th.emplace_back(work, k);
    work(0);
    for (auto& t : th) t.join();
    if (ecode) fail(ecode, emsg);
    for (int32_t i = 0; i < n; ++i) {
      auto* h = new dte_forest();
      h->f = std::move(*res[i]);
      out[i] = h;
    }
  });
}

int dte_forest_free(dte_forest* forest) {
  return guarded([&] {
    if (forest) {
      (void)hipSetDevice(forest->f.device);
      delete forest;
    }
  });
}


int dte_forest_info_get(const dte_forest* forest, dte_forest_info* out) {
  return guarded([&] {
    if (!forest || !out) fail(DTE_ERR_INVALID_ARGUMENT, "NULL pointer argument");
    Forest& f = const_cast<dte_forest*>(forest)->f;
    DTE_HIP(hipSetDevice(f.device));
    if (f.n_leaves < 0) {  // counted once from the node records, then kept
      std::vector<int32_t> feat(f.n_nodes);
      DTE_HIP(hipMemcpy(feat.data(), f.feature.p, 4 * f.n_nodes, hipMemcpyDeviceToHost));
      f.n_leaves = count_leaves(feat);
    }
    std::memset(out, 0, sizeof(*out));
    out->n_trees = f.n_trees;
    out->n_classes = f.n_classes;
    out->n_features = f.n_features;
    out->task = f.task;
    out->algorithm = f.algorithm;
    out->fixed_shift = f.fixed_shift;
    out->n_nodes = f.n_nodes;
    out->n_leaves = f.n_leaves;
    out->n_probabilities = out->n_leaves * (f.task == 0 ? f.n_classes : 1);
    out->device_bytes = (int64_t)(f.pnodes.bytes() + f.tree_start.bytes() + f.feature.bytes() + f.left.bytes() +
                                  f.depth.bytes() + f.count.bytes() + f.n_entries.bytes() + f.threshold.bytes() +
                                  f.score.bytes() + f.value.bytes() + f.rowhash.bytes() +
                                  f.class_counts.bytes() + f.sum_fixed.bytes());
  });
}

int dte_forest_export(const dte_forest* forest, dte_node* nodes, int64_t nodes_cap, float* probs,
                      int64_t probs_cap) {
  return guarded([&] {
    if (!forest || !nodes || !probs) fail(DTE_ERR_INVALID_ARGUMENT, "NULL pointer argument");
    const Forest& f = forest->f;
    DTE_HIP(hipSetDevice(f.device));
    HostModel m = pull_model(f);
    const int64_t n = f.n_nodes;
    const int T = f.n_trees;
    const int P = f.task == 0 ? f.n_classes : 1;
    if (nodes_cap < n) fail(DTE_ERR_INVALID_ARGUMENT, "nodes_cap too small");
    if (probs_cap < count_leaves(m.feature) * P) fail(DTE_ERR_INVALID_ARGUMENT, "probs_cap too small");
    // canonical index -> export index: roots first (gpudte_algorithm.cc:349-357),
    // then every tree's remaining nodes in level order.
    std::vector<int64_t> ex(n);
    int64_t next = T;
    for (int t = 0; t < T; ++t) {
      ex[m.ts[t]] = t;
      for (int64_t g = m.ts[t] + 1; g < m.ts[t + 1]; ++g) ex[g] = next++;
    }
    std::vector<int64_t> inv(n);
    for (int64_t g = 0; g < n; ++g) inv[ex[g]] = g;
    int64_t pcur = 0;
    for (int64_t e = 0; e < n; ++e) {
      const int64_t g = inv[e];
      dte_node& d = nodes[e];
      if (m.feature[g] >= 0) {
        d.child_start = (int32_t)ex[m.left[g]];
        d.child_count = 2;
        d.probability_start = -1;
        d.attribute = m.feature[g];
        d.split_point = m.thr[g];
      } else {
        d.child_start = -1;
        d.child_count = 0;
        d.probability_start = (int32_t)pcur;
        d.attribute = -1;
        d.split_point = -1.0f;
        if (f.task == 0)
          for (int c = 0; c < P; ++c) probs[pcur++] = (float)m.cls[g * P + c];
        else
          probs[pcur++] = m.value[g];
      }
    }
  });
}

int dte_forest_import(const dte_node* nodes, int64_t n_nodes, const float* probs, int64_t n_probs, int32_t n_trees,
                      int32_t n_classes, int32_t n_features, int32_t task, int32_t device_id, dte_forest** out) {
  return guarded([&] {
    if (!nodes || !probs || !out) fail(DTE_ERR_INVALID_ARGUMENT, "NULL pointer argument");
    if (n_trees <= 0 || n_nodes < n_trees) fail(DTE_ERR_INVALID_ARGUMENT, "bad tree / node counts");
    if (task != 0 && task != 1) fail(DTE_ERR_INVALID_ARGUMENT, "unknown task");
    const int P = task == 0 ? n_classes : 1;
    if (P <= 0) fail(DTE_ERR_INVALID_ARGUMENT, "n_classes must be positive");
    // Canonicalise: per tree, breadth-first from root t, children in order.
    std::vector<int64_t> order;
    order.reserve(n_nodes);
    std::vector<int64_t> ts(n_trees + 1, 0);
    std::vector<uint8_t> seen(n_nodes, 0);
    for (int t = 0; t < n_trees; ++t) {
      ts[t] = (int64_t)order.size();
      size_t head = order.size();
      order.push_back(t);
      seen[t] = 1;
      while (head < order.size()) {
        const dte_node& d = nodes[order[head++]];
        if (d.child_count == 0) {
          if (d.probability_start < 0 || (int64_t)d.probability_start + P > n_probs)
            fail(DTE_ERR_INVALID_ARGUMENT, "leaf probability index out of range");
          continue;
        }
        if (d.child_count != 2) fail(DTE_ERR_INVALID_ARGUMENT, "child_count must be 0 or 2");
        if (d.attribute < 0 || d.attribute >= n_features) fail(DTE_ERR_INVALID_ARGUMENT, "attribute out of range");
        for (int c = 0; c < 2; ++c) {
          const int64_t ch = (int64_t)d.child_start + c;
          if (ch < 0 || ch >= n_nodes || seen[ch]) fail(DTE_ERR_INVALID_ARGUMENT, "bad child link");
          seen[ch] = 1;
          order.push_back(ch);
        }
      }
    }
    ts[n_trees] = (int64_t)order.size();
    const int64_t n = (int64_t)order.size();
    std::vector<int64_t> canon(n_nodes, -1);
    for (int64_t g = 0; g < n; ++g) canon[order[g]] = g;
    std::vector<int32_t> feature(n), left(n), depth(n, 0);
    std::vector<float> thr(n), value(n, 0.0f);
    std::vector<uint32_t> cls(task == 0 ? (size_t)n * P : 0, 0);
    for (int64_t g = 0; g < n; ++g) {
      const dte_node& d = nodes[order[g]];
      if (d.child_count == 0) {
        feature[g] = -1;
        left[g] = -1;
        thr[g] = -1.0f;
        if (task == 0)
          for (int c = 0; c < P; ++c) {
            const float v = probs[d.probability_start + c];
            cls[g * P + c] = v > 0 ? (uint32_t)v : 0u;
          }
        else
          value[g] = probs[d.probability_start];
      } else {
        feature[g] = d.attribute;
        thr[g] = d.split_point;
        left[g] = (int32_t)canon[d.child_start];
      }
    }
    for (int t = 0; t < n_trees; ++t)
      for (int64_t g = ts[t]; g < ts[t + 1]; ++g)
        if (feature[g] >= 0) depth[left[g]] = depth[left[g] + 1] = depth[g] + 1;
    set_device(device_id);
    hipStream_t s = library_stream(device_id);
    auto* h = new dte_forest();
    std::unique_ptr<dte_forest> guard(h);
    Forest& f = h->f;
    f.device = device_id;
    f.task = task;
    f.n_classes = P;
    f.n_features = n_features;
    f.n_trees = n_trees;
    f.n_nodes = n;
    f.n_leaves = count_leaves(feature);
    f.has_stats = false;
    f.tree_start.alloc(n_trees + 1);
    f.feature.alloc(n);
    f.left.alloc(n);
    f.depth.alloc(n);
    f.threshold.alloc(n);
    f.value.alloc(n);
    f.class_counts.alloc(task == 0 ? (size_t)n * P : 1);
    DTE_HIP(hipMemcpyAsync(f.tree_start.p, ts.data(), 8 * ts.size(), hipMemcpyHostToDevice, s));
    DTE_HIP(hipMemcpyAsync(f.feature.p, feature.data(), 4 * n, hipMemcpyHostToDevice, s));
    DTE_HIP(hipMemcpyAsync(f.left.p, left.data(), 4 * n, hipMemcpyHostToDevice, s));
    DTE_HIP(hipMemcpyAsync(f.depth.p, depth.data(), 4 * n, hipMemcpyHostToDevice, s));
    DTE_HIP(hipMemcpyAsync(f.threshold.p, thr.data(), 4 * n, hipMemcpyHostToDevice, s));
    DTE_HIP(hipMemcpyAsync(f.value.p, value.data(), 4 * n, hipMemcpyHostToDevice, s));
    if (task == 0) DTE_HIP(hipMemcpyAsync(f.class_counts.p, cls.data(), 4 * cls.size(), hipMemcpyHostToDevice, s));
    build_predict_nodes(f, s);
    DTE_HIP(hipStreamSynchronize(s));
    *out = guard.release();
  });
}

// ------------------------------------------------------- forest files --
extern "C++" {
namespace {
constexpr char kForestMagic[8] = {'D', 'T', 'E', 'F', 'R', 'S', 'T', '1'};
constexpr uint32_t kForestVersion = 1;

uint64_t fnv1a64(const uint8_t* p, size_t n, uint64_t h = 0xcbf29ce484222325ull) {
  for (size_t i = 0; i < n; ++i) {
    h ^= p[i];
    h *= 0x100000001b3ull;
  }
  return h;
}

struct ForestHeader {
  int32_t n_trees, n_classes, n_features, task, algorithm, fixed_shift, has_stats, tree_offset;
  int64_t n_nodes, fit_rows;
  uint64_t seed;
};
constexpr size_t kForestHeaderBytes = 8 + 4 + 8 * 4 + 3 * 8;

// Bytes of the arrays that follow the header.
size_t forest_payload(const ForestHeader& h) {
  const size_t n = (size_t)h.n_nodes;
  size_t b = 8 * ((size_t)h.n_trees + 1) + n * (4 + 4 + 4 + 4 + 4);
  if (h.task == 0) b += 4 * n * (size_t)h.n_classes;
  if (h.has_stats) b += n * (4 + 4 + 4 + 8 + 8);
  return b;
}

struct Writer {
  std::vector<uint8_t> buf;
  void put(const void* p, size_t n) {
    const uint8_t* c = (const uint8_t*)p;
    buf.insert(buf.end(), c, c + n);
  }
  template <typename T>
  void put(const T& v) { put(&v, sizeof(T)); }
};

struct Reader {
  const uint8_t* p;
  size_t n, off = 0;
  void get(void* dst, size_t k) {
    if (off + k > n) fail(DTE_ERR_INVALID_ARGUMENT, "forest file is truncated");
    std::memcpy(dst, p + off, k);
    off += k;
  }
  template <typename T>
  T get() {
    T v;
    get(&v, sizeof(T));
    return v;
  }
};
}  // namespace
}  // extern "C++"

int dte_forest_save(const dte_forest* forest, const char* path) {
  return guarded([&] {
    if (!forest || !path) fail(DTE_ERR_INVALID_ARGUMENT, "NULL pointer argument");
    const Forest& f = forest->f;
    DTE_HIP(hipSetDevice(f.device));
    hipStream_t s = library_stream(f.device);
    const int64_t n = f.n_nodes;
    ForestHeader h{f.n_trees, f.n_classes, f.n_features, f.task, f.algorithm, f.fixed_shift, f.has_stats ? 1 : 0,
                   f.tree_offset, n, f.fit_rows, f.seed};
    std::vector<uint8_t> arrays(forest_payload(h));
    size_t off = 0;
    auto cp = [&](const void* src, size_t bytes) {
      DTE_HIP(hipMemcpyAsync(arrays.data() + off, src, bytes, hipMemcpyDeviceToHost, s));
      off += bytes;
    };
    cp(f.tree_start.p, 8 * ((size_t)f.n_trees + 1));
    cp(f.feature.p, 4 * n);
    cp(f.threshold.p, 4 * n);
    cp(f.left.p, 4 * n);
    cp(f.depth.p, 4 * n);
    cp(f.value.p, 4 * n);
    if (f.task == 0) cp(f.class_counts.p, 4 * (size_t)n * f.n_classes);
    if (f.has_stats) {
      cp(f.count.p, 4 * n);
      cp(f.n_entries.p, 4 * n);
      cp(f.score.p, 4 * n);
      cp(f.rowhash.p, 8 * n);
      cp(f.sum_fixed.p, 8 * n);
    }
    DTE_HIP(hipStreamSynchronize(s));
    Writer w;
    w.put(kForestMagic, 8);
    w.put(kForestVersion);
    w.put(h.n_trees);
    w.put(h.n_classes);
    w.put(h.n_features);
    w.put(h.task);
    w.put(h.algorithm);
    w.put(h.fixed_shift);
    w.put(h.has_stats);
    w.put(h.tree_offset);
    w.put(h.n_nodes);
    w.put(h.fit_rows);
    w.put(h.seed);
    w.put(arrays.data(), arrays.size());
    w.put(fnv1a64(w.buf.data(), w.buf.size()));
    FILE* fp = std::fopen(path, "wb");
    if (!fp) fail(DTE_ERR_IO, std::string("cannot open for writing: ") + path);
    const size_t wrote = std::fwrite(w.buf.data(), 1, w.buf.size(), fp);
    const int closed = std::fclose(fp);
    if (wrote != w.buf.size() || closed != 0) fail(DTE_ERR_IO, std::string("write failed: ") + path);
  });
}

int dte_forest_load(const char* path, int32_t device_id, dte_forest** out) {
  return guarded([&] {
    if (!path || !out) fail(DTE_ERR_INVALID_ARGUMENT, "NULL pointer argument");
    *out = nullptr;
    std::vector<uint8_t> file;
    {
      FILE* fp = std::fopen(path, "rb");
      if (!fp) fail(DTE_ERR_IO, std::string("cannot open: ") + path);
      uint8_t chunk[1 << 16];
      size_t k;
      while ((k = std::fread(chunk, 1, sizeof(chunk), fp)) > 0) file.insert(file.end(), chunk, chunk + k);
      const bool err = std::ferror(fp) != 0;
      std::fclose(fp);
      if (err) fail(DTE_ERR_IO, std::string("read failed: ") + path);
    }
    if (file.size() < kForestHeaderBytes + 8) fail(DTE_ERR_INVALID_ARGUMENT, "not a forest file (too short)");
    if (std::memcmp(file.data(), kForestMagic, 8) != 0) fail(DTE_ERR_INVALID_ARGUMENT, "not a forest file (magic)");
    uint64_t stored;
    std::memcpy(&stored, file.data() + file.size() - 8, 8);
    if (fnv1a64(file.data(), file.size() - 8) != stored)
      fail(DTE_ERR_INVALID_ARGUMENT, "forest file checksum mismatch");
    Reader r{file.data(), file.size() - 8, 8};
    if (r.get<uint32_t>() != kForestVersion) fail(DTE_ERR_INVALID_ARGUMENT, "unsupported forest file version");
    ForestHeader h;
    h.n_trees = r.get<int32_t>();
    h.n_classes = r.get<int32_t>();
    h.n_features = r.get<int32_t>();
    h.task = r.get<int32_t>();
    h.algorithm = r.get<int32_t>();
    h.fixed_shift = r.get<int32_t>();
    h.has_stats = r.get<int32_t>();
    h.tree_offset = r.get<int32_t>();
    h.n_nodes = r.get<int64_t>();
    h.fit_rows = r.get<int64_t>();
    h.seed = r.get<uint64_t>();
    if (h.n_trees <= 0 || h.n_nodes < h.n_trees || h.n_features <= 0 || (h.task != 0 && h.task != 1) ||
        (h.algorithm != DTE_RANDOM_FOREST && h.algorithm != DTE_EXTRA_TREES) || (h.has_stats != 0 && h.has_stats != 1) ||
        (h.task == 0 && (h.n_classes <= 0 || h.n_classes > 32)) || (h.task == 1 && h.n_classes != 1) ||
        h.n_nodes > (int64_t)file.size() || h.tree_offset < 0 || h.fit_rows < 0)
      fail(DTE_ERR_INVALID_ARGUMENT, "forest file header out of range");
    if (r.off + forest_payload(h) != r.n) fail(DTE_ERR_INVALID_ARGUMENT, "forest file size does not match its header");
    const int64_t n = h.n_nodes;
    const int T = h.n_trees;
    std::vector<int64_t> ts(T + 1);
    std::vector<int32_t> feature(n), left(n), depth(n);
    r.get(ts.data(), 8 * ts.size());
    r.get(feature.data(), 4 * n);
    const size_t thr_off = r.off;
    r.off += 4 * n;
    r.get(left.data(), 4 * n);
    r.get(depth.data(), 4 * n);
    // structure: trees tile [0, n); children adjacent inside their tree
    if (ts[0] != 0 || ts[T] != n) fail(DTE_ERR_INVALID_ARGUMENT, "forest file tree table out of range");
    for (int t = 0; t < T; ++t) {
      if (ts[t + 1] <= ts[t]) fail(DTE_ERR_INVALID_ARGUMENT, "forest file tree table out of range");
      if (depth[ts[t]] != 0) fail(DTE_ERR_INVALID_ARGUMENT, "forest file root depth");
      for (int64_t g = ts[t]; g < ts[t + 1]; ++g) {
        if (feature[g] < 0) {
          if (feature[g] != -1) fail(DTE_ERR_INVALID_ARGUMENT, "forest file feature out of range");
          continue;
        }
        if (feature[g] >= h.n_features) fail(DTE_ERR_INVALID_ARGUMENT, "forest file feature out of range");
        if (left[g] <= g || (int64_t)left[g] + 1 >= ts[t + 1])
          fail(DTE_ERR_INVALID_ARGUMENT, "forest file child link out of range");
        if (depth[left[g]] != depth[g] + 1 || depth[left[g] + 1] != depth[g] + 1)
          fail(DTE_ERR_INVALID_ARGUMENT, "forest file depths inconsistent");
      }
    }
    set_device(device_id);
    hipStream_t s = library_stream(device_id);
    auto* hd = new dte_forest();
    std::unique_ptr<dte_forest> guard(hd);
    Forest& f = hd->f;
    f.device = device_id;
    f.task = h.task;
    f.n_classes = h.n_classes;
    f.n_features = h.n_features;
    f.n_trees = T;
    f.fixed_shift = h.fixed_shift;
    f.algorithm = h.algorithm;
    f.n_nodes = n;
    f.n_leaves = count_leaves(feature);
    f.has_stats = h.has_stats != 0;
    f.seed = h.seed;
    f.tree_offset = h.tree_offset;
    f.fit_rows = h.fit_rows;
    const uint8_t* base = file.data();
    auto up = [&](auto& buf, size_t count, size_t src_off) {
      buf.alloc(count);
      DTE_HIP(hipMemcpyAsync(buf.p, base + src_off, count * sizeof(*buf.p), hipMemcpyHostToDevice, s));
    };
    size_t o = kForestHeaderBytes;
    up(f.tree_start, (size_t)T + 1, o);
    o += 8 * ((size_t)T + 1);
    up(f.feature, n, o);
    o += 4 * n;
    up(f.threshold, n, thr_off);
    o += 4 * n;
    up(f.left, n, o);
    o += 4 * n;
    up(f.depth, n, o);
    o += 4 * n;
    up(f.value, n, o);
    o += 4 * n;
    if (h.task == 0) {
      up(f.class_counts, (size_t)n * h.n_classes, o);
      o += 4 * (size_t)n * h.n_classes;
    } else {
      f.class_counts.alloc(1);
    }
    if (f.has_stats) {
      up(f.count, n, o);
      o += 4 * n;
      up(f.n_entries, n, o);
      o += 4 * n;
      up(f.score, n, o);
      o += 4 * n;
      up(f.rowhash, n, o);
      o += 8 * n;
      up(f.sum_fixed, n, o);
      o += 8 * n;
    }
    build_predict_nodes(f, s);
    DTE_HIP(hipStreamSynchronize(s));
    *out = guard.release();
  });
}

int dte_forest_canonical(const dte_forest* forest, int64_t* tree_start, int32_t* feature, float* threshold,
                         int32_t* left, int32_t* depth, int32_t* count, int32_t* n_entries, float* score,
                         uint64_t* rowhash, uint32_t* class_counts, int64_t* sum_fixed, float* value) {
  return guarded([&] {
    if (!forest) fail(DTE_ERR_INVALID_ARGUMENT, "forest is NULL");
    const Forest& f = forest->f;
    DTE_HIP(hipSetDevice(f.device));
    hipStream_t s = library_stream(f.device);
    const int64_t n = f.n_nodes;
    auto cp = [&](void* dst, const void* src, size_t bytes) {
      if (dst && src) DTE_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
    };
    cp(tree_start, f.tree_start.p, 8 * (f.n_trees + 1));
    cp(feature, f.feature.p, 4 * n);
    cp(threshold, f.threshold.p, 4 * n);
    cp(left, f.left.p, 4 * n);
    cp(depth, f.depth.p, 4 * n);
    cp(value, f.value.p, 4 * n);
    if (f.has_stats) {
      cp(count, f.count.p, 4 * n);
      cp(n_entries, f.n_entries.p, 4 * n);
      cp(score, f.score.p, 4 * n);
      cp(rowhash, f.rowhash.p, 8 * n);
      cp(sum_fixed, f.sum_fixed.p, 8 * n);
    }
    if (f.task == 0) cp(class_counts, f.class_counts.p, 4 * (size_t)n * f.n_classes);
    DTE_HIP(hipStreamSynchronize(s));
  });
}

int dte_oob_evaluate(const dte_forest* forest, const float* samples_dev, const float* targets, int64_t n_rows,
                     int32_t permuted_feature, double* out) {
  return guarded([&] {
    if (!forest || !samples_dev || !targets || !out) fail(DTE_ERR_INVALID_ARGUMENT, "NULL pointer argument");
    const Forest& f = forest->f;
    if (f.algorithm != DTE_RANDOM_FOREST || f.fit_rows <= 0)
      fail(DTE_ERR_INVALID_ARGUMENT, "out-of-bag evaluation needs a random forest fitted by this library");
    if (n_rows != f.fit_rows) fail(DTE_ERR_INVALID_ARGUMENT, "n_rows differs from the fit's rows");
    if (permuted_feature < -1 || permuted_feature >= f.n_features)
      fail(DTE_ERR_INVALID_ARGUMENT, "permuted_feature out of range");
    set_device(f.device);
    oob_evaluate(f, samples_dev, targets, n_rows, permuted_feature, out, library_stream(f.device));
  });
}

int dte_release_cached_memory(int64_t* bytes_released) {
  return guarded([&] {
    const size_t b = pool_trim();
    if (bytes_released) *bytes_released = (int64_t)b;
  });
}

int dte_predict(const dte_forest* forest, const float* samples, int64_t n_rows, float* out) {
  return guarded([&] {
    if (!forest || !samples || !out) fail(DTE_ERR_INVALID_ARGUMENT, "NULL pointer argument");
    if (n_rows < 0) fail(DTE_ERR_INVALID_ARGUMENT, "n_rows < 0");
    const Forest& f = forest->f;
    if (n_rows == 0) return;
    set_device(f.device);
    hipStream_t s = library_stream(f.device);
    DevBuf<float> X((size_t)n_rows * f.n_features);
    DTE_HIP(hipMemcpyAsync(X.p, samples, sizeof(float) * n_rows * f.n_features, hipMemcpyHostToDevice, s));
    if (f.task == 0) {
      DevBuf<int32_t> v((size_t)n_rows * f.n_classes);
      predict_forest(f, X.p, n_rows, n_rows, 0, v.p, nullptr, s);
      std::vector<int32_t> h((size_t)n_rows * f.n_classes);
      DTE_HIP(hipMemcpyAsync(h.data(), v.p, 4 * h.size(), hipMemcpyDeviceToHost, s));
      DTE_HIP(hipStreamSynchronize(s));
      for (size_t i = 0; i < h.size(); ++i) out[i] = (float)h[i];
    } else {
      DevBuf<float> v(n_rows);
      predict_forest(f, X.p, n_rows, n_rows, 0, nullptr, v.p, s);
      DTE_HIP(hipMemcpyAsync(out, v.p, 4 * n_rows, hipMemcpyDeviceToHost, s));
      DTE_HIP(hipStreamSynchronize(s));
    }
  });
}

int dte_predict_noalloc(const dte_forest* forest, const float* samples, int64_t n_rows, float* out) {
  return guarded([&] {
    if (!forest || !samples || !out) fail(DTE_ERR_INVALID_ARGUMENT, "NULL pointer argument");
    if (n_rows < 0) fail(DTE_ERR_INVALID_ARGUMENT, "n_rows < 0");
    const Forest& f = forest->f;
    if (n_rows == 0) return;
    const int64_t W = f.task == 0 ? f.n_classes : 1;
    const int64_t F = f.n_features;
    // samples (256-byte aligned) then results, inside the scratch
    if (n_rows > (int64_t)(kNoAllocScratchBytes / 4)) fail(DTE_ERR_INVALID_ARGUMENT, "dte_predict_noalloc: too many rows");
    const int64_t x_bytes = (4 * n_rows * F + 255) / 256 * 256;
    const int64_t o_bytes = 4 * n_rows * W;
    if (x_bytes + o_bytes > (int64_t)kNoAllocScratchBytes)
      fail(DTE_ERR_INVALID_ARGUMENT,
           "dte_predict_noalloc: n_rows x (n_features + result width) x 4 bytes = " + std::to_string(x_bytes + o_bytes) +
               " exceeds its " + std::to_string(kNoAllocScratchBytes) + "-byte device scratch; use dte_predict");
    if (f.device >= kMaxScratchDevices) fail(DTE_ERR_NO_DEVICE, "dte_predict_noalloc: device id above 63");
    set_device(f.device);
    char* base = static_cast<char*>(noalloc_scratch(f.device));
    std::lock_guard<std::mutex> lk(g_scratch_mu[f.device]);
    hipStream_t s = library_stream(f.device);
    float* X = reinterpret_cast<float*>(base);
    void* o = base + x_bytes;
    DTE_HIP(hipMemcpyAsync(X, samples, 4 * n_rows * F, hipMemcpyHostToDevice, s));
    if (f.task == 0) {
      predict_forest(f, X, n_rows, n_rows, 0, static_cast<int32_t*>(o), nullptr, s);  // clears the votes itself
      votes_to_float(static_cast<int32_t*>(o), n_rows * W, s);
    } else {
      predict_forest(f, X, n_rows, n_rows, 0, nullptr, static_cast<float*>(o), s);
    }
    DTE_HIP(hipMemcpyAsync(out, o, o_bytes, hipMemcpyDeviceToHost, s));
    DTE_HIP(hipStreamSynchronize(s));
  });
}

int dte_predict_device(const dte_forest* forest, const float* samples_dev, int64_t n_rows, int64_t ld,
                       int32_t row_major, int32_t* votes_dev, float* sums_dev, void* stream) {
  return guarded([&] {
    if (!forest || !samples_dev) fail(DTE_ERR_INVALID_ARGUMENT, "NULL pointer argument");
    const Forest& f = forest->f;
    if (f.task == 0 && !votes_dev) fail(DTE_ERR_INVALID_ARGUMENT, "votes_dev is NULL");
    if (f.task == 1 && !sums_dev) fail(DTE_ERR_INVALID_ARGUMENT, "sums_dev is NULL");
    if (row_major ? ld < f.n_features : ld < n_rows) fail(DTE_ERR_INVALID_ARGUMENT, "leading dimension too small");
    set_device(f.device);
    hipStream_t s = stream ? (hipStream_t)stream : library_stream(f.device);
    predict_forest(f, samples_dev, n_rows, ld, row_major, votes_dev, sums_dev, s);
    if (!stream) DTE_HIP(hipStreamSynchronize(s));
  });
}

int dte_device_alloc(int32_t device_id, int64_t bytes, void** out) {
  return guarded([&] {
    if (!out || bytes < 0) fail(DTE_ERR_INVALID_ARGUMENT, "bad arguments");
    set_device(device_id);
    void* p = nullptr;
    DTE_HIP(hipMalloc(&p, bytes > 0 ? (size_t)bytes : 1));
    *out = p;
  });
}

int dte_device_free(int32_t device_id, void* ptr) {
  return guarded([&] {
    set_device(device_id);
    if (ptr) DTE_HIP(hipFree(ptr));
  });
}

int dte_memcpy_to_device(int32_t device_id, void* dst, const void* src, int64_t bytes) {
  return guarded([&] {
    if ((!dst || !src) && bytes > 0) fail(DTE_ERR_INVALID_ARGUMENT, "NULL pointer argument");
    set_device(device_id);
    if (bytes > 0) DTE_HIP(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyHostToDevice));
  });
}

int dte_memcpy_to_host(int32_t device_id, void* dst, const void* src, int64_t bytes) {
  return guarded([&] {
    if ((!dst || !src) && bytes > 0) fail(DTE_ERR_INVALID_ARGUMENT, "NULL pointer argument");
    set_device(device_id);
    if (bytes > 0) DTE_HIP(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyDeviceToHost));
  });
}

int dte_profiling(int32_t enable) {
  return guarded([&] {
    std::lock_guard<std::mutex> lk(g_prof_mu);
    g_prof_on = enable != 0;
  });
}

int dte_profile_reset(void) {
  return guarded([&] {
    std::lock_guard<std::mutex> lk(g_prof_mu);
    drain_pending();
    g_stats.clear();
  });
}

int dte_profile_query(int32_t index, char* name, int32_t name_cap, int64_t* launches, double* total_ms,
                      double* algorithmic_bytes) {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  drain_pending();
  const int n = (int)g_stats.size();
  if (index >= 0 && index < n) {
    const KernelStat& k = g_stats[index];
    if (name && name_cap > 0) {
      const size_t c = std::min<size_t>((size_t)name_cap - 1, k.name.size());
      std::memcpy(name, k.name.data(), c);
      name[c] = 0;
    }
    if (launches) *launches = k.launches;
    if (total_ms) *total_ms = k.total_ms;
    if (algorithmic_bytes) *algorithmic_bytes = k.bytes;
  }
  return n;
}

}  // extern "C"
