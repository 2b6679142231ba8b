// Internal declarations shared by the build, predict, serve and C-ABI units
// of libdte.  Not part of the public ABI (include/dte.h is).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <string>
#include <vector>

#include "../../include/dte.h"
#include "dte_numerics.h"

namespace dte {

// ------------------------------------------------------------- errors ----
struct Error {
  int code;
  std::string msg;
};
void set_error(int code, const std::string& msg);
// Runs fn, turning exceptions into a status code and the thread's last error.
int run_guarded(const std::function<void()>& fn);
// hipSetDevice + one-time library initialisation of that device.
void select_device(int device);
[[noreturn]] void fail(int code, const std::string& msg);

#define DTE_HIP(expr)                                                                  \
  do {                                                                                 \
    hipError_t _e = (expr);                                                            \
    if (_e != hipSuccess)                                                              \
      ::dte::fail(_e == hipErrorOutOfMemory ? DTE_ERR_OUT_OF_MEMORY : DTE_ERR_HIP,      \
                  std::string(#expr) + ": " + hipGetErrorString(_e));                  \
  } while (0)

// ------------------------------------------------------- scratch pool --
// Device blocks of a fit's temporaries are cached per (device, stream) when
// released and handed back to later requests on the same stream, so that the
// growing per-level buffers and repeated fits do not call hipMalloc / hipFree
// (which synchronise the device) inside the level loop.  Reuse on the same
// stream is ordered by the stream.  Pooling is on for the allocations of a
// thread inside a PoolScope; forests handed to the caller are allocated
// outside the pool (PoolSuspend).
struct PoolBlock {
  void* p;
  size_t bytes;
};
// Takes a block of at least `bytes` for the thread's pool stream (granted
// size in *got, owning device and stream in *dev, *owner).
void* pool_alloc(size_t bytes, size_t* got, int* dev, hipStream_t* owner);
void pool_free(void* p, size_t bytes, int dev, hipStream_t owner);
// Frees every cached block (all devices); returns the bytes released.
size_t pool_trim();
bool pool_on();  // inside a PoolScope (and not suspended) on this thread

struct PoolScope {
  bool prev_on;
  hipStream_t prev_s;
  explicit PoolScope(hipStream_t s);
  ~PoolScope();
};
struct PoolSuspend {
  bool prev_on;
  PoolSuspend();
  ~PoolSuspend();
};

// Device buffer owned by RAII (the reference frees by hand in
// GpuParams::finalize, gpudte_algorithm_shared.cc:30-52, and leaks on throw).
template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;           // elements requested
  size_t block = 0;       // bytes of the pool block (pooled buffers)
  bool pooled = false;
  int dev = 0;
  hipStream_t owner = nullptr;
  DevBuf() = default;
  explicit DevBuf(size_t count) { alloc(count); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n), block(o.block), pooled(o.pooled), dev(o.dev), owner(o.owner) {
    o.p = nullptr;
    o.n = 0;
  }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      n = o.n;
      block = o.block;
      pooled = o.pooled;
      dev = o.dev;
      owner = o.owner;
      o.p = nullptr;
      o.n = 0;
    }
    return *this;
  }
  ~DevBuf() { release(); }
  void alloc(size_t count) {
    release();
    if (count == 0) count = 1;
    if (pool_on()) {
      p = (T*)pool_alloc(count * sizeof(T), &block, &dev, &owner);
      pooled = true;
      n = count;
    } else {
      DTE_HIP(hipMalloc((void**)&p, count * sizeof(T)));
      pooled = false;
      n = count;
    }
  }
  // Grow to at least `count` elements; contents are not preserved.
  void reserve(size_t count) {
    if (count > n) alloc(count + count / 4);
  }
  void release() {
    if (p) {
      if (pooled)
        pool_free(p, block, dev, owner);
      else
        (void)hipFree(p);
    }
    p = nullptr;
    n = 0;
    block = 0;
    pooled = false;
  }
  size_t bytes() const { return n * sizeof(T); }
};

// Kernel timing of one named kernel family, collected with HIP events on the
// library's stream when profiling is on.
struct KernelStat {
  std::string name;
  int64_t launches = 0;
  double total_ms = 0.0;
  double bytes = 0.0;  // algorithmic bytes
};

void profile_record(const char* name, hipEvent_t a, hipEvent_t b, double bytes);
void profile_add_bytes(const char* name, double bytes);
void profile_count(const char* name, int64_t n);
void profile_host(const char* name, double ms);
bool profiling_enabled();

// HIP events around the launches of one kernel family on stream s (only
// while profiling is on); recorded under `name` with `bytes` algorithmic bytes.
struct Timer {
  hipEvent_t a = nullptr, b = nullptr;
  const char* name;
  double bytes;
  hipStream_t s;
  Timer(const char* nm, double by, hipStream_t st) : name(nm), bytes(by), s(st) {
    if (profiling_enabled()) {
      DTE_HIP(hipEventCreate(&a));
      DTE_HIP(hipEventCreate(&b));
      DTE_HIP(hipEventRecord(a, s));
    }
  }
  ~Timer() {
    if (a) {
      (void)hipEventRecord(b, s);
      profile_record(name, a, b, bytes);
    }
  }
};

// Predict-side node record: 16 bytes, one 128-bit load per visited node.
//   internal: feature >= 0, thr, left child index (right = left + 1)
//   leaf:     feature = -1, thr = regression leaf value, left = vote class
struct PNode {
  int32_t feature;
  float thr;
  int32_t left;
  int32_t pad;
};

// A fitted forest resident in HBM.  Node arrays are canonical: tree-major,
// level order inside each tree, children adjacent.
struct Forest {
  int device = 0;
  int task = 0;
  int n_classes = 1;
  int n_features = 0;
  int n_trees = 0;
  int fixed_shift = 0;
  int algorithm = 0;
  int64_t n_nodes = 0;
  // Leaves; -1 until known.  A built tree has one more leaf than splits
  // (every split makes two children), so a fit knows it without a pass.
  int64_t n_leaves = -1;
  // Bootstrap record of a fit (random forest): tree t drew its rows with
  // seed and global tree id tree_offset + t from fit_rows rows; fit_rows = 0
  // when unknown (imported forests).
  uint64_t seed = 0;
  int32_t tree_offset = 0;
  int64_t fit_rows = 0;
  DevBuf<int64_t> tree_start;  // n_trees + 1
  DevBuf<PNode> pnodes;        // n_nodes
  // Build diagnostics (absent for imported forests).
  bool has_stats = false;
  DevBuf<int32_t> feature, left, depth, count, n_entries;
  DevBuf<float> threshold, score, value;
  DevBuf<uint64_t> rowhash;
  DevBuf<uint32_t> class_counts;  // n_nodes * n_classes (classification)
  DevBuf<int64_t> sum_fixed;
};

struct FitInput {
  const float* X;  // device, feature-major, canonical (-0 -> +0)
  const float* y;  // device or host targets (see caller)
  int64_t n_rows;
  int32_t n_features;
};

// Entry points implemented in dte_build.hip / dte_predict.hip.
Forest* fit_forest(const float* X_dev, const float* y_host, int64_t n_rows, int32_t n_features,
                   const dte_params& p, hipStream_t stream);
void build_predict_nodes(Forest& f, hipStream_t stream);
// Several extra-trees forests of one task in one build: every level step
// serves the open nodes of all forests (see dte_fit_many).
std::vector<Forest*> fit_joint(const float* const* X_dev, const float* const* y_host, const int64_t* n_rows,
                               const int32_t* n_features, const dte_params* params, int n, hipStream_t stream);
void predict_forest(const Forest& f, const float* X_dev, int64_t n_rows, int64_t ld,
                    int row_major, int32_t* votes_dev, float* sums_dev, hipStream_t stream);
// In place: int32 vote counts -> the same values as float.
void votes_to_float(int32_t* votes, int64_t n, hipStream_t stream);
// Out-of-bag evaluation (dte_oob_evaluate): pairs evaluated, correct pairs
// (classification) or the sum of squared errors (regression).
void oob_evaluate(const Forest& f, const float* X_dev, const float* targets, int64_t n_rows, int perm_feature,
                  double out[2], hipStream_t stream);
void device_init(int device);
hipStream_t library_stream(int device);
const LnTable& host_ln_table();

constexpr int kBlock = 256;

// Defaults of the split-search tuning fields of dte_params (0 selects them).
constexpr int kDefCutTeam = 512;
constexpr int kMaxCutTeam = 512;  // capacity of the wave-per-item tier (64 lanes x 8)
constexpr int kDefCutBlock = 2048;
constexpr int kDefBucketTarget = 384;
constexpr int kDefCutErt = 4096;

}  // namespace dte

// Opaque handle of the C ABI.
struct dte_forest {
  dte::Forest f;
};
