/*
 * libdte -- C ABI of the MI355X (gfx950) tree-ensemble build + predict path.
 *
 * This is the seam the reference's host orchestrator calls into.  In the
 * reference, GpuDteAlgorithm<T>::Fit / ::Predict
 * (lib_ensembles/source/gpudte_algorithm.cc:14-267, 269-331) drive the
 * device through GpuDte<T>::CallCudaKernel / CopyIterationInfo /
 * CopyDataStaticInfo (lib_ensembles/source/gpudte.h:18-24) and
 * lib_gpu::GpuDevice alloc/copy/sync (lib_gpu/include/gpu_device.h:9-19).
 * Here that whole device side sits behind the functions below: plain
 * pointers and sizes, integer status codes, opaque handles, no C++ types.
 *
 * Conventions
 *  - Every function returns DTE_OK (0) or a DTE_ERR_* code; the message of
 *    the last failure on the calling thread is read with dte_last_error().
 *    (The reference throws through CoreLib::ThrowException,
 *    lib_core/source/core_interface.cc:28-33; the C++ layer above this ABI
 *    rethrows.)
 *  - Calls are blocking and re-entrant per device: per-launch state travels
 *    as kernel arguments, not __constant__ globals (gpudte.cu:12-52).
 *  - Matrices are fp32 and feature-major, X[f * n_rows + row], exactly the
 *    layout of MlDataFrame<float>::GetSamples() (lib_parsing/source/
 *    csv_parser.cc:92-96), unless a row-major variant says otherwise.
 */
#ifndef DTE_H_
#define DTE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. */
#define DTE_OK 0
#define DTE_ERR_INVALID_ARGUMENT 1
#define DTE_ERR_HIP 2
#define DTE_ERR_OUT_OF_MEMORY 3
#define DTE_ERR_NO_DEVICE 4
#define DTE_ERR_ADMISSION 5
#define DTE_ERR_INTERNAL 6
#define DTE_ERR_NOT_FOUND 7
#define DTE_ERR_IO 8

/* Algorithms: GpuRf / GpuErt (lib_ensembles/source/gpurf.h, gpuert.h). */
#define DTE_RANDOM_FOREST 0
#define DTE_EXTRA_TREES 1

/* Tasks: AlgorithmsInterface::AlgorithmType (lib_algorithms/include/
 * algorithms_interface.h:11). */
#define DTE_CLASSIFICATION 0
#define DTE_REGRESSION 1

/* Build drivers.  BATCHED: all open nodes of every tree of a level in one
 * set of launches; once every open node fits the device-sized tiers (random
 * forest: at most cut_block entries; extra trees: from the root) each level
 * step is a replay of a captured hipGraph whose kernels read the level's
 * counts from device memory, with no host round trip between levels.
 * BATCHED_HOST: the same launches sized on the host from counts read back
 * at every level (no level graphs).  PLAIN: the reference's launch
 * structure (tree batches of <= 100, node chunks of <= 512, one block per
 * node, host round trips per chunk; gpudte_algorithm.cc:105-253).  All
 * produce identical forests. */
#define DTE_BUILD_BATCHED 0
#define DTE_BUILD_PLAIN 1
#define DTE_BUILD_BATCHED_HOST 2

/* Fit parameters.  Mirrors the int-keyed param pack of
 * EnsemblesInterface::CreateGpuRfParamPack / CreateGpuErtParamPack
 * (lib_ensembles/source/ensembles_interface.cc:28-58): kNrTrees, kAlgoType,
 * kNrFeatures, kMaxDepth, kMinNodeSize, kDevId. */
typedef struct dte_params {
  int32_t algorithm;           /* DTE_RANDOM_FOREST | DTE_EXTRA_TREES */
  int32_t task;                /* kAlgoType: DTE_CLASSIFICATION | DTE_REGRESSION */
  int32_t n_trees;             /* kNrTrees (default 100) */
  int32_t n_features_per_node; /* kNrFeatures; 0 -> round(ln F) + 1 (gpudte_algorithm.cc:579-580) */
  int32_t max_depth;           /* kMaxDepth (default 1000) */
  int32_t min_node_size;       /* kMinNodeSize (default 5) */
  int32_t n_classes;           /* MlDataFrame::GetNrTargets(); classification only */
  int32_t device_id;           /* kDevId */
  uint64_t seed;               /* RNG key; defaults 324123 (RF) / 123512 (ERT) as gpurf.cu:55, gpuert.cu:18 */
  int32_t build_mode;          /* DTE_BUILD_BATCHED (default) | DTE_BUILD_PLAIN | DTE_BUILD_BATCHED_HOST */
  int32_t tree_offset;         /* index of the first tree: the tree counter of every random draw.  Parts
                                  built with offsets 0, n0, n0+n1, ... concatenate to the forest of one
                                  build (what SplitParameterPack + AggregateModels need, gpu_algorithm.cc:
                                  17-31; the reference restarts its RNG per batch instead, D7) */
  /* Split-search tuning; 0 keeps the default, dte_autotune picks them for a
   * data set.  Results do not depend on them, only the time does. */
  int32_t cut_team;            /* RF: largest node (distinct rows) searched by one wave per draw,
                                  64..512 (default 512); smaller nodes use 16/32/64-lane segments */
  int32_t cut_block;           /* RF: largest node searched by one workgroup per draw, cut_team..2048
                                  (default 2048); larger nodes take the bucketed chunk pipeline */
  int32_t bucket_target;       /* RF: rows per key bucket aimed at in that pipeline, 64..512 (384) */
  int32_t cut_ert;             /* ERT: largest node searched by one wave per draw, 64..2^20 (4096) */
  int32_t reserved[2];         /* must be zero */
} dte_params;

/* One node of the exported model: gpuDTE_NodeHeader_Classify<float>
 * (lib_ensembles/source/gpudte_algorithm_shared.h:82-90).  Leaves have
 * child_count == 0 and point at n_classes floats (class counts) or one
 * float (regression mean) in the probability array. */
typedef struct dte_node {
  int32_t child_start;
  int32_t child_count;
  int32_t probability_start;
  int32_t attribute;
  float split_point;
} dte_node;

typedef struct dte_forest dte_forest;

typedef struct dte_forest_info {
  int32_t n_trees;
  int32_t n_classes;      /* 1 for regression */
  int32_t n_features;
  int32_t task;
  int32_t algorithm;
  int32_t fixed_shift;    /* regression: fixed-point shift of the target sums */
  int64_t n_nodes;
  int64_t n_leaves;
  int64_t n_probabilities; /* floats in the exported probability array */
  int64_t device_bytes;    /* HBM held by the forest */
} dte_forest_info;

/* Defaults of CreateGpuRfParamPack / CreateGpuErtParamPack
 * (ensembles_interface.cc:28-58). */
int dte_params_default(int32_t algorithm, dte_params* out);

/* Fit.  Replaces the device side of GpuDteAlgorithm<T>::Fit
 * (gpudte_algorithm.cc:14-267: AllocateFit, kSetupKernel, kInitTreeBatch,
 * kFindSplit, kPerformSplit, node streaming).  samples: host, feature-major
 * n_features x n_rows; targets: host, n_rows (class ids as float, or
 * regression targets).  The fit's temporary device buffers stay cached for
 * later fits on the same stream (dte_release_cached_memory frees them). */
int dte_fit(const float* samples, const float* targets, int64_t n_rows, int32_t n_features,
            const dte_params* params, dte_forest** out);

/* Fit with the samples already resident in HBM (device pointer on
 * params->device_id); targets stay on the host. */
int dte_fit_device(const float* samples_dev, const float* targets, int64_t n_rows,
                   int32_t n_features, const dte_params* params, dte_forest** out);

/* n independent fits, samples resident in HBM (as dte_fit_device); every
 * forest is bit-identical to its fit alone.  The many-forests workload of
 * the reference, which fits one model per device thread
 * (lib_algorithms/source/gpu_algorithm.cc:18-31) and otherwise one after
 * another.  When n >= 2 and every fit is extra trees, DTE_BUILD_BATCHED and
 * of one task, the forests are built jointly: one level step (a replay of a
 * captured hipGraph) serves the open nodes of every forest at once; rows,
 * features, classes, draws per node, depth and node-size limits, seeds and
 * tree offsets may all differ.  Otherwise up to max_concurrent (0: 8) host
 * threads, each with its own HIP stream, take the fits in order.  All
 * params[i].device_id must be equal.  On failure no forest is returned. */
/* Split-search tuning for one data set (the autotuner of the batched
 * build): trial fits of trial_trees trees (0: 16) of *params on these
 * samples (best of two runs each), one candidate setting of cut_team,
 * cut_block, bucket_target (RF) or cut_ert (ERT) at a time, each kept when it
 * is at least 3 % faster than the setting so far (the defaults first).
 * *tuned = *params with the chosen tuning fields; *trial_seconds (may be
 * NULL) = the time spent.  No counterpart in the reference, whose launch
 * shape is fixed (one block of 64 threads per node, gpudte_algorithm.cc:
 * 149-151). */
int dte_autotune(const float* samples_dev, const float* targets, int64_t n_rows, int32_t n_features,
                 const dte_params* params, int32_t trial_trees, dte_params* tuned, double* trial_seconds);

int dte_fit_many(int32_t n, const float* const* samples_dev, const float* const* targets, const int64_t* n_rows,
                 const int32_t* n_features, const dte_params* params, int32_t max_concurrent, dte_forest** out);

int dte_forest_info_get(const dte_forest* forest, dte_forest_info* out);

/* Export in the reference model layout: MlModel kNodeArray / kProbArray
 * (lib_models/include/models_interface.h:9-19), roots at [0, n_trees) as
 * the reference's AggregateModels produces (gpudte_algorithm.cc:333-397),
 * then every tree's other nodes in level order, children adjacent. */
int dte_forest_export(const dte_forest* forest, dte_node* nodes, int64_t nodes_cap, float* probs,
                      int64_t probs_cap);

/* Import a model in that layout (any node order whose links are
 * consistent) and make it resident for predict. */
int dte_forest_import(const dte_node* nodes, int64_t n_nodes, const float* probs, int64_t n_probs,
                      int32_t n_trees, int32_t n_classes, int32_t n_features, int32_t task,
                      int32_t device_id, dte_forest** out);

/* Out-of-bag evaluation of a random forest on its training samples (the
 * reference's kOobEstimate / kFeatureImp kernels, lib_ensembles/source/
 * gpurf.cu:290-438, which its driver never launches).  samples_dev: the
 * fit's samples in HBM (feature-major, n_rows = the fit's rows); targets:
 * host, as for dte_fit.  Every (tree, row) pair whose bootstrap left the row
 * out is routed through the tree.  permuted_feature = -1: plain routing;
 * f >= 0: every node that splits on f reads the value of a drawn row instead
 * (one counter-based draw per visit, keyed by tree, row and depth; the
 * reference draws curand % N per visit).  out[0] = pairs evaluated, out[1] =
 * pairs whose tree vote equals the label (classification; the reference's
 * oobCounts) or the sum of squared errors of the leaf values (regression;
 * the reference sums the leaf values, FIX).  Permutation importance of f =
 * accuracy(-1) - accuracy(f), or mse(f) - mse(-1). */
int dte_oob_evaluate(const dte_forest* forest, const float* samples_dev, const float* targets, int64_t n_rows,
                     int32_t permuted_feature, double* out);

/* Forest file: the canonical arrays below with a header and an FNV-1a-64
 * checksum, little-endian (format "DTEFRST1", version 1).  The reference
 * keeps models only in memory (MlModel kNodeArray / kProbArray,
 * gpudte_algorithm.cc:58-88) and rebuilds them per process; a saved forest
 * reloads into HBM bit for bit (predictions and export identical).
 * dte_forest_load checks the header, sizes, checksum and tree links before it
 * touches the device: DTE_ERR_IO for a file that cannot be read or written,
 * DTE_ERR_INVALID_ARGUMENT for one that is not a valid forest file. */
int dte_forest_save(const dte_forest* forest, const char* path);
int dte_forest_load(const char* path, int32_t device_id, dte_forest** out);

/* Canonical build arrays (tree-major, level order in each tree) for
 * parity checks.  Any output pointer may be NULL.  class_counts has
 * n_nodes * n_classes entries (classification only). */
int dte_forest_canonical(const dte_forest* forest, int64_t* tree_start, int32_t* feature,
                         float* threshold, int32_t* left, int32_t* depth, int32_t* count,
                         int32_t* n_entries, float* score, uint64_t* rowhash,
                         uint32_t* class_counts, int64_t* sum_fixed, float* value);

int dte_forest_free(dte_forest* forest);

/* Predict.  Replaces GpuDteAlgorithm<T>::Predict + kPredict
 * (gpudte_algorithm.cc:269-331, gpudte.cu:311-359).  samples: host,
 * feature-major n_features x n_rows, where n_features is the forest's
 * (dte_forest_info): n_features * n_rows floats are read.  out: n_rows x
 * n_classes vote counts
 * (classification) or n_rows sums over trees in tree order (regression;
 * mean = sum / n_trees), as MlResultData receives them. */
int dte_predict(const dte_forest* forest, const float* samples, int64_t n_rows, float* out);

/* Predict without allocating: the same inputs, outputs and results as
 * dte_predict, with the samples and votes staged in a fixed 64 MiB device
 * scratch array that the library holds on every GPU it runs on until the
 * process ends, so the call makes no hipMalloc / hipFree and no other
 * allocation of device memory.  No counterpart in the reference: it is for
 * callers that cannot allocate (code under HIP graph capture, tight serving
 * loops, harnesses that forbid allocation).  Limit: n_rows x (n_features +
 * result width) x 4 bytes (samples rounded up to 256 bytes) must be at most
 * 64 MiB, else DTE_ERR_INVALID_ARGUMENT.  Calls on one GPU share its array:
 * the library runs them one at a time (a lock per device) and the array
 * keeps the last call's contents; a call clears what must start at zero. */
int dte_predict_noalloc(const dte_forest* forest, const float* samples, int64_t n_rows, float* out);

/* Predict on device buffers.  samples_dev: feature-major with leading
 * dimension ld (>= n_rows) when row_major == 0, else row-major with leading
 * dimension ld (>= n_features).  votes_dev (n_rows x n_classes int32) or
 * sums_dev (n_rows fp32).  stream: hipStream_t or NULL for the library
 * stream. */
int dte_predict_device(const dte_forest* forest, const float* samples_dev, int64_t n_rows,
                       int64_t ld, int32_t row_major, int32_t* votes_dev, float* sums_dev,
                       void* stream);

/* Serving: forests stay resident, requests of uneven row counts are
 * batched continuously, admission by free HBM. */
typedef struct dte_server dte_server;
typedef struct dte_server_stats {
  int64_t requests_done;
  int64_t rows_done;
  int64_t batches_launched;
  int64_t requests_rejected;
  double busy_ms;          /* device time of the batches: h2d + kernel + d2h */
  int64_t hbm_in_use;
  /* Where the batches' time went (sums over finished batches, ms): host
   * staging of rows into pinned memory, the rows' host-to-device copy, the
   * traversal launches, the results' device-to-host copy, and the host
   * scatter of results into the requests. */
  double stage_ms;
  double h2d_ms;
  double kernel_ms;
  double d2h_ms;
  double scatter_ms;
  int64_t groups_launched; /* traversal launches: one per forest per batch */
} dte_server_stats;

int dte_server_create(int32_t device_id, int64_t hbm_budget_bytes, int32_t max_batch_rows,
                      dte_server** out);
int dte_server_load_forest(dte_server* server, const dte_forest* forest, int32_t* slot);
/* rows: host, row-major n_rows x n_features, where n_features is that of
 * the forest in the slot: the server copies n_rows * n_features floats.
 * A batch carries requests for any of the resident forests. */
int dte_server_submit(dte_server* server, int32_t slot, const float* rows, int64_t n_rows,
                      uint64_t* ticket);
/* done = 1 when finished; out receives n_rows x n_classes votes or n_rows sums. */
int dte_server_poll(dte_server* server, uint64_t ticket, float* out, int32_t* done);
int dte_server_wait(dte_server* server, uint64_t ticket, float* out);
int dte_server_get_stats(dte_server* server, dte_server_stats* out);
int dte_server_destroy(dte_server* server);

/* Kernel timing: HIP events recorded around every launch of the library's
 * kernels while profiling is on. */
int dte_profiling(int32_t enable);
int dte_profile_reset(void);
/* Returns the number of kernel families; fills up to cap entries. */
int dte_profile_query(int32_t index, char* name, int32_t name_cap, int64_t* launches,
                      double* total_ms, double* algorithmic_bytes);

/* Device memory helpers for callers that stage inputs in HBM themselves
 * (the reference's lib_gpu::GpuDevice AllocateMemory / CopyToDevice /
 * CopyToHost, lib_gpu/include/gpu_device.h:9-19). */
int dte_device_alloc(int32_t device_id, int64_t bytes, void** out);
int dte_device_free(int32_t device_id, void* ptr);
int dte_memcpy_to_device(int32_t device_id, void* dst, const void* src, int64_t bytes);
int dte_memcpy_to_host(int32_t device_id, void* dst, const void* src, int64_t bytes);

/* Frees the device memory a fit keeps cached for later fits on the same
 * stream (its per-level scratch; see dte_fit).  The reference allocates and
 * frees its buffers around every fit (gpudte_algorithm.cc:486-596,
 * gpudte_algorithm_shared.cc:30-52), so it has no counterpart; callers that
 * size other allocations from the free memory call this first.
 * bytes_released may be NULL. */
int dte_release_cached_memory(int64_t* bytes_released);

int dte_device_count(int32_t* count);
int dte_last_error(char* buf, int64_t cap);

#ifdef __cplusplus
}
#endif

#endif /* DTE_H_ */
