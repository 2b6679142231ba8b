// Forest traversal for MI355X.  Replaces gpudte_predict (lib_ensembles/
// source/gpudte.cu:311-359) and the host launch loop of
// GpuDteAlgorithm<T>::Predict (gpudte_algorithm.cc:292-311):
//
//  * nodes are 16-byte records (feature, threshold, left / leaf payload), one
//    128-bit load per visited node;
//  * classification: every (row, tree) pair is walked exactly once (FIX of
//    the truncated grid, gpudte_algorithm.cc:297-305) and votes are integer
//    atomics, so the result is order-free;
//  * regression: one thread per row sums its trees in tree order (fixed fp32
//    order instead of float atomics, gpudte.cu:354-357).
#include <hip/hip_runtime.h>

#include "dte_internal.h"

namespace dte {
namespace {

constexpr int kRowsPerBlock = 64;

__device__ __forceinline__ int walk(const PNode* __restrict__ nodes, int64_t root, const float* __restrict__ X,
                                    int64_t row, int64_t ld, int row_major, float* leaf_value) {
  PNode nd = nodes[root];
  while (nd.feature >= 0) {
    const float x = row_major ? X[row * ld + nd.feature] : X[(int64_t)nd.feature * ld + row];
    nd = nodes[nd.left + side_of(x, nd.thr)];
  }
  *leaf_value = nd.thr;
  return nd.left;
}

// Block = 64 rows x 4 tree lanes; each wave walks every 4th tree for its
// row, votes are gathered in LDS and flushed with one atomic per class.
__global__ void __launch_bounds__(kBlock) k_predict_cls(const PNode* __restrict__ nodes,
                                                        const int64_t* __restrict__ tree_start, int T,
                                                        const float* __restrict__ X, int64_t n_rows, int64_t ld,
                                                        int row_major, int C, int trees_per_block,
                                                        int32_t* __restrict__ votes) {
  extern __shared__ uint32_t s_votes[];  // kRowsPerBlock * C
  const int lane = threadIdx.x & (kRowsPerBlock - 1);
  const int group = threadIdx.x / kRowsPerBlock;
  const int64_t row = (int64_t)blockIdx.x * kRowsPerBlock + lane;
  const int t0 = blockIdx.y * trees_per_block;
  const int t1 = min(T, t0 + trees_per_block);
  for (int i = threadIdx.x; i < kRowsPerBlock * C; i += kBlock) s_votes[i] = 0;
  __syncthreads();
  if (row < n_rows) {
    for (int t = t0 + group; t < t1; t += kBlock / kRowsPerBlock) {
      float v;
      const int cls = walk(nodes, tree_start[t], X, row, ld, row_major, &v);
      atomicAdd(&s_votes[lane * C + cls], 1u);
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kRowsPerBlock * C; i += kBlock) {
    const int r = i / C;
    const int64_t gr = (int64_t)blockIdx.x * kRowsPerBlock + r;
    const uint32_t v = s_votes[i];
    if (gr < n_rows && v) atomicAdd(&votes[gr * C + (i % C)], (int32_t)v);
  }
}

__global__ void __launch_bounds__(kBlock) k_predict_reg(const PNode* __restrict__ nodes,
                                                        const int64_t* __restrict__ tree_start, int T,
                                                        const float* __restrict__ X, int64_t n_rows, int64_t ld,
                                                        int row_major, float* __restrict__ sums) {
  for (int64_t row = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; row < n_rows;
       row += (int64_t)gridDim.x * blockDim.x) {
    float acc = 0.0f;
    for (int t = 0; t < T; ++t) {
      float v;
      walk(nodes, tree_start[t], X, row, ld, row_major, &v);
      acc = acc + v;
    }
    sums[row] = acc;
  }
}

}  // namespace

void predict_forest(const Forest& f, const float* X, int64_t n_rows, int64_t ld, int row_major,
                    int32_t* votes, float* sums, hipStream_t s) {
  if (n_rows <= 0) return;
  if (f.task == 0) {
    DTE_HIP(hipMemsetAsync(votes, 0, sizeof(int32_t) * n_rows * f.n_classes, s));
    const int64_t row_blocks = (n_rows + kRowsPerBlock - 1) / kRowsPerBlock;
    // Enough tree chunks to fill the chip even for small requests.
    int chunks = 1;
    while (row_blocks * chunks < 2048 && chunks * 32 < f.n_trees) chunks *= 2;
    const int tpb = (f.n_trees + chunks - 1) / chunks;
    const dim3 grid((unsigned)row_blocks, (unsigned)((f.n_trees + tpb - 1) / tpb));
    Timer tm("predict", 0.0, s);
    hipLaunchKernelGGL(k_predict_cls, grid, dim3(kBlock), sizeof(uint32_t) * kRowsPerBlock * f.n_classes, s,
                       f.pnodes.p, f.tree_start.p, f.n_trees, X, n_rows, ld, row_major, f.n_classes, tpb, votes);
  } else {
    const int grid = (int)std::min<int64_t>((n_rows + kBlock - 1) / kBlock, 65535);
    Timer tm("predict", 0.0, s);
    hipLaunchKernelGGL(k_predict_reg, dim3(grid), dim3(kBlock), 0, s, f.pnodes.p, f.tree_start.p, f.n_trees, X,
                       n_rows, ld, row_major, sums);
  }
  DTE_HIP(hipGetLastError());
}

}  // namespace dte
