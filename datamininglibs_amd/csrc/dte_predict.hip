// Forest traversal for MI355X.  Replaces gpudte_predict (lib_ensembles/
// source/gpudte.cu:311-359) and the host launch loop of
// GpuDteAlgorithm<T>::Predict (gpudte_algorithm.cc:292-311):
//
//  * nodes are 16-byte records (feature, threshold, left / leaf payload), one
//    128-bit load per visited node;
//  * classification: every (row, tree) pair is walked exactly once (FIX of
//    the truncated grid, gpudte_algorithm.cc:297-305) and votes are integer
//    atomics, so the result is order-free;
//  * regression: every row's trees are summed in tree order (fixed fp32
//    order instead of float atomics, gpudte.cu:354-357).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "dte_internal.h"

namespace dte {
namespace {

constexpr int kRowsPerBlock = 64;
constexpr int kWave = 64;

// One visited node = one wide load.  Left alone, the compiler splits the
// record into dword loads and sinks the threshold's load past the feature's
// LDS lookup: two memory round trips per level instead of one.  The empty
// asm statement (no instruction) takes the three fields as operands, so
// their loads stay together and merge into one.
__device__ __forceinline__ PNode load_node(const PNode* __restrict__ nodes, int64_t i) {
  const int4 v = reinterpret_cast<const int4*>(nodes)[i];
  asm volatile("" : : "v"(v.x), "v"(v.y), "v"(v.z));
  return PNode{v.x, __int_as_float(v.y), v.z, v.w};
}

__device__ __forceinline__ int walk(const PNode* __restrict__ nodes, int64_t root, const float* __restrict__ X,
                                    int64_t row, int64_t ld, int row_major, float* leaf_value) {
  PNode nd = load_node(nodes, root);
  while (nd.feature >= 0) {
    const float x = row_major ? X[row * ld + nd.feature] : X[(int64_t)nd.feature * ld + row];
    nd = load_node(nodes, nd.left + side_of(x, nd.thr));
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

// The same with the block's R rows staged in LDS first (coalesced reads of
// X, feature-major or row-major), so that every visit reads its feature
// value from LDS: block = R rows x (256 / R) tree lanes.
template <int R>
__global__ void __launch_bounds__(kBlock) k_predict_cls_lds(const PNode* __restrict__ nodes,
                                                            const int64_t* __restrict__ tree_start, int T,
                                                            const float* __restrict__ X, int64_t n_rows, int64_t ld,
                                                            int row_major, int C, int F, int trees_per_block,
                                                            int32_t* __restrict__ votes) {
  extern __shared__ uint32_t s_mem[];
  uint32_t* s_votes = s_mem;                            // R * C
  float* s_x = reinterpret_cast<float*>(s_mem + R * C);  // R rows of F + 1 floats
  const int FP = F + 1;
  const int lane = threadIdx.x % R;
  const int group = threadIdx.x / R;
  const int64_t row0 = (int64_t)blockIdx.x * R;
  const int nr = (int)min<int64_t>(R, n_rows - row0);
  const int t0 = blockIdx.y * trees_per_block;
  const int t1 = min(T, t0 + trees_per_block);
  for (int i = threadIdx.x; i < R * C; i += kBlock) s_votes[i] = 0;
  if (row_major) {
    for (int i = threadIdx.x; i < nr * F; i += kBlock) {
      const int r = i / F, f = i - r * F;
      s_x[r * FP + f] = X[(row0 + r) * ld + f];
    }
  } else {
    for (int i = threadIdx.x; i < R * F; i += kBlock) {
      const int f = i / R, r = i - f * R;
      if (r < nr) s_x[r * FP + f] = X[(int64_t)f * ld + row0 + r];
    }
  }
  __syncthreads();
  if (lane < nr) {
    const float* xr = s_x + lane * FP;
    for (int t = t0 + group; t < t1; t += kBlock / R) {
      PNode nd = load_node(nodes, tree_start[t]);
      while (nd.feature >= 0) nd = load_node(nodes, nd.left + side_of(xr[nd.feature], nd.thr));
      atomicAdd(&s_votes[lane * C + nd.left], 1u);
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < R * C; i += kBlock) {
    const int r = i / C;
    const uint32_t v = s_votes[i];
    if (r < nr && v) atomicAdd(&votes[(row0 + r) * C + (i % C)], (int32_t)v);
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

// Regression with a block's R rows staged in LDS.  Trees go in chunks of
// kRegChunk: the 256 / R tree lanes of a row write the chunk's leaf values
// to LDS, then one lane per row adds them in tree order, so every row's sum
// is formed in the same order as the one-thread-per-row kernel.
constexpr int kRegChunk = 32;
template <int R>
__global__ void __launch_bounds__(kBlock) k_predict_reg_lds(const PNode* __restrict__ nodes,
                                                            const int64_t* __restrict__ tree_start, int T,
                                                            const float* __restrict__ X, int64_t n_rows, int64_t ld,
                                                            int row_major, int F, float* __restrict__ sums) {
  extern __shared__ float s_mem_f[];
  float* s_val = s_mem_f;              // R x kRegChunk leaf values
  float* s_x = s_mem_f + R * kRegChunk;  // R rows of F + 1 floats
  const int FP = F + 1;
  const int lane = threadIdx.x % R;
  const int group = threadIdx.x / R;
  const int64_t row0 = (int64_t)blockIdx.x * R;
  const int nr = (int)min<int64_t>(R, n_rows - row0);
  if (row_major) {
    for (int i = threadIdx.x; i < nr * F; i += kBlock) {
      const int r = i / F, f = i - r * F;
      s_x[r * FP + f] = X[(row0 + r) * ld + f];
    }
  } else {
    for (int i = threadIdx.x; i < R * F; i += kBlock) {
      const int f = i / R, r = i - f * R;
      if (r < nr) s_x[r * FP + f] = X[(int64_t)f * ld + row0 + r];
    }
  }
  __syncthreads();
  const float* xr = s_x + lane * FP;
  float acc = 0.0f;
  for (int c0 = 0; c0 < T; c0 += kRegChunk) {
    const int c1 = min(T, c0 + kRegChunk);
    if (lane < nr)
      for (int t = c0 + group; t < c1; t += kBlock / R) {
        PNode nd = load_node(nodes, tree_start[t]);
        while (nd.feature >= 0) nd = load_node(nodes, nd.left + side_of(xr[nd.feature], nd.thr));
        s_val[lane * kRegChunk + (t - c0)] = nd.thr;
      }
    __syncthreads();
    if (group == 0 && lane < nr)
      for (int t = c0; t < c1; ++t) acc = acc + s_val[lane * kRegChunk + (t - c0)];
    __syncthreads();
  }
  if (group == 0 && lane < nr) sums[row0 + lane] = acc;
}

// ---------------------------------------------------- out-of-bag pass --
// Stream of the per-visit row draws of the permutation pass.
constexpr uint32_t kStreamImp = 0x00706d69u;

// In-bag bits: bit r of tree t is set iff tree t's bootstrap drew row r
// (the draws of the fit's bootstrap, k_boot_count).  One workgroup per
// (tree, range of kBagRange rows): every draw of the tree is generated, the
// range's bits are set in LDS and written out once.
constexpr int kBagThreads = 1024;
constexpr int64_t kBagRange = (int64_t)1 << 20;  // rows per workgroup: 128 KB of bits
__global__ void __launch_bounds__(kBagThreads) k_inbag_bits_lds(uint32_t* bits, int64_t N, int64_t words,
                                                                int32_t tree0, uint64_t seed) {
  __shared__ uint32_t s_bits[kBagRange / 32];
  const int t = blockIdx.x;
  const int64_t r0 = (int64_t)blockIdx.y * kBagRange;
  const int64_t nw = (min(kBagRange, N - r0) + 31) / 32;  // words of this range
  for (int64_t w = threadIdx.x; w < nw; w += kBagThreads) s_bits[w] = 0;
  __syncthreads();
  const uint32_t tg = (uint32_t)(tree0 + t);
  for (int64_t i = threadIdx.x; i < N; i += kBagThreads) {
    const int64_t r = (int64_t)(rand_u32(seed, kStreamBoot, tg, 0, (uint64_t)i) % (uint32_t)N) - r0;
    if (r >= 0 && r < kBagRange) atomicOr(&s_bits[r >> 5], 1u << (r & 31));
  }
  __syncthreads();
  uint32_t* bt = bits + (int64_t)t * words + r0 / 32;
  for (int64_t w = threadIdx.x; w < nw; w += kBagThreads) bt[w] = s_bits[w];
}

// Every (tree, row) pair whose row is out of the tree's bag is routed
// through the tree (gpurf.cu:290-353); with perm >= 0, nodes that split on
// feature perm read the value of row draw % N instead, one counter-based
// draw per visit keyed by (tree, row, depth) (gpurf.cu:356-438).
// Classification counts the pairs whose tree vote equals the label;
// regression sums the squared errors of the leaf values.
template <int TASK>
__global__ void __launch_bounds__(kBlock) k_oob(const PNode* __restrict__ nodes, const int64_t* __restrict__ tree_start,
                                                const float* __restrict__ X, int64_t N, const uint32_t* __restrict__ bits,
                                                int64_t words, const uint32_t* __restrict__ label,
                                                const float* __restrict__ y, int32_t tree0, uint64_t seed, int perm,
                                                unsigned long long* counts, double* sse) {
  const int t = blockIdx.y;
  const uint32_t tg = (uint32_t)(tree0 + t);
  const uint32_t* bt = bits + (int64_t)t * words;
  const int64_t root = tree_start[t];
  unsigned long long n_pairs = 0, n_ok = 0;
  double e2 = 0.0;
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < N; r += (int64_t)gridDim.x * blockDim.x) {
    if ((bt[r >> 5] >> (r & 31)) & 1u) continue;  // in the bag
    PNode nd = load_node(nodes, root);
    uint32_t depth = 0;
    while (nd.feature >= 0) {
      int64_t src = r;
      if (nd.feature == perm) src = rand_u32(seed, kStreamImp, tg, (uint32_t)r, depth) % (uint32_t)N;
      nd = load_node(nodes, nd.left + side_of(X[(int64_t)nd.feature * N + src], nd.thr));
      ++depth;
    }
    ++n_pairs;
    if (TASK == 0) {
      n_ok += (uint32_t)nd.left == label[r] ? 1ull : 0ull;
    } else {
      const double e = (double)nd.thr - (double)y[r];
      e2 += e * e;
    }
  }
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) {
    n_pairs += __shfl_xor(n_pairs, off, kWave);
    n_ok += __shfl_xor(n_ok, off, kWave);
    if (TASK == 1) e2 += __shfl_xor(e2, off, kWave);
  }
  if ((threadIdx.x & (kWave - 1)) == 0 && n_pairs) {
    atomicAdd(&counts[0], n_pairs);
    if (TASK == 0) atomicAdd(&counts[1], n_ok);
    else atomicAdd(sse, e2);
  }
}

// The same pairs with a block's R rows staged in LDS (feature-major X read
// coalesced), R rows x 256 / R tree lanes, trees [t0, t1) of the block's
// chunk; the permuted feature still reads its drawn row from X.
template <int TASK, int R>
__global__ void __launch_bounds__(kBlock) k_oob_lds(const PNode* __restrict__ nodes,
                                                    const int64_t* __restrict__ tree_start, int T,
                                                    const float* __restrict__ X, int64_t N, int F,
                                                    const uint32_t* __restrict__ bits, int64_t words,
                                                    const uint32_t* __restrict__ label, const float* __restrict__ y,
                                                    int32_t tree0, uint64_t seed, int perm, int trees_per_block,
                                                    unsigned long long* counts, double* sse) {
  extern __shared__ float s_x[];  // R rows of F + 1 floats
  const int FP = F + 1;
  const int lane = threadIdx.x % R;
  const int group = threadIdx.x / R;
  const int64_t row0 = (int64_t)blockIdx.x * R;
  const int nr = (int)min<int64_t>(R, N - row0);
  const int t0 = blockIdx.y * trees_per_block;
  const int t1 = min(T, t0 + trees_per_block);
  for (int i = threadIdx.x; i < R * F; i += kBlock) {
    const int f = i / R, r = i - f * R;
    if (r < nr) s_x[r * FP + f] = X[(int64_t)f * N + row0 + r];
  }
  __syncthreads();
  unsigned long long n_pairs = 0, n_ok = 0;
  double e2 = 0.0;
  if (lane < nr) {
    const int64_t r = row0 + lane;
    const float* xr = s_x + lane * FP;
    const uint32_t lab = TASK == 0 ? label[r] : 0u;
    const double yr = TASK == 1 ? (double)y[r] : 0.0;
    for (int t = t0 + group; t < t1; t += kBlock / R) {
      if ((bits[(int64_t)t * words + (r >> 5)] >> (r & 31)) & 1u) continue;  // in the bag
      const uint32_t tg = (uint32_t)(tree0 + t);
      PNode nd = load_node(nodes, tree_start[t]);
      uint32_t depth = 0;
      while (nd.feature >= 0) {
        float v;
        if (nd.feature == perm)
          v = X[(int64_t)perm * N + rand_u32(seed, kStreamImp, tg, (uint32_t)r, depth) % (uint32_t)N];
        else
          v = xr[nd.feature];
        nd = load_node(nodes, nd.left + side_of(v, nd.thr));
        ++depth;
      }
      ++n_pairs;
      if (TASK == 0) {
        n_ok += (uint32_t)nd.left == lab ? 1ull : 0ull;
      } else {
        const double e = (double)nd.thr - yr;
        e2 += e * e;
      }
    }
  }
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) {
    n_pairs += __shfl_xor(n_pairs, off, kWave);
    n_ok += __shfl_xor(n_ok, off, kWave);
    if (TASK == 1) e2 += __shfl_xor(e2, off, kWave);
  }
  if ((threadIdx.x & (kWave - 1)) == 0 && n_pairs) {
    atomicAdd(&counts[0], n_pairs);
    if (TASK == 0) atomicAdd(&counts[1], n_ok);
    else atomicAdd(sse, e2);
  }
}

__global__ void __launch_bounds__(kBlock) k_zero_u32(uint32_t* __restrict__ p, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) p[i] = 0u;
}

__global__ void __launch_bounds__(kBlock) k_votes_to_float(int32_t* __restrict__ p, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const float v = (float)p[i];
    reinterpret_cast<float*>(p)[i] = v;
  }
}

unsigned elementwise_grid(int64_t n) { return (unsigned)std::min<int64_t>((n + kBlock - 1) / kBlock, 4096); }

// Rows per block of the LDS-staged kernels: 64 while the block's rows (and
// vote counters, C) fit 48 KB, else 32, else 0 (rows read from HBM).
int staged_rows(int F, int C) {
  if (4 * 64 * (F + 1 + C) <= 48 * 1024) return 64;
  if (4 * 32 * (F + 1 + C) <= 48 * 1024) return 32;
  return 0;
}

// Tree chunks so that row blocks x chunks fill the chip even for few rows.
int tree_chunks(int64_t row_blocks, int n_trees) {
  int chunks = 1;
  while (row_blocks * chunks < 2048 && chunks * 32 < n_trees) chunks *= 2;
  return chunks;
}

}  // namespace

void oob_evaluate(const Forest& f, const float* X, const float* targets, int64_t N, int perm, double out[2],
                  hipStream_t s) {
  const int T = f.n_trees;
  if (T > 65535) fail(DTE_ERR_INVALID_ARGUMENT, "the out-of-bag pass handles at most 65535 trees");
  const int64_t words = (N + 31) / 32;
  DevBuf<uint32_t> bits((size_t)words * T);  // every word is written by k_inbag_bits_lds
  const int gx = (int)std::min<int64_t>((N + kBlock - 1) / kBlock, 1024);
  hipLaunchKernelGGL(k_inbag_bits_lds, dim3((unsigned)T, (unsigned)((N + kBagRange - 1) / kBagRange)),
                     dim3(kBagThreads), 0, s, bits.p, N, words, f.tree_offset, f.seed);
  DevBuf<uint32_t> label;
  DevBuf<float> yv;
  if (f.task == 0) {
    std::vector<uint32_t> lab(N);
    for (int64_t i = 0; i < N; ++i) {
      const int c = (int)targets[i];
      if (c < 0 || c >= f.n_classes || (float)c != targets[i]) fail(DTE_ERR_INVALID_ARGUMENT, "class label out of range");
      lab[i] = (uint32_t)c;
    }
    label.alloc(N);
    DTE_HIP(hipMemcpyAsync(label.p, lab.data(), 4 * N, hipMemcpyHostToDevice, s));
    DTE_HIP(hipStreamSynchronize(s));
  } else {
    yv.alloc(N);
    DTE_HIP(hipMemcpyAsync(yv.p, targets, 4 * N, hipMemcpyHostToDevice, s));
    DTE_HIP(hipStreamSynchronize(s));
  }
  DevBuf<unsigned long long> counts(2);
  DevBuf<double> sse(1);
  DTE_HIP(hipMemsetAsync(counts.p, 0, counts.bytes(), s));
  DTE_HIP(hipMemsetAsync(sse.p, 0, sse.bytes(), s));
  const int R = staged_rows(f.n_features, 0);
  if (R) {
    const int64_t row_blocks = (N + R - 1) / R;
    const int tpb = (T + tree_chunks(row_blocks, T) - 1) / tree_chunks(row_blocks, T);
    const dim3 grid((unsigned)row_blocks, (unsigned)((T + tpb - 1) / tpb));
    const size_t lds = sizeof(float) * (size_t)R * (f.n_features + 1);
    const uint32_t* lab = label.p;
    const float* yp = yv.p;
    const int Fn = f.n_features;
    if (f.task == 0 && R == 64)
      hipLaunchKernelGGL((k_oob_lds<0, 64>), grid, dim3(kBlock), lds, s, f.pnodes.p, f.tree_start.p, T, X, N, Fn,
                         bits.p, words, lab, yp, f.tree_offset, f.seed, perm, tpb, counts.p, sse.p);
    else if (f.task == 0)
      hipLaunchKernelGGL((k_oob_lds<0, 32>), grid, dim3(kBlock), lds, s, f.pnodes.p, f.tree_start.p, T, X, N, Fn,
                         bits.p, words, lab, yp, f.tree_offset, f.seed, perm, tpb, counts.p, sse.p);
    else if (R == 64)
      hipLaunchKernelGGL((k_oob_lds<1, 64>), grid, dim3(kBlock), lds, s, f.pnodes.p, f.tree_start.p, T, X, N, Fn,
                         bits.p, words, lab, yp, f.tree_offset, f.seed, perm, tpb, counts.p, sse.p);
    else
      hipLaunchKernelGGL((k_oob_lds<1, 32>), grid, dim3(kBlock), lds, s, f.pnodes.p, f.tree_start.p, T, X, N, Fn,
                         bits.p, words, lab, yp, f.tree_offset, f.seed, perm, tpb, counts.p, sse.p);
  } else if (f.task == 0) {
    hipLaunchKernelGGL(k_oob<0>, dim3((unsigned)gx, (unsigned)T), dim3(kBlock), 0, s, f.pnodes.p, f.tree_start.p, X,
                       N, bits.p, words, label.p, (const float*)nullptr, f.tree_offset, f.seed, perm, counts.p, sse.p);
  } else {
    hipLaunchKernelGGL(k_oob<1>, dim3((unsigned)gx, (unsigned)T), dim3(kBlock), 0, s, f.pnodes.p, f.tree_start.p, X,
                       N, bits.p, words, (const uint32_t*)nullptr, yv.p, f.tree_offset, f.seed, perm, counts.p, sse.p);
  }
  DTE_HIP(hipGetLastError());
  unsigned long long c[2];
  double e = 0.0;
  DTE_HIP(hipMemcpyAsync(c, counts.p, sizeof(c), hipMemcpyDeviceToHost, s));
  DTE_HIP(hipMemcpyAsync(&e, sse.p, sizeof(e), hipMemcpyDeviceToHost, s));
  DTE_HIP(hipStreamSynchronize(s));
  out[0] = (double)c[0];
  out[1] = f.task == 0 ? (double)c[1] : e;
}

void votes_to_float(int32_t* votes, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_votes_to_float, dim3(elementwise_grid(n)), dim3(kBlock), 0, s, votes, n);
  DTE_HIP(hipGetLastError());
}

void predict_forest(const Forest& f, const float* X, int64_t n_rows, int64_t ld, int row_major,
                    int32_t* votes, float* sums, hipStream_t s) {
  if (n_rows <= 0) return;
  if (f.task == 0) {
    // votes cleared by a kernel (no memset node if the caller captures)
    const int64_t nv = n_rows * f.n_classes;
    hipLaunchKernelGGL(k_zero_u32, dim3(elementwise_grid(nv)), dim3(kBlock), 0, s, reinterpret_cast<uint32_t*>(votes), nv);
    // Rows staged in LDS when a block's rows and votes fit 48 KB: 64 rows
    // while F + C <= 191, 32 rows while F + C <= 383.
    const int F = f.n_features, C = f.n_classes;
    const int R = staged_rows(F, C);
    if (R) {
      const int64_t row_blocks = (n_rows + R - 1) / R;
      const int chunks = tree_chunks(row_blocks, f.n_trees);
      const int tpb = (f.n_trees + chunks - 1) / chunks;
      const dim3 grid((unsigned)row_blocks, (unsigned)((f.n_trees + tpb - 1) / tpb));
      const size_t lds = sizeof(uint32_t) * (size_t)R * (C + F + 1);
      Timer tm("predict", 0.0, s);
      if (R == 64)
        hipLaunchKernelGGL(k_predict_cls_lds<64>, grid, dim3(kBlock), lds, s, f.pnodes.p, f.tree_start.p, f.n_trees,
                           X, n_rows, ld, row_major, C, F, tpb, votes);
      else
        hipLaunchKernelGGL(k_predict_cls_lds<32>, grid, dim3(kBlock), lds, s, f.pnodes.p, f.tree_start.p, f.n_trees,
                           X, n_rows, ld, row_major, C, F, tpb, votes);
      DTE_HIP(hipGetLastError());
      return;
    }
    const int64_t row_blocks = (n_rows + kRowsPerBlock - 1) / kRowsPerBlock;
    const int chunks = tree_chunks(row_blocks, f.n_trees);
    const int tpb = (f.n_trees + chunks - 1) / chunks;
    const dim3 grid((unsigned)row_blocks, (unsigned)((f.n_trees + tpb - 1) / tpb));
    Timer tm("predict", 0.0, s);
    hipLaunchKernelGGL(k_predict_cls, grid, dim3(kBlock), sizeof(uint32_t) * kRowsPerBlock * f.n_classes, s,
                       f.pnodes.p, f.tree_start.p, f.n_trees, X, n_rows, ld, row_major, f.n_classes, tpb, votes);
  } else {
    // Rows staged in LDS (with the chunk's leaf values) when they fit 48 KB.
    const int F = f.n_features;
    const int R = staged_rows(F, kRegChunk);
    if (R) {
      const int64_t row_blocks = (n_rows + R - 1) / R;
      const size_t lds = sizeof(float) * (size_t)R * (kRegChunk + F + 1);
      Timer tm("predict", 0.0, s);
      if (R == 64)
        hipLaunchKernelGGL(k_predict_reg_lds<64>, dim3((unsigned)row_blocks), dim3(kBlock), lds, s, f.pnodes.p,
                           f.tree_start.p, f.n_trees, X, n_rows, ld, row_major, F, sums);
      else
        hipLaunchKernelGGL(k_predict_reg_lds<32>, dim3((unsigned)row_blocks), dim3(kBlock), lds, s, f.pnodes.p,
                           f.tree_start.p, f.n_trees, X, n_rows, ld, row_major, F, sums);
    } else {
      const int grid = (int)std::min<int64_t>((n_rows + kBlock - 1) / kBlock, 65535);
      Timer tm("predict", 0.0, s);
      hipLaunchKernelGGL(k_predict_reg, dim3(grid), dim3(kBlock), 0, s, f.pnodes.p, f.tree_start.p, f.n_trees, X,
                         n_rows, ld, row_major, sums);
    }
  }
  DTE_HIP(hipGetLastError());
}

}  // namespace dte
