// Dense ranks of every feature column on the GPU (random-forest
// preprocessing, once per fit):
//   K[f][row] = index of X[f][row] among the distinct values of feature f
//   SV[SVoff[f] + k] = k-th smallest distinct value of feature f
// All F columns are sorted at once by an LSD radix sort on order-preserving
// float keys: 4 passes of 8 bits; each pass = per-tile digit histogram, one
// device-wide scan over [feature][digit][tile] (whose prefix at
// (f, 0, 0) is exactly f*N, so offsets are absolute), and a stable
// tile-local scatter through two 4-bit LDS passes.
#include <hip/hip_runtime.h>

#include "dte_block.h"
#include "dte_internal.h"

namespace dte {

void scan_u32_ext(const uint32_t* in, int64_t n, uint32_t* out, hipStream_t s, std::vector<DevBuf<uint32_t>>& tmp);

namespace {

constexpr int kRIpt = 8;
constexpr int kRTile = kBlock * kRIpt;  // 2048

__device__ __forceinline__ uint32_t order_key(float v) {
  const uint32_t u = __float_as_uint(v);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float key_value(uint32_t k) {
  const uint32_t u = (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k;
  return __uint_as_float(u);
}

__global__ void k_rank_init(const float* X, int64_t total, uint32_t* key, uint32_t* val, int64_t N) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    key[i] = order_key(X[i]);
    val[i] = (uint32_t)(i % N);
  }
}

// hist[(f * 256 + digit) * tiles + tile]
__global__ void __launch_bounds__(kBlock) k_rank_hist(const uint32_t* key, int64_t N, int tiles, int shift,
                                                      uint32_t* hist) {
  __shared__ uint32_t s_h[256];
  const int f = blockIdx.y, tile = blockIdx.x;
  s_h[threadIdx.x] = 0;
  __syncthreads();
  const int64_t base = (int64_t)f * N + (int64_t)tile * kRTile;
  const int64_t lim = min((int64_t)kRTile, N - (int64_t)tile * kRTile);
  for (int i = threadIdx.x; i < lim; i += kBlock) atomicAdd(&s_h[(key[base + i] >> shift) & 255u], 1u);
  __syncthreads();
  hist[((int64_t)f * 256 + threadIdx.x) * tiles + tile] = s_h[threadIdx.x];
}

__global__ void __launch_bounds__(kBlock) k_rank_scatter(const uint32_t* kin, const uint32_t* vin, int64_t N,
                                                         int tiles, int shift, const uint32_t* off, uint32_t* kout,
                                                         uint32_t* vout) {
  __shared__ uint32_t s_k[2][kRTile];
  __shared__ uint16_t s_v[2][kRTile];
  __shared__ uint16_t s_cnt[16 * kBlock];
  __shared__ uint32_t s_w[kWaves];
  __shared__ uint32_t s_dstart[256];
  __shared__ uint32_t s_goff[256];
  const int f = blockIdx.y, tile = blockIdx.x, t = threadIdx.x;
  const int64_t base = (int64_t)f * N + (int64_t)tile * kRTile;
  const int n = (int)min((int64_t)kRTile, N - (int64_t)tile * kRTile);
  for (int i = t; i < n; i += kBlock) {
    s_k[0][i] = kin[base + i];
    s_v[0][i] = (uint16_t)i;
  }
  s_goff[t] = off[((int64_t)f * 256 + t) * tiles + tile];
  __syncthreads();
  // stable local sort on the 8-bit digit (two 4-bit passes on the digit)
  lds_sort_pass<kRIpt, 256>(s_k[0], s_v[0], s_k[1], s_v[1], n, shift, s_cnt, s_w);
  lds_sort_pass<kRIpt, 256>(s_k[1], s_v[1], s_k[0], s_v[0], n, shift + 4, s_cnt, s_w);
  // local start of each digit
  if (t < 256) s_dstart[t] = 0xFFFFFFFFu;
  __syncthreads();
  for (int i = t; i < n; i += kBlock) {
    const uint32_t d = (s_k[0][i] >> shift) & 255u;
    if (i == 0 || ((s_k[0][i - 1] >> shift) & 255u) != d) s_dstart[d] = (uint32_t)i;
  }
  __syncthreads();
  for (int i = t; i < n; i += kBlock) {
    const uint32_t d = (s_k[0][i] >> shift) & 255u;
    const int64_t pos = (int64_t)s_goff[d] + (i - (int)s_dstart[d]);
    kout[pos] = s_k[0][i];
    vout[pos] = vin[base + s_v[0][i]];
  }
}

// flag[i] = 1 where a new distinct value starts (and at every column start)
__global__ void k_rank_flags(const uint32_t* key, int64_t total, int64_t N, uint32_t* flag) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x)
    flag[i] = ((i % N) == 0 || key[i] != key[i - 1]) ? 1u : 0u;
}

// dense[i] = exclusive scan of flags; K and SV from it.
__global__ void k_rank_emit(const uint32_t* key, const uint32_t* row, const uint32_t* dense, int64_t total, int64_t N,
                            uint32_t* K, float* SV) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t f = i / N;
    const bool start = (i % N) == 0 || key[i] != key[i - 1];
    // dense[] counts run starts before i: a run start has index dense[i],
    // any later element of the same run has dense[i] - 1.
    const uint32_t rank_global = start ? dense[i] : dense[i] - 1u;
    const uint32_t col_base = dense[f * N];
    K[f * N + row[i]] = rank_global - col_base;
    if (start) SV[rank_global] = key_value(key[i]);
  }
}

__global__ void k_rank_meta(const uint32_t* dense, int64_t N, int F, int64_t* SVoff, uint32_t* S) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= F) return;
  const uint32_t b = dense[(int64_t)f * N];
  const uint32_t e = dense[(int64_t)(f + 1) * N];  // dense has F*N + 1 entries
  SVoff[f] = b;
  S[f] = e - b;
}

int grid_of(int64_t n) {
  int64_t g = (n + kBlock - 1) / kBlock;
  return (int)std::max<int64_t>(1, std::min<int64_t>(g, 65535 * 4));
}

}  // namespace

// Computes K (F*N), SV (sum S_f), SVoff (F), S (F) on the device.
void dense_ranks_device(const float* X, int64_t N, int F, DevBuf<uint32_t>& K, DevBuf<float>& SV,
                        DevBuf<int64_t>& SVoff, DevBuf<uint32_t>& S, hipStream_t s,
                        std::vector<DevBuf<uint32_t>>& scan_tmp) {
  const int64_t total = N * (int64_t)F;
  const int tiles = (int)((N + kRTile - 1) / kRTile);
  DevBuf<uint32_t> ka(total), va(total), kb(total), vb(total);
  const int64_t nh = (int64_t)F * 256 * tiles;
  DevBuf<uint32_t> hist(nh), off(nh + 1);
  hipLaunchKernelGGL(k_rank_init, dim3(grid_of(total)), dim3(kBlock), 0, s, X, total, ka.p, va.p, N);
  uint32_t *kin = ka.p, *vin = va.p, *kout = kb.p, *vout = vb.p;
  for (int shift = 0; shift < 32; shift += 8) {
    hipLaunchKernelGGL(k_rank_hist, dim3(tiles, F), dim3(kBlock), 0, s, kin, N, tiles, shift, hist.p);
    scan_u32_ext(hist.p, nh, off.p, s, scan_tmp);
    hipLaunchKernelGGL(k_rank_scatter, dim3(tiles, F), dim3(kBlock), 0, s, kin, vin, N, tiles, shift, off.p, kout,
                       vout);
    std::swap(kin, kout);
    std::swap(vin, vout);
  }
  // kin/vin: sorted keys and rows per feature
  hist.release();
  off.release();
  DevBuf<uint32_t> flag(total), dense(total + 1);
  hipLaunchKernelGGL(k_rank_flags, dim3(grid_of(total)), dim3(kBlock), 0, s, kin, total, N, flag.p);
  scan_u32_ext(flag.p, total, dense.p, s, scan_tmp);
  uint32_t n_distinct = 0;
  DTE_HIP(hipMemcpyAsync(&n_distinct, dense.p + total, 4, hipMemcpyDeviceToHost, s));
  DTE_HIP(hipStreamSynchronize(s));
  K.alloc(total);
  SV.alloc(n_distinct);
  SVoff.alloc(F);
  S.alloc(F);
  hipLaunchKernelGGL(k_rank_emit, dim3(grid_of(total)), dim3(kBlock), 0, s, kin, vin, dense.p, total, N, K.p, SV.p);
  hipLaunchKernelGGL(k_rank_meta, dim3((F + kBlock - 1) / kBlock), dim3(kBlock), 0, s, dense.p, N, F, SVoff.p, S.p);
  DTE_HIP(hipGetLastError());
}

}  // namespace dte
