// Team-level primitives for gfx950 (wave64).  A "team" is either one wave
// (TEAM = 64: synchronised with a wavefront-scope fence, several teams per
// workgroup work on independent items) or a whole 256-thread workgroup
// (TEAM = 256: __syncthreads).  Scans go through DPP-backed wave shuffles
// and one LDS round for the cross-wave part.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace dte {

constexpr int kWave = 64;
constexpr int kWaves = 256 / kWave;

template <int TEAM>
__device__ __forceinline__ void team_sync() {
  if constexpr (TEAM == kWave) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

template <int TEAM>
__device__ __forceinline__ int team_tid() {
  return TEAM == kWave ? (int)(threadIdx.x & (kWave - 1)) : (int)threadIdx.x;
}

// Wave64 scans and reductions through DPP (GFX9 row shifts and row
// broadcasts): plain VALU data movement, no LDS crossbar.  They assume all
// 64 lanes are active, which every caller guarantees.
// DPP controls: row_shr:n = 0x110 + n, row_bcast:15 = 0x142, row_bcast:31 =
// 0x143.  A lane whose source is out of its row, or whose row is masked
// off, gets `ident`.
template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t v, uint32_t ident) {
  return (uint32_t)__builtin_amdgcn_update_dpp((int)ident, (int)v, CTRL, ROW_MASK, 0xf, false);
}

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x) {
  x += dpp_u32<0x111>(x, 0u);
  x += dpp_u32<0x112>(x, 0u);
  x += dpp_u32<0x114>(x, 0u);
  x += dpp_u32<0x118>(x, 0u);
  x += dpp_u32<0x142, 0xa>(x, 0u);
  x += dpp_u32<0x143, 0xc>(x, 0u);
  return x;
}

template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ int64_t dpp_i64(int64_t v) {
  const uint64_t u = (uint64_t)v;
  const uint32_t lo = dpp_u32<CTRL, ROW_MASK>((uint32_t)u, 0u);
  const uint32_t hi = dpp_u32<CTRL, ROW_MASK>((uint32_t)(u >> 32), 0u);
  return (int64_t)(((uint64_t)hi << 32) | lo);
}

__device__ __forceinline__ int64_t wave_incl_scan_i64(int64_t x) {
  x += dpp_i64<0x111>(x);
  x += dpp_i64<0x112>(x);
  x += dpp_i64<0x114>(x);
  x += dpp_i64<0x118>(x);
  x += dpp_i64<0x142, 0xa>(x);
  x += dpp_i64<0x143, 0xc>(x);
  return x;
}

__device__ __forceinline__ uint32_t wave_last(uint32_t v) { return (uint32_t)__builtin_amdgcn_readlane((int)v, 63); }
__device__ __forceinline__ int64_t wave_last_i64(int64_t v) {
  const uint64_t u = (uint64_t)v;
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, 63);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), 63);
  return (int64_t)(((uint64_t)hi << 32) | lo);
}

// Exclusive scan over the team; s_w holds kWaves words (unused for a wave
// team).
template <int TEAM>
__device__ __forceinline__ uint32_t team_excl_scan(uint32_t v, uint32_t* s_w, uint32_t* total) {
  const uint32_t inc = wave_incl_scan(v);
  if constexpr (TEAM == kWave) {
    if (total) *total = wave_last(inc);
    return inc - v;
  } else {
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    if (lane == kWave - 1) s_w[w] = inc;
    __syncthreads();
    uint32_t pre = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < kWaves; ++i) {
      const uint32_t s = s_w[i];
      pre += (i < w) ? s : 0u;
      tot += s;
    }
    __syncthreads();
    if (total) *total = tot;
    return pre + inc - v;
  }
}

template <int TEAM>
__device__ __forceinline__ int64_t team_excl_scan_i64(int64_t v, int64_t* s_w, int64_t* total) {
  const int64_t inc = wave_incl_scan_i64(v);
  if constexpr (TEAM == kWave) {
    if (total) *total = wave_last_i64(inc);
    return inc - v;
  } else {
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    if (lane == kWave - 1) s_w[w] = inc;
    __syncthreads();
    int64_t pre = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < kWaves; ++i) {
      const int64_t s = s_w[i];
      pre += (i < w) ? s : 0;
      tot += s;
    }
    __syncthreads();
    if (total) *total = tot;
    return pre + inc - v;
  }
}

// Exclusive scan of N lanes of counters at once; s_w holds N * kWaves words.
template <int TEAM, int N>
__device__ __forceinline__ void team_excl_scan_vec(uint32_t (&v)[N], uint32_t* s_w, uint32_t (&tot)[N]) {
  uint32_t inc[N];
#pragma unroll
  for (int c = 0; c < N; ++c) inc[c] = wave_incl_scan(v[c]);
  if constexpr (TEAM == kWave) {
#pragma unroll
    for (int c = 0; c < N; ++c) {
      tot[c] = wave_last(inc[c]);
      v[c] = inc[c] - v[c];
    }
  } else {
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
#pragma unroll
    for (int c = 0; c < N; ++c)
      if (lane == kWave - 1) s_w[c * kWaves + w] = inc[c];
    __syncthreads();
#pragma unroll
    for (int c = 0; c < N; ++c) {
      uint32_t pre = 0, t = 0;
#pragma unroll
      for (int i = 0; i < kWaves; ++i) {
        const uint32_t s = s_w[c * kWaves + i];
        pre += (i < w) ? s : 0u;
        t += s;
      }
      tot[c] = t;
      v[c] = pre + inc[c] - v[c];
    }
    __syncthreads();
  }
}

// Exclusive scan of N int64 lanes at once, with totals; s_w holds
// N * kWaves words (unused for a wave team).
template <int TEAM, int N>
__device__ __forceinline__ void team_excl_scan_i64_vec(int64_t (&v)[N], int64_t* s_w, int64_t (&tot)[N]) {
  int64_t inc[N];
#pragma unroll
  for (int c = 0; c < N; ++c) inc[c] = wave_incl_scan_i64(v[c]);
  if constexpr (TEAM == kWave) {
#pragma unroll
    for (int c = 0; c < N; ++c) {
      tot[c] = wave_last_i64(inc[c]);
      v[c] = inc[c] - v[c];
    }
  } else {
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
#pragma unroll
    for (int c = 0; c < N; ++c)
      if (lane == kWave - 1) s_w[c * kWaves + w] = inc[c];
    __syncthreads();
#pragma unroll
    for (int c = 0; c < N; ++c) {
      int64_t pre = 0, t = 0;
#pragma unroll
      for (int i = 0; i < kWaves; ++i) {
        const int64_t x = s_w[c * kWaves + i];
        pre += (i < w) ? x : 0;
        t += x;
      }
      tot[c] = t;
      v[c] = pre + inc[c] - v[c];
    }
    __syncthreads();
  }
}

__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
  v = max(v, dpp_u32<0x111>(v, 0u));
  v = max(v, dpp_u32<0x112>(v, 0u));
  v = max(v, dpp_u32<0x114>(v, 0u));
  v = max(v, dpp_u32<0x118>(v, 0u));
  v = max(v, dpp_u32<0x142, 0xa>(v, 0u));
  v = max(v, dpp_u32<0x143, 0xc>(v, 0u));
  return wave_last(v);
}
__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
  v = min(v, dpp_u32<0x111>(v, ~0u));
  v = min(v, dpp_u32<0x112>(v, ~0u));
  v = min(v, dpp_u32<0x114>(v, ~0u));
  v = min(v, dpp_u32<0x118>(v, ~0u));
  v = min(v, dpp_u32<0x142, 0xa>(v, ~0u));
  v = min(v, dpp_u32<0x143, 0xc>(v, ~0u));
  return wave_last(v);
}

template <int TEAM>
__device__ __forceinline__ uint32_t team_reduce_max(uint32_t v, uint32_t* s_w) {
  v = wave_max_u32(v);
  if constexpr (TEAM == kWave) {
    return v;
  } else {
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    if (lane == 0) s_w[w] = v;
    __syncthreads();
    uint32_t r = s_w[0];
#pragma unroll
    for (int i = 1; i < kWaves; ++i) r = max(r, s_w[i]);
    __syncthreads();
    return r;
  }
}

template <int TEAM>
__device__ __forceinline__ uint32_t team_reduce_min(uint32_t v, uint32_t* s_w) {
  v = wave_min_u32(v);
  if constexpr (TEAM == kWave) {
    return v;
  } else {
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    if (lane == 0) s_w[w] = v;
    __syncthreads();
    uint32_t r = s_w[0];
#pragma unroll
    for (int i = 1; i < kWaves; ++i) r = min(r, s_w[i]);
    __syncthreads();
    return r;
  }
}

// Best split candidate: larger gain wins, equal gain -> smaller key (lower
// threshold).  key = UINT32_MAX means "none".
struct BestCand {
  double g;
  uint32_t key;
  uint32_t prev;
};

__device__ __forceinline__ bool cand_better(double ga, uint32_t ka, double gb, uint32_t kb) {
  return ga > gb || (ga == gb && ka < kb);
}

// One step of the best-candidate scan: combine with the DPP source lane
// (lanes without a source see key = none, which never wins).
template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ void best_step(BestCand& b) {
  const uint64_t gb = __builtin_bit_cast(uint64_t, b.g);
  const uint32_t glo = dpp_u32<CTRL, ROW_MASK>((uint32_t)gb, 0u);
  const uint32_t ghi = dpp_u32<CTRL, ROW_MASK>((uint32_t)(gb >> 32), 0u);
  const uint32_t k = dpp_u32<CTRL, ROW_MASK>(b.key, 0xFFFFFFFFu);
  const uint32_t p = dpp_u32<CTRL, ROW_MASK>(b.prev, 0xFFFFFFFFu);
  const double g = __builtin_bit_cast(double, ((uint64_t)ghi << 32) | glo);
  if (k != 0xFFFFFFFFu && cand_better(g, k, b.g, b.key)) {
    b.g = g;
    b.key = k;
    b.prev = p;
  }
}

// Best candidate of the wave (higher gain, then lower key); every lane
// gets it.  The "better" relation is a total order, so a scan-style
// combination in any association gives the maximum.
__device__ __forceinline__ BestCand wave_best(BestCand b) {
  best_step<0x111>(b);
  best_step<0x112>(b);
  best_step<0x114>(b);
  best_step<0x118>(b);
  best_step<0x142, 0xa>(b);
  best_step<0x143, 0xc>(b);
  const uint64_t gb = __builtin_bit_cast(uint64_t, b.g);
  const uint32_t glo = wave_last((uint32_t)gb), ghi = wave_last((uint32_t)(gb >> 32));
  return BestCand{__builtin_bit_cast(double, ((uint64_t)ghi << 32) | glo), wave_last(b.key), wave_last(b.prev)};
}

// Team reduction of the best candidate; every lane gets the result.
template <int TEAM>
__device__ __forceinline__ BestCand team_best(BestCand b, double* s_g, uint32_t* s_k, uint32_t* s_p) {
  b = wave_best(b);
  if constexpr (TEAM == kWave) {
    return b;
  } else {
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    if (lane == 0) {
      s_g[w] = b.g;
      s_k[w] = b.key;
      s_p[w] = b.prev;
    }
    __syncthreads();
    BestCand r{s_g[0], s_k[0], s_p[0]};
#pragma unroll
    for (int i = 1; i < kWaves; ++i)
      if (cand_better(s_g[i], s_k[i], r.g, r.key)) r = BestCand{s_g[i], s_k[i], s_p[i]};
    __syncthreads();
    return r;
  }
}

// Positions per lane when n elements are spread over the whole team.
template <int TEAM>
__device__ __forceinline__ int lane_span(int n) {
  return (n + TEAM - 1) / TEAM;
}

// One stable LSD pass over n <= TEAM*IPT (key, val) pairs held in LDS, on the
// 4-bit digit at `shift`.  Lane t owns positions [t*ipt, t*ipt+ipt) with
// ipt = ceil(n / TEAM), so small inputs keep every lane busy.
// s_cnt: 16 * TEAM uint16 counters.  Reads (kin, vin), writes (kout, vout).
template <int IPT, int TEAM>
__device__ __forceinline__ void lds_sort_pass(const uint32_t* kin, const uint16_t* vin, uint32_t* kout,
                                              uint16_t* vout, int n, int shift, uint16_t* s_cnt, uint32_t* s_w) {
  const int t = team_tid<TEAM>();
  const int ipt = lane_span<TEAM>(n);
  const int lo = t * ipt;
  uint32_t k[IPT];
  uint16_t v[IPT];
  uint32_t d[IPT];
#pragma unroll
  for (int i = 0; i < 16; ++i) s_cnt[i * TEAM + t] = 0;
#pragma unroll
  for (int i = 0; i < IPT; ++i) {
    const int p = lo + i;
    if (i < ipt && p < n) {
      k[i] = kin[p];
      v[i] = vin[p];
      d[i] = (k[i] >> shift) & 15u;
      s_cnt[d[i] * TEAM + t] += 1;
    }
  }
  team_sync<TEAM>();
  // Exclusive scan of the 16 x TEAM counter matrix in digit-major order; lane
  // t covers flat entries [16t, 16t+16).
  uint32_t loc[16];
  uint32_t sum = 0;
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    loc[i] = s_cnt[t * 16 + i];
    sum += loc[i];
  }
  uint32_t base = team_excl_scan<TEAM>(sum, s_w, nullptr);
  team_sync<TEAM>();
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    s_cnt[t * 16 + i] = (uint16_t)base;
    base += loc[i];
  }
  team_sync<TEAM>();
#pragma unroll
  for (int i = 0; i < IPT; ++i) {
    const int p = lo + i;
    if (i < ipt && p < n) {
      const uint32_t pos = s_cnt[d[i] * TEAM + t];
      s_cnt[d[i] * TEAM + t] = (uint16_t)(pos + 1);
      kout[pos] = k[i];
      vout[pos] = v[i];
    }
  }
  team_sync<TEAM>();
}

// Counting sort of n <= TEAM*IPT (key, val) pairs whose keys lie in
// [0, span), with two 16-bit counters per word of `cnt` (span / 2 words;
// n < 65536, so counts and offsets fit a half word).  Equal keys come out in
// any order: callers only use the order of distinct keys.
template <int IPT, int TEAM>
__device__ __forceinline__ void lds_count_sort(const uint32_t* kin, const uint16_t* vin, uint32_t* kout,
                                               uint16_t* vout, int n, int span, uint32_t* cnt, uint32_t* s_w) {
  const int t = team_tid<TEAM>();
  const int words = (span + 1) >> 1;
  for (int i = t; i < words; i += TEAM) cnt[i] = 0;
  team_sync<TEAM>();
  uint32_t r[IPT];
#pragma unroll
  for (int q = 0; q < IPT; ++q) {
    const int e = t + q * TEAM;
    if (e < n) {
      const uint32_t k = kin[e];
      const uint32_t sh = (k & 1u) << 4;
      r[q] = (atomicAdd(&cnt[k >> 1], 1u << sh) >> sh) & 0xFFFFu;
    }
  }
  team_sync<TEAM>();
  const int per = (words + TEAM - 1) / TEAM;
  const int lo = t * per, hi = min(lo + per, words);
  uint32_t loc = 0;
  for (int i = lo; i < hi; ++i) {
    const uint32_t w = cnt[i];
    loc += (w & 0xFFFFu) + (w >> 16);
  }
  uint32_t base = team_excl_scan<TEAM>(loc, s_w, nullptr);
  for (int i = lo; i < hi; ++i) {
    const uint32_t w = cnt[i];
    const uint32_t c0 = w & 0xFFFFu;
    cnt[i] = base | ((base + c0) << 16);
    base += c0 + (w >> 16);
  }
  team_sync<TEAM>();
#pragma unroll
  for (int q = 0; q < IPT; ++q) {
    const int e = t + q * TEAM;
    if (e < n) {
      const uint32_t k = kin[e];
      const uint32_t sh = (k & 1u) << 4;
      const uint32_t pos = ((cnt[k >> 1] >> sh) & 0xFFFFu) + r[q];
      kout[pos] = k;
      vout[pos] = vin[e];
    }
  }
  team_sync<TEAM>();
}

// Sorts n <= TEAM*IPT (key, val) pairs whose keys lie below 2^bits, with
// 2^bits above the counter span: a counting sort on the top digit
// (key >> shift, shift = bits - log2(span)) from (k0, v0) into (k1, v1), then
// every run of equal top digits ordered by each entry's rank inside its run
// (smaller key, or equal key and earlier position), written back to (k0, v0).
// Returns false, with (k0, v0) untouched, when a run holds more than max_run
// entries: the caller sorts by radix passes instead.
template <int IPT, int TEAM>
__device__ __noinline__ bool lds_digit_sort(uint32_t* k0, uint16_t* v0, uint32_t* k1, uint16_t* v1, int n,
                                               int bits, int span, int max_run, uint32_t* cnt, uint32_t* s_w) {
  const int t = team_tid<TEAM>();
  int lg = 0;
  while ((1 << (lg + 1)) <= span) ++lg;
  const int shift = bits - lg;
  const int words = span >> 1;
  for (int i = t; i < words; i += TEAM) cnt[i] = 0;
  team_sync<TEAM>();
  uint32_t r[IPT];
#pragma unroll
  for (int q = 0; q < IPT; ++q) {
    const int e = t + q * TEAM;
    if (e < n) {
      const uint32_t dg = k0[e] >> shift;
      const uint32_t sh = (dg & 1u) << 4;
      r[q] = (atomicAdd(&cnt[dg >> 1], 1u << sh) >> sh) & 0xFFFFu;
    }
  }
  team_sync<TEAM>();
  const int per = (words + TEAM - 1) / TEAM;
  const int lo = t * per, hi = min(lo + per, words);
  uint32_t loc = 0, run_max = 0;
  for (int i = lo; i < hi; ++i) {
    const uint32_t w = cnt[i];
    loc += (w & 0xFFFFu) + (w >> 16);
    run_max = max(run_max, max(w & 0xFFFFu, w >> 16));
  }
  run_max = team_reduce_max<TEAM>(run_max, s_w);
  if (run_max > (uint32_t)max_run) return false;  // team-uniform
  uint32_t base = team_excl_scan<TEAM>(loc, s_w, nullptr);
  for (int i = lo; i < hi; ++i) {
    const uint32_t w = cnt[i];
    const uint32_t c0 = w & 0xFFFFu;
    cnt[i] = base | ((base + c0) << 16);
    base += c0 + (w >> 16);
  }
  team_sync<TEAM>();
  auto start = [&](uint32_t dg) -> int {
    if ((int)dg >= span) return n;
    return (int)((cnt[dg >> 1] >> ((dg & 1u) << 4)) & 0xFFFFu);
  };
#pragma unroll
  for (int q = 0; q < IPT; ++q) {
    const int e = t + q * TEAM;
    if (e < n) {
      const uint32_t k = k0[e];
      const uint32_t dg = k >> shift;
      const int pos = start(dg) + (int)r[q];
      k1[pos] = k;
      v1[pos] = v0[e];
    }
  }
  team_sync<TEAM>();
#pragma unroll
  for (int q = 0; q < IPT; ++q) {
    const int p = t + q * TEAM;
    if (p < n) {
      const uint32_t kp = k1[p];
      const uint32_t dg = kp >> shift;
      const int s0 = start(dg), s1 = start(dg + 1u);
      int rank = 0;
      for (int x = s0; x < s1; ++x) {
        const uint32_t kx = k1[x];
        rank += (kx < kp || (kx == kp && x < p)) ? 1 : 0;
      }
      k0[s0 + rank] = kp;
      v0[s0 + rank] = v1[p];
    }
  }
  team_sync<TEAM>();
  return true;
}

// Compatibility names used by the whole-block kernels.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* s_w, uint32_t* total) {
  return team_excl_scan<256>(v, s_w, total);
}
__device__ __forceinline__ int64_t block_excl_scan_i64(int64_t v, int64_t* s_w, int64_t* total) {
  return team_excl_scan_i64<256>(v, s_w, total);
}

}  // namespace dte
