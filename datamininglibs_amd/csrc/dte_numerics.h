// Deterministic numerics shared by every build and predict kernel of the
// tree-ensemble path, usable from host and device code.
//
// Everything here is defined as a fixed sequence of IEEE-754 operations
// (fma only where written; the library is compiled with -ffp-contract=off),
// so host and gfx950 produce identical bits.
//
//  * counter-based RNG: Philox4x32-10 (Salmon et al., SC'11).  Replaces the
//    reference's per-launch-position cuRAND MRG32k3a states
//    (lib_ensembles/source/gpurf.cu:48-56, gpuert.cu:44-52).
//  * x ln x on integer counts: the reference's lnFunc (gpudte.cu:70-74)
//    evaluated with a 128-entry reduction table plus a degree-7 log1p
//    polynomial instead of the device log.
//  * information gain in bits (gpudte.cu:76-117) and variance reduction
//    (gpuert.cu:370-375) from exact integer statistics.
#pragma once
#include <stdint.h>
#include <float.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define DTE_HD __host__ __device__ __forceinline__
#else
#define DTE_HD inline
#endif

#include <cmath>
#include <cstring>

namespace dte {

// Streams of the counter-based draws: counter = (idx/4, node, tree, stream),
// key = (seed lo, seed hi); a draw is word idx%4 of the Philox block.
constexpr uint32_t kStreamBoot = 0x746f6f62u;  // bootstrap rows
constexpr uint32_t kStreamFeat = 0x74616566u;  // feature draws of a node
constexpr uint32_t kStreamThr = 0x00726874u;   // extra-trees threshold samples

constexpr double kLn2 = 0.6931471805599453;
constexpr float kMissing = -FLT_MAX;  // reference "missing value" marker (gpudte.cu:331)

DTE_HD void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                          uint32_t k1, uint32_t out[4]) {
#pragma unroll
  for (int round = 0; round < 10; ++round) {
    if (round > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    const uint32_t n1 = (uint32_t)p1;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    const uint32_t n3 = (uint32_t)p0;
    c0 = n0;
    c1 = n1;
    c2 = n2;
    c3 = n3;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

DTE_HD uint32_t rand_u32(uint64_t seed, uint32_t stream, uint32_t tree, uint32_t node,
                         uint64_t idx) {
  uint32_t o[4];
  philox4x32_10((uint32_t)(idx >> 2), node, tree, stream, (uint32_t)seed, (uint32_t)(seed >> 32), o);
  const uint32_t w = (uint32_t)(idx & 3);
  return w == 0 ? o[0] : (w == 1 ? o[1] : (w == 2 ? o[2] : o[3]));
}

DTE_HD uint64_t dbl_bits(double d) {
#if defined(__HIP_DEVICE_COMPILE__)
  return (uint64_t)__double_as_longlong(d);
#else
  uint64_t b;
  std::memcpy(&b, &d, 8);
  return b;
#endif
}
DTE_HD double bits_dbl(uint64_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __longlong_as_double((long long)b);
#else
  double d;
  std::memcpy(&d, &b, 8);
  return d;
#endif
}
DTE_HD double fma_d(double a, double b, double c) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_fma(a, b, c);
#else
  return std::fma(a, b, c);
#endif
}
DTE_HD double ldexp_d(double a, int e) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_ldexp(a, e);
#else
  return std::ldexp(a, e);
#endif
}

// Reduction table of x ln x: inv[j] = 1/(1+(j+0.5)/128), logc[j] = -ln inv[j].
struct LnTable {
  double inv[128];
  double logc[128];
};

// Table-free ln for the table build: atanh series, fixed op order.
inline double series_ln(double v) {
  double s = (v - 1.0) / (v + 1.0);
  double z = s * s;
  double p = 1.0 / 49.0;
  for (int k = 23; k >= 0; --k) p = p * z + 1.0 / (double)(2 * k + 1);
  return 2.0 * s * p;
}

inline void build_ln_table(LnTable* t) {
  for (int j = 0; j < 128; ++j) {
    t->inv[j] = 1.0 / (1.0 + ((double)j + 0.5) / 128.0);
    t->logc[j] = -series_ln(t->inv[j]);
  }
}

// x ln x for an integer count x (0 and 1 give 0).  `inv`/`logc` point at the
// reduction table (LDS copy on the device).
DTE_HD double xlnx(uint64_t x, const double* inv, const double* logc) {
  if (x <= 1) return 0.0;
  const double d = (double)x;
  const uint64_t b = dbl_bits(d);
  const int e = (int)(b >> 52) - 1023;
  const uint64_t mant = b & 0x000FFFFFFFFFFFFFull;
  const int j = (int)(mant >> 45);
  const double m = bits_dbl(mant | 0x3FF0000000000000ull);
  const double r = fma_d(m, inv[j], -1.0);
  double p = fma_d(r, 1.0 / 7.0, -1.0 / 6.0);
  p = fma_d(r, p, 1.0 / 5.0);
  p = fma_d(r, p, -1.0 / 4.0);
  p = fma_d(r, p, 1.0 / 3.0);
  p = fma_d(r, p, -1.0 / 2.0);
  const double l1p = fma_d(r * r, p, r);
  const double lnx = (double)e * kLn2 + (logc[j] + l1p);
  return d * lnx;
}

// Information gain in bits (gpudte.cu:76-117), canonical form.
// Every x ln x term is turned into an exact int64 fixed-point value
//   T(x) = round_half_even(xlnx(x) * 2^F),  F = ig_shift(n, C),
// so that sums of terms are exact and order-free (an incremental update of
// one class gives the same bits as a full recomputation).  With
//   P = T(n) - sum_c T(n_c)
//   Q = (sum_c T(l_c) - T(nl)) + (sum_c T(n_c - l_c) - T(nr))
// the gain is  ldexp((double)(P + Q), -F) * (1 / (n ln 2)).
DTE_HD int bit_length64(uint64_t v) {
  int b = 0;
  while (v) {
    ++b;
    v >>= 1;
  }
  return b;
}

DTE_HD int ig_shift(uint64_t n, int C, const double* inv, const double* logc) {
  const double b = (double)(2 * C + 2) * (xlnx(n, inv, logc) + 1.0);
  return 62 - bit_length64((uint64_t)b);
}

DTE_HD int64_t d2ll_rn(double v) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __double2ll_rn(v);
#else
  return std::llrint(v);
#endif
}

DTE_HD int64_t xlnx_fixed(uint64_t x, int F, const double* inv, const double* logc) {
  if (x <= 1) return 0;
  return d2ll_rn(ldexp_d(xlnx(x, inv, logc), F));
}

DTE_HD double ig_from_fixed(int64_t pq, int F, double inv_nln2) {
  return ldexp_d((double)pq, -F) * inv_nln2;
}

// Variance reduction nl*nr/(nl+nr) * (mean_l - mean_r)^2 from exact fixed-point
// sums, in target units squared.
DTE_HD double variance_gain(int64_t nl, int64_t sl, int64_t nr, int64_t sr, int shift) {
  const double ml = (double)sl / (double)nl;
  const double mr = (double)sr / (double)nr;
  const double diff = ml - mr;
  const double inv_n = 1.0 / (double)(nl + nr);
  const double g = (((double)nl * (double)nr) * inv_n) * (diff * diff);
  return ldexp_d(g, -2 * shift);
}

// Threshold between consecutive distinct values a < b: (a+b)/2 in fp32
// (gpurf.cu:744); if rounding lands on a, b is used so that {x < thr} is
// exactly the scored left side.
DTE_HD float midpoint(float a, float b) {
  float m = (a + b) * 0.5f;
  if (!(a < m && m <= b)) m = b;
  return m;
}

// Routing rule of a split (gpudte.cu:216, 331-336): missing -> left,
// otherwise x < thr -> left.  0 = left, 1 = right.
DTE_HD int side_of(float x, float thr) {
  if (x == kMissing) return 0;
  return x < thr ? 0 : 1;
}

// Row hash summed (mod 2^64, weighted) over a node's entries: pins node
// membership without storing it.
DTE_HD uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Number of bits of fixed-point shift for regression targets: the largest s
// with max|y| * 2^s * total_weight < 2^62 (capped at 60).
inline int fixed_shift(double max_abs, int64_t total_weight) {
  int cl = 0;
  while (cl < 63 && ((int64_t)1 << cl) < total_weight) ++cl;
  int e = 0;
  if (max_abs > 0) std::frexp(max_abs, &e);
  int s = 62 - cl - e;
  if (s > 60) s = 60;
  return s;
}

}  // namespace dte
