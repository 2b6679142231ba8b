// Level-wise tree-ensemble build for MI355X (gfx950).
//
// Replaces the device side of the reference's build (lib_ensembles/source/
// gpudte_algorithm.cc:14-267 driving gpurf.cu / gpuert.cu / gpudte.cu):
//
//   reference                               here
//   kSetupKernel (cuRAND states)            counter-based Philox, no state
//   kInitTreeBatch (bootstrap, 1 blk/tree)  k_boot_count + k_boot_compact
//   kFindSplit (1 64-lane block per node)   k_search_rf / k_search_ert over
//                                           (node, feature draw) items of all
//                                           trees of a level
//   kPerformSplit (atomic slots)            k_partition: stable compaction
//   host node cache + frame swapping        level arrays resident in HBM
//
// Layout in HBM during a fit:
//   X     fp32 [F][N]    feature-major samples (canonical: -0 -> +0)
//   K     u32  [F][N]    dense rank of X[f][row] among the distinct values of
//                        feature f (random forest only)
//   SV    fp32 [sum S_f] distinct sorted values of each feature
//   rows  u32  [E]       entries: in-bag rows of every tree, ascending row id
//                        inside every node (E = sum over trees of distinct
//                        bootstrap rows), ping-pong A/B between levels
//   aux   u32  [E]       label << 16 | bootstrap multiplicity
//   ey    i64  [E]       fixed-point target of the entry (regression)
//   nodes: level-major SoA (tree, id, depth, count, m, stats, hash, split)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <thread>
#include <type_traits>
#include <cstdlib>
#include <chrono>
#include <memory>

#include "dte_block.h"
#include "dte_internal.h"

namespace dte {

__device__ LnTable g_ln_table;
const LnTable& host_ln_table();

namespace {

constexpr int kIpt = 8;                 // elements per thread in the LDS sort
constexpr int kTeamIpt = 8;             // entries per lane of the wave-per-item tier
constexpr int kMaxDigitRun = 16;        // longest run of equal top digits ordered in place
static_assert(kWave * kTeamIpt >= kMaxCutTeam, "wave tier capacity");
constexpr int kCap = kBlock * kIpt;     // LDS chunk capacity (2048)
constexpr int kHistWords = 8192;        // LDS histogram words (32 KB)
constexpr int kMaxBuckets = 2048;       // bucket table of the big-item path
constexpr uint32_t kNone = 0xFFFFFFFFu;

__device__ __forceinline__ void load_ln_table(double* s_inv, double* s_logc) {
  const int t = threadIdx.x;
  if (t < 128) s_inv[t] = g_ln_table.inv[t];
  else s_logc[t - 128] = g_ln_table.logc[t - 128];
}

// ------------------------------------------------------------- args -----
constexpr int kMaxTiers = 7;
struct DataArgs {
  const float* X;
  const uint32_t* K;
  const float* SV;
  const int64_t* SVoff;
  const uint32_t* S;
  int64_t N;
  int F;
  int C;
  int shift;
  uint64_t seed;
  int min_node;
  int max_depth;
  int igF;           // fixed-point shift of the x ln x terms (per fit)
  const int64_t* T;  // T[x] = xlnx_fixed(x, igF) for x <= Tmax (classification)
  int64_t Tmax;
  // Joint build of several extra-trees forests in one set of level steps:
  // the data of batch tree t is forests[forest_of_tree[t]], whose D0 / R
  // are its mandatory and retry draws; C is the joint class width in every
  // forest's entry (classes past a forest's own count hold zero counts,
  // which add nothing to the entropy sums).  NULL for a fit of one forest.
  const struct DataArgs* forests;
  const int32_t* forest_of_tree;
  int D0, R;
};

__device__ __forceinline__ DataArgs forest_data(const DataArgs& d, int tree) {
  return d.forests ? d.forests[d.forest_of_tree[tree]] : d;
}

// Counts and offsets of the level being built, in device memory, for the
// level steps that run without a host round trip (captured in a hipGraph and
// replayed level after level): the kernels read them at entry instead of
// taking them as launch arguments.
struct DevLevel {
  int64_t off;                    // level-major offset of the level
  int64_t n;                      // nodes of the level
  int64_t depth;
  unsigned int n_open;            // open nodes (total of the open-flag scan)
  unsigned int n_retry;           // open nodes that draw retries
  unsigned int n_next;            // nodes of the next level
  unsigned int tier[kMaxTiers];   // split-search tier sizes, first phase
  unsigned int rtier[kMaxTiers];  // the same, retry phase
  unsigned int ptier[kMaxTiers];  // partition tier sizes
};

struct LevelArgs {
  // level node slice (level-local index i)
  const int32_t* tree;   // batch tree index
  const uint32_t* id;    // tree-local level-order id
  const int32_t* count;  // sum of weights
  const int32_t* m;      // entries
  const uint32_t* cls;   // class counts [i * C]
  const int64_t* sum;    // regression fixed-point sum
  const int64_t* start;  // entry start (current buffer)
  const uint32_t* open_idx;  // open position -> level index
  const int32_t* tree_global;
  const uint32_t* row;   // current entry buffer
  const uint32_t* aux;
  const int64_t* ey;
  // device-described level: the slice pointers above are the level-major
  // bases and *off_dev the level's offset (NULL: already offset)
  const int64_t* off_dev;
};

struct ItemArgs {
  const uint32_t* list;  // optional: item group -> open position
  int64_t n_groups;
  int per_group;         // draws per group in this phase
  int j0;                // first draw index of this phase
  int D;                 // result stride per open node
  float* res_score;
  float* res_thr;
  int32_t* res_feat;
  unsigned long long* bytes;  // algorithmic bytes counter (may be NULL)
  int node_major;        // 1: a grid of n_groups teams, each team walks its own group's draws in order
  uint32_t* scr_key;     // big-item scratch, one slab per block
  uint32_t* scr_aux;
  int64_t* scr_y;
  int64_t scr_stride;
  const unsigned int* n_dev;  // device-described level: n_groups = *n_dev
};

// Level slice and item count of a device-described level (a no-op for
// launches whose arguments carry them).
__device__ __forceinline__ void resolve_level(LevelArgs& lv, int C) {
  if (lv.off_dev) {
    const int64_t off = *lv.off_dev;
    lv.tree += off;
    lv.id += off;
    lv.count += off;
    lv.m += off;
    if (lv.cls) lv.cls += off * C;
    if (lv.sum) lv.sum += off;
    lv.off_dev = nullptr;
  }
}
__device__ __forceinline__ void resolve_items(ItemArgs& it) {
  if (it.n_dev) it.n_groups = *it.n_dev;
}

// Draw index of item position q of a node whose forest data is dd: j0 + q
// for a fit of one forest.  A joint build pads every node to the widest
// forest's draws: the first phase covers draws 0 .. D0-1 of the node's own
// forest, the retry phase (j0 > 0) its draws D0 .. D0+R-1; -1 marks a
// position past them.
__device__ __forceinline__ int item_draw(const DataArgs& d, const DataArgs& dd, const ItemArgs& it, int q) {
  if (!d.forests) return it.j0 + q;
  if (it.j0 == 0) return q < dd.D0 ? q : -1;
  return q < dd.R ? dd.D0 + q : -1;
}

// ------------------------------------------------- per-node context -----
// Fixed-point x ln x with the fit-wide shift F.  Counts never exceed the
// root's total weight, so a per-fit table T[x] = xlnx_fixed(x, F) for
// x <= Tmax (built by k_build_T with the same function) answers every
// lookup with one cached load; counts above a capped table fall back to the
// computation -- the same bits either way.
__device__ __attribute__((noinline)) int64_t xlnx_fixed_slow(uint64_t x, int F, const double* inv,
                                                           const double* logc) {
  return xlnx_fixed(x, F, inv, logc);
}

struct LnCtx {
  const double* inv;
  const double* logc;
  const int64_t* T;
  int64_t Tmax;
  int F;
  __device__ __forceinline__ int64_t t(uint64_t x) const {
    return x <= (uint64_t)Tmax ? T[x] : xlnx_fixed_slow(x, F, inv, logc);
  }
};

// The same with a copy of T[0, nTL) in LDS, for kernels whose nodes are
// small (their lookups mostly fall below nTL).
struct LnCtxL {
  const double* inv;
  const double* logc;
  const int64_t* T;
  int64_t Tmax;
  int F;
  const int64_t* TL;
  int64_t nTL;
  __device__ __forceinline__ int64_t t(uint64_t x) const {
    if (x < (uint64_t)nTL) return TL[x];
    return x <= (uint64_t)Tmax ? T[x] : xlnx_fixed_slow(x, F, inv, logc);
  }
};

// Stages T[0, min(n, Tmax + 1)) into LDS; returns the staged length (0 when
// the fit has no table: regression).
template <int NT>
__device__ __forceinline__ int64_t load_t_prefix(int64_t* s_T, const int64_t* T, int64_t Tmax) {
  if (!T) return 0;
  const int64_t n = Tmax + 1 < NT ? Tmax + 1 : NT;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) s_T[i] = T[i];
  return n;
}

struct NodeCtx {
  int64_t n;        // sum of weights
  int64_t ssum;     // regression: fixed-point target sum
  int64_t P;        // classification: T(n) - sum_c T(n_c)
  double inv;       // 1 / (n ln 2)
  int C;
};

template <int MAXC, int TASK>
struct Carry {
  uint32_t run[MAXC];  // classification: class counts before the chunk
  int64_t run_n;       // weights before the chunk
  int64_t run_s;       // regression: target sum before the chunk
  uint32_t last_key;   // last key before the chunk (valid if run_n > 0)
  double best_g;
  uint32_t best_key;
  uint32_t best_prev;
};

// Per-team LDS of the sorted-chunk evaluator.
template <int MAXC, int TASK, int TEAM, int IPT>
struct TeamLds {
  static constexpr int CAP = TEAM * IPT;
  // 16-bit counters: 16 radix digits per thread, or the packed counters of a
  // counting sort over a key span of up to SPAN.
  static constexpr int SPAN = 16 * TEAM > (4 * CAP < 2048 ? 4 * CAP : 2048) ? 16 * TEAM
                                                                           : (4 * CAP < 2048 ? 4 * CAP : 2048);
  union {
    struct {
      uint32_t k[2][CAP];
      uint16_t v[2][CAP];
      uint16_t cnt[SPAN];
    } srt;
    uint32_t hist[TEAM == 256 ? kHistWords : 1];
    uint32_t bpos[TEAM == 256 ? kMaxBuckets : 1];
  } u;
  uint32_t aux[CAP];
  int64_t y[TASK == 1 ? CAP : 1];
  // Per-lane class columns [c * TEAM + t] of the walk; a wave team of up to
  // 16 classes keeps them in the sort counters, free once the chunk is sorted.
  static constexpr bool kColsInCnt = TEAM == kWave && MAXC * TEAM * 4 <= SPAN * 2;
  uint32_t pre[TASK == 0 && !kColsInCnt ? MAXC * TEAM : 1];
  uint32_t ncls[MAXC];
  uint32_t ctot[MAXC];
  uint32_t last[TEAM == kWave ? 1 : TEAM];  // histogram walk of whole-workgroup teams
  __device__ __forceinline__ uint32_t* cols() {
    if constexpr (kColsInCnt) {
      return reinterpret_cast<uint32_t*>(u.srt.cnt);
    } else {
      return pre;
    }
  }
  __device__ __forceinline__ const uint32_t* cols() const { return const_cast<TeamLds*>(this)->cols(); }
  uint32_t w[8 * kWaves + kWaves];
  int64_t w64[kWaves];
  int64_t w64v[4 * kWaves];
  double bg[kWaves];
  uint32_t bk[kWaves];
  uint32_t bp[kWaves];
  Carry<MAXC, TASK> cr;
};

// Exclusive scan of the per-lane class columns pre[c][t] across the team,
// plus the carry; chunk totals go to ctot[c].
template <int G, int MAXC, int TASK, int TEAM, int IPT>
__device__ __forceinline__ void scan_class_group(TeamLds<MAXC, TASK, TEAM, IPT>& L, int c0, int C) {
  const int t = team_tid<TEAM>();
  uint32_t v[G], tot[G];
#pragma unroll
  for (int q = 0; q < G; ++q) v[q] = (c0 + q < C) ? L.cols()[(c0 + q) * TEAM + t] : 0u;
  team_excl_scan_vec<TEAM, G>(v, L.w, tot);
#pragma unroll
  for (int q = 0; q < G; ++q)
    if (c0 + q < C) {
      L.cols()[(c0 + q) * TEAM + t] = v[q] + L.cr.run[c0 + q];
      if (t == 0) L.ctot[c0 + q] = tot[q];
    }
}

// Groups of 8 columns, then one group of 4, 2 or 1 for the rest, so that C
// classes take C scans.
template <int MAXC, int TASK, int TEAM, int IPT>
__device__ void scan_class_columns(TeamLds<MAXC, TASK, TEAM, IPT>& L, int C) {
  int c0 = 0;
  if constexpr (MAXC >= 8)
    for (; c0 + 8 <= C; c0 += 8) scan_class_group<8>(L, c0, C);
  while (c0 < C) {
    const int r = C - c0;
    if (MAXC >= 4 && r > 2) {
      scan_class_group<4>(L, c0, C);
      c0 += 4;
    } else if (r == 2) {
      scan_class_group<2>(L, c0, C);
      c0 += 2;
    } else {
      scan_class_group<1>(L, c0, C);
      c0 += 1;
    }
  }
}

// Incremental state of one lane's walk: exact fixed-point sums of the left
// and right x ln x terms.
struct WalkState {
  int64_t sl, sr;
  int64_t nl;
};

template <int MAXC, int TASK, int TEAM, int IPT, class LX>
__device__ __forceinline__ WalkState walk_init(const TeamLds<MAXC, TASK, TEAM, IPT>& L, const NodeCtx& nc,
                                               const LX& lx) {
  const int t = team_tid<TEAM>();
  WalkState w{0, 0, 0};
  for (int c = 0; c < nc.C; ++c) {
    const uint32_t l = L.cols()[c * TEAM + t];
    w.nl += l;
    w.sl += lx.t(l);
    w.sr += lx.t(L.ncls[c] - l);
  }
  return w;
}

template <class LX>
__device__ __forceinline__ double walk_gain(const WalkState& w, const NodeCtx& nc, const LX& lx) {
  const int64_t q = (w.sl - lx.t((uint64_t)w.nl)) + (w.sr - lx.t((uint64_t)(nc.n - w.nl)));
  return ig_from_fixed(nc.P + q, lx.F, nc.inv);
}

template <int MAXC, int TASK, int TEAM, int IPT, class LX>
__device__ __forceinline__ void walk_add(WalkState& w, TeamLds<MAXC, TASK, TEAM, IPT>& L, int lab, uint32_t add,
                                         const LX& lx) {
  const int t = team_tid<TEAM>();
  const uint32_t lo = L.cols()[lab * TEAM + t];
  const uint32_t ln = lo + add;
  L.cols()[lab * TEAM + t] = ln;
  const uint32_t tot = L.ncls[lab];
  w.sl += lx.t(ln) - lx.t(lo);
  w.sr += lx.t(tot - ln) - lx.t(tot - lo);
  w.nl += add;
}

template <int MAXC, int TASK>
__device__ __forceinline__ void carry_best(Carry<MAXC, TASK>& cr, const BestCand& b) {
  if (b.key != kNone && cand_better(b.g, b.key, cr.best_g, cr.best_key)) {
    cr.best_g = b.g;
    cr.best_key = b.key;
    cr.best_prev = b.prev;
  }
}

// Classification walk of a key-sorted chunk on a one-wave team.  The class
// counts before each position come from the per-lane class columns (lane t
// owns a run of consecutive positions) and go to LDS; the x ln x reads and
// the running sums then go row by row, lane t on position 64 * i + t, so
// that the 64 reads of one instruction are at neighbouring counts (a few
// cache lines) instead of 64 runs ipt positions apart.  The candidates and
// their order are those of the run-per-lane walk: positions ascend within a
// lane, and the best candidate is a maximum under a total order.
template <int MAXC, int TASK, int IPT, class LX>
__device__ void eval_sorted_wave(const uint32_t* keys, const uint16_t* vals, int n, uint32_t kbase,
                                 const NodeCtx& nc, TeamLds<MAXC, TASK, kWave, IPT>& L, const LX& lx) {
  const int t = threadIdx.x & (kWave - 1);
  const int ipt = lane_span<kWave>(n);
  const int lo = t * ipt;
  const int hi = min(lo + ipt, n);
  // Counts before each position, in the key buffer the sort left free.
  uint32_t* lbp = keys == L.u.srt.k[0] ? L.u.srt.k[1] : L.u.srt.k[0];
  for (int c = 0; c < nc.C; ++c) L.cols()[c * kWave + t] = 0;
  for (int p = lo; p < hi; ++p) {
    const uint32_t a = L.aux[vals[p]];
    L.cols()[(a >> 16) * kWave + t] += a & 0xFFFFu;
  }
  scan_class_columns<MAXC, TASK, kWave, IPT>(L, nc.C);
  for (int p = lo; p < hi; ++p) {
    const uint32_t a = L.aux[vals[p]];
    const int c = (int)(a >> 16);
    const uint32_t l = L.cols()[c * kWave + t];
    lbp[p] = l;
    L.cols()[c * kWave + t] = l + (a & 0xFFFFu);
  }
  // Left and right sums and the weight before the chunk (one class per lane).
  int64_t s0 = 0;
  uint32_t n0 = 0;
  if (t < nc.C) {
    n0 = L.cr.run[t];
    s0 = lx.t(n0) + lx.t(L.ncls[t] - n0);
  }
  s0 = wave_last_i64(wave_incl_scan_i64(s0));
  n0 = wave_last(wave_incl_scan(n0));
  team_sync<kWave>();
  // Row pass 1: every x ln x read of the chunk, independent of each other.
  const int rows = (n + kWave - 1) / kWave;
  int64_t dd[IPT], tq[IPT];
  uint32_t nrun = n0;
#pragma unroll
  for (int i = 0; i < IPT; ++i) {
    dd[i] = 0;
    tq[i] = 0;
    if (i < rows) {
      const int p = i * kWave + t;
      uint32_t w = 0, lb = 0, tc = 0;
      if (p < n) {
        const uint32_t a = L.aux[vals[p]];
        w = a & 0xFFFFu;
        lb = lbp[p];
        tc = L.ncls[a >> 16];
      }
      const uint32_t inc = wave_incl_scan(w);
      const uint32_t nlb = nrun + inc - w;
      nrun += wave_last(inc);
      if (p < n) {
        dd[i] = (lx.t(lb + w) - lx.t(lb)) + (lx.t(tc - lb - w) - lx.t(tc - lb));
        tq[i] = nlb > 0 ? lx.t((uint64_t)nlb) + lx.t((uint64_t)(nc.n - nlb)) : -1;
      }
    }
  }
  // Row pass 2: running sums, the gain at each boundary between distinct keys.
  BestCand tb{0.0, kNone, kNone};
  int64_t best_pq = 0;
  int64_t srun = s0;
#pragma unroll
  for (int i = 0; i < IPT; ++i) {
    if (i < rows) {
      const int p = i * kWave + t;
      const int64_t inc = wave_incl_scan_i64(dd[i]);
      const int64_t before = srun + inc - dd[i];
      srun += wave_last_i64(inc);
      if (p < n && tq[i] >= 0) {
        const uint32_t k = keys[p] + kbase;
        const uint32_t prevk = p > 0 ? keys[p - 1] + kbase : L.cr.last_key;
        if (prevk < k) {
          const int64_t pq = nc.P + before - tq[i];
          if (pq > best_pq) {
            const double g = ig_from_fixed(pq, lx.F, nc.inv);
            if (g > tb.g) {
              tb = BestCand{g, k, prevk};
              best_pq = pq;
            }
          }
        }
      }
    }
  }
  const BestCand b = team_best<kWave>(tb, L.bg, L.bk, L.bp);
  if (t == 0) {
    carry_best(L.cr, b);
    for (int c = 0; c < nc.C; ++c) {
      L.cr.run[c] += L.ctot[c];
      L.cr.run_n += L.ctot[c];
    }
    if (n > 0) L.cr.last_key = keys[n - 1] + kbase;
  }
  team_sync<kWave>();
}

// Evaluate every boundary of a key-sorted chunk (keys relative to kbase,
// vals index the chunk's payload in L.aux / L.y).
template <int MAXC, int TASK, int TEAM, int IPT, class LX>
__device__ void eval_sorted_chunk(const uint32_t* keys, const uint16_t* vals, int n, uint32_t kbase,
                                  const NodeCtx& nc, TeamLds<MAXC, TASK, TEAM, IPT>& L, const LX& lx) {
  const int t = team_tid<TEAM>();
  const int ipt = lane_span<TEAM>(n);
  const int lo = t * ipt;
  const int hi = min(lo + ipt, n);
  BestCand tb{0.0, kNone, kNone};
  if constexpr (TASK == 0 && TEAM == kWave) {
    eval_sorted_wave<MAXC, TASK, IPT>(keys, vals, n, kbase, nc, L, lx);
  } else if constexpr (TASK == 0) {
    for (int c = 0; c < nc.C; ++c) L.cols()[c * TEAM + t] = 0;
    for (int p = lo; p < hi; ++p) {
      const uint32_t a = L.aux[vals[p]];
      L.cols()[(a >> 16) * TEAM + t] += a & 0xFFFFu;
    }
    scan_class_columns<MAXC, TASK, TEAM, IPT>(L, nc.C);
    // Three passes so that the x ln x lookups of all positions are
    // independent: (1) counts before each position, (2) the lane's deltas of
    // the left / right sums (plus one exclusive scan that gives every lane
    // its starting sums), (3) running sums and the gain at each boundary.
    uint32_t lb[IPT], tc[IPT], wv[IPT];
    int64_t nlb[IPT];
    if (lo < hi) {
      int64_t nl = 0;
      for (int c = 0; c < nc.C; ++c) nl += L.cols()[c * TEAM + t];
#pragma unroll
      for (int i = 0; i < IPT; ++i) {
        const int p = lo + i;
        if (p < hi) {
          const uint32_t a = L.aux[vals[p]];
          const int c = (int)(a >> 16);
          const uint32_t wt = a & 0xFFFFu;
          lb[i] = L.cols()[c * TEAM + t];
          L.cols()[c * TEAM + t] = lb[i] + wt;
          tc[i] = L.ncls[c];
          wv[i] = wt;
          nlb[i] = nl;
          nl += wt;
        }
      }
    }
    int64_t sc[4] = {0, 0, 0, 0};  // lane deltas of sl, sr; sums at the chunk start
    int64_t dl[IPT], dr[IPT];
    int64_t tq[IPT];  // T(nl) + T(n - nl) before each position, read with the deltas
#pragma unroll
    for (int i = 0; i < IPT; ++i) {
      if (lo + i < hi) {
        dl[i] = lx.t(lb[i] + wv[i]) - lx.t(lb[i]);
        dr[i] = lx.t(tc[i] - lb[i] - wv[i]) - lx.t(tc[i] - lb[i]);
        tq[i] = nlb[i] > 0 ? lx.t((uint64_t)nlb[i]) + lx.t((uint64_t)(nc.n - nlb[i])) : 0;
        sc[0] += dl[i];
        sc[1] += dr[i];
      }
    }
    if (t < nc.C) {  // one class per lane
      sc[2] = lx.t(L.cr.run[t]);
      sc[3] = lx.t(L.ncls[t] - L.cr.run[t]);
    }
    int64_t tot[4];
    team_excl_scan_i64_vec<TEAM, 4>(sc, L.w64v, tot);
    if (lo < hi) {
      int64_t sl = tot[2] + sc[0], sr = tot[3] + sc[1];
      uint32_t prevk = (lo > 0) ? keys[lo - 1] + kbase : L.cr.last_key;
      // The gain is a non-decreasing function of the exact sum P + q, and a
      // later boundary only wins with a strictly larger gain: a sum below
      // the lane's best sum cannot win.
      int64_t best_pq = 0;
#pragma unroll
      for (int i = 0; i < IPT; ++i) {
        const int p = lo + i;
        if (p < hi) {
          const uint32_t k = keys[p] + kbase;
          if (nlb[i] > 0 && prevk < k) {
            const int64_t pq = nc.P + sl + sr - tq[i];
            if (pq > best_pq) {
              const double g = ig_from_fixed(pq, lx.F, nc.inv);
              if (g > tb.g) {
                tb = BestCand{g, k, prevk};
                best_pq = pq;
              }
            }
          }
          sl += dl[i];
          sr += dr[i];
          prevk = k;
        }
      }
    }
    const BestCand b = team_best<TEAM>(tb, L.bg, L.bk, L.bp);
    if (t == 0) {
      carry_best(L.cr, b);
      for (int c = 0; c < nc.C; ++c) {
        L.cr.run[c] += L.ctot[c];
        L.cr.run_n += L.ctot[c];
      }
      if (n > 0) L.cr.last_key = keys[n - 1] + kbase;
    }
    team_sync<TEAM>();
  } else {
    int64_t ln = 0, ls = 0;
    for (int p = lo; p < hi; ++p) {
      const uint16_t v = vals[p];
      const int64_t w = (int64_t)(L.aux[v] & 0xFFFFu);
      ln += w;
      ls += w * L.y[v];
    }
    int64_t tot_n, tot_s;
    int64_t pn = team_excl_scan_i64<TEAM>(ln, L.w64, &tot_n);
    int64_t ps = team_excl_scan_i64<TEAM>(ls, L.w64, &tot_s);
    pn += L.cr.run_n;
    ps += L.cr.run_s;
    if (lo < hi) {
      uint32_t prevk = (lo > 0) ? keys[lo - 1] + kbase : L.cr.last_key;
      for (int p = lo; p < hi; ++p) {
        const uint32_t k = keys[p] + kbase;
        if (pn > 0 && prevk < k) {
          const double g = variance_gain(pn, ps, nc.n - pn, nc.ssum - ps, 0);
          if (g > tb.g) tb = BestCand{g, k, prevk};
        }
        const uint16_t v = vals[p];
        const int64_t w = (int64_t)(L.aux[v] & 0xFFFFu);
        pn += w;
        ps += w * L.y[v];
        prevk = k;
      }
    }
    const BestCand b = team_best<TEAM>(tb, L.bg, L.bk, L.bp);
    if (t == 0) {
      carry_best(L.cr, b);
      L.cr.run_n += tot_n;
      L.cr.run_s += tot_s;
      if (n > 0) L.cr.last_key = keys[n - 1] + kbase;
    }
    team_sync<TEAM>();
  }
}

// One bucket of W key slots through an LDS histogram (bucket holds more
// elements than the sort capacity: heavy ties).  Classification:
// hist[slot * C + class]; regression: counts in words [0, W), 64-bit sums
// from word W on.  Whole-workgroup teams only.
template <int MAXC, int TASK, int IPT>
__device__ void eval_hist_bucket(const uint32_t* skey, const uint32_t* saux, const int64_t* sy, int64_t n_el,
                                 uint32_t kbase, int W, const NodeCtx& nc, TeamLds<MAXC, TASK, 256, IPT>& L,
                                 const LnCtx& lx) {
  const int t = threadIdx.x;
  const int C = nc.C;
  uint32_t* s_hist = L.u.hist;
  const int zero_words = (TASK == 0) ? W * C : 3 * W;
  for (int i = t; i < zero_words; i += 256) s_hist[i] = 0;
  __syncthreads();
  unsigned long long* hsum = (unsigned long long*)(s_hist + W);
  for (int64_t e = t; e < n_el; e += 256) {
    const uint32_t slot = skey[e] - kbase;
    const uint32_t a = saux[e];
    const uint32_t w = a & 0xFFFFu;
    if (TASK == 0) {
      atomicAdd(&s_hist[slot * C + (a >> 16)], w);
    } else {
      atomicAdd(&s_hist[slot], w);
      atomicAdd(&hsum[slot], (unsigned long long)((int64_t)w * sy[e]));
    }
  }
  __syncthreads();
  const int spt = (W + 255) / 256;
  const int s0 = t * spt, s1 = min(s0 + spt, W);
  uint32_t my_last = 0;  // last non-empty slot key + 1 in my range, 0 = none
  BestCand tb{0.0, kNone, kNone};
  if constexpr (TASK == 0) {
    for (int c = 0; c < C; ++c) L.cols()[c * 256 + t] = 0;
    for (int sl = s0; sl < s1; ++sl) {
      uint32_t any = 0;
      for (int c = 0; c < C; ++c) {
        const uint32_t h = s_hist[sl * C + c];
        L.cols()[c * 256 + t] += h;
        any |= h;
      }
      if (any) my_last = kbase + sl + 1;
    }
    L.last[t] = my_last;
    scan_class_columns<MAXC, TASK, 256, IPT>(L, C);
    uint32_t prev_last = 0;
    for (int i = 0; i < t; ++i) prev_last = max(prev_last, L.last[i]);
    if (s0 < s1) {
      WalkState w = walk_init<MAXC, TASK, 256, IPT>(L, nc, lx);
      uint32_t prevk = prev_last ? prev_last - 1 : L.cr.last_key;
      for (int sl = s0; sl < s1; ++sl) {
        uint32_t any = 0;
        for (int c = 0; c < C; ++c) any |= s_hist[sl * C + c];
        if (!any) continue;
        const uint32_t k = kbase + sl;
        if (w.nl > 0 && prevk < k) {
          const double g = walk_gain(w, nc, lx);
          if (g > tb.g) tb = BestCand{g, k, prevk};
        }
        for (int c = 0; c < C; ++c) {
          const uint32_t h = s_hist[sl * C + c];
          if (h) walk_add<MAXC, TASK, 256, IPT>(w, L, c, h, lx);
        }
        prevk = k;
      }
    }
    const BestCand b = team_best<256>(tb, L.bg, L.bk, L.bp);
    if (t == 0) {
      carry_best(L.cr, b);
      uint32_t last = 0;
      for (int i = 0; i < 256; ++i) last = max(last, L.last[i]);
      for (int c = 0; c < C; ++c) {
        L.cr.run[c] += L.ctot[c];
        L.cr.run_n += L.ctot[c];
      }
      if (last) L.cr.last_key = last - 1;
    }
    __syncthreads();
  } else {
    int64_t ln = 0, ls = 0;
    for (int sl = s0; sl < s1; ++sl) {
      const uint32_t hc = s_hist[sl];
      if (hc) {
        ln += hc;
        ls += (int64_t)hsum[sl];
        my_last = kbase + sl + 1;
      }
    }
    L.last[t] = my_last;
    int64_t tot_n, tot_s;
    int64_t pn = team_excl_scan_i64<256>(ln, L.w64, &tot_n);
    int64_t ps = team_excl_scan_i64<256>(ls, L.w64, &tot_s);
    uint32_t prev_last = 0;
    for (int i = 0; i < t; ++i) prev_last = max(prev_last, L.last[i]);
    pn += L.cr.run_n;
    ps += L.cr.run_s;
    uint32_t prevk = prev_last ? prev_last - 1 : L.cr.last_key;
    for (int sl = s0; sl < s1; ++sl) {
      const uint32_t hc = s_hist[sl];
      if (!hc) continue;
      const uint32_t k = kbase + sl;
      if (pn > 0 && prevk < k) {
        const double g = variance_gain(pn, ps, nc.n - pn, nc.ssum - ps, 0);
        if (g > tb.g) tb = BestCand{g, k, prevk};
      }
      pn += hc;
      ps += (int64_t)hsum[sl];
      prevk = k;
    }
    const BestCand b = team_best<256>(tb, L.bg, L.bk, L.bp);
    if (t == 0) {
      carry_best(L.cr, b);
      uint32_t last = 0;
      for (int i = 0; i < 256; ++i) last = max(last, L.last[i]);
      L.cr.run_n += tot_n;
      L.cr.run_s += tot_s;
      if (last) L.cr.last_key = last - 1;
    }
    __syncthreads();
  }
}

// Sort a chunk (keys relative to kbase, span < 2^bits) held in LDS buffer 0
// and evaluate it.
template <int MAXC, int TASK, int TEAM, int IPT, bool DIGIT = false, class LX = LnCtx>
__device__ void sort_and_eval(int n, int bits, uint32_t kbase, const NodeCtx& nc,
                              TeamLds<MAXC, TASK, TEAM, IPT>& L, const LX& lx) {
  constexpr int kSpan = TeamLds<MAXC, TASK, TEAM, IPT>::SPAN;
  if (bits > 4 && (1 << bits) <= kSpan) {
    lds_count_sort<IPT, TEAM>(L.u.srt.k[0], L.u.srt.v[0], L.u.srt.k[1], L.u.srt.v[1], n, 1 << bits,
                              reinterpret_cast<uint32_t*>(L.u.srt.cnt), L.w);
    eval_sorted_chunk<MAXC, TASK, TEAM, IPT>(L.u.srt.k[1], L.u.srt.v[1], n, kbase, nc, L, lx);
    return;
  }
  // Wide key spans: one counting pass on the top digit, then short runs
  // ordered in place; radix passes when a run is long (clustered keys).
  if (DIGIT && bits > 8 &&
      lds_digit_sort<IPT, TEAM>(L.u.srt.k[0], L.u.srt.v[0], L.u.srt.k[1], L.u.srt.v[1], n, bits, kSpan, kMaxDigitRun,
                                reinterpret_cast<uint32_t*>(L.u.srt.cnt), L.w)) {
    eval_sorted_chunk<MAXC, TASK, TEAM, IPT>(L.u.srt.k[0], L.u.srt.v[0], n, kbase, nc, L, lx);
    return;
  }
  int cur = 0;
  for (int sh = 0; sh < bits; sh += 4) {
    lds_sort_pass<IPT, TEAM>(L.u.srt.k[cur], L.u.srt.v[cur], L.u.srt.k[cur ^ 1], L.u.srt.v[cur ^ 1], n, sh,
                             L.u.srt.cnt, L.w);
    cur ^= 1;
  }
  eval_sorted_chunk<MAXC, TASK, TEAM, IPT>(L.u.srt.k[cur], L.u.srt.v[cur], n, kbase, nc, L, lx);
}

__device__ __forceinline__ int bit_length(uint32_t v) { return v ? 32 - __clz((int)v) : 0; }

__device__ __forceinline__ void load_tables(double* s_inv, double* s_logc) {
  for (int i = threadIdx.x; i < 128; i += blockDim.x) {
    s_inv[i] = g_ln_table.inv[i];
    s_logc[i] = g_ln_table.logc[i];
  }
}

// Node statistics of one item, computed by every lane of the team.
template <int MAXC, int TASK, int TEAM, int IPT, class LX>
__device__ __forceinline__ bool node_context(const DataArgs& d, const LevelArgs& lv, uint32_t i, NodeCtx& nc,
                                             TeamLds<MAXC, TASK, TEAM, IPT>& L, const LX& lx) {
  const int t = team_tid<TEAM>();
  nc.C = d.C;
  nc.n = lv.count[i];
  nc.ssum = 0;
  nc.P = 0;
  nc.inv = 0.0;
  if constexpr (TASK == 0) {
    const uint32_t my = t < d.C ? lv.cls[(int64_t)i * d.C + t] : 0u;
    if (t < d.C) L.ncls[t] = my;
    const uint32_t mx = team_reduce_max<TEAM>(my, L.w);
    const int64_t tn = t < d.C ? lx.t(my) : 0;
    int64_t tsum = 0;  // sum of T(n_c) over the classes (lanes < C)
    (void)team_excl_scan_i64<TEAM>(tn, L.w64, &tsum);
    nc.P = lx.t((uint64_t)nc.n) - tsum;
    nc.inv = 1.0 / ((double)nc.n * kLn2);
    return (int64_t)mx == nc.n;  // pure node
  } else {
    nc.ssum = lv.sum[i];
    return false;
  }
}

// Evaluate one RF item on a team; returns through L.cr.  BIG: the item may
// exceed the team's LDS capacity (whole-workgroup teams of the plain driver):
// bucket it through the workgroup's scratch slab.
template <int MAXC, int TASK, int TEAM, int IPT, bool BIG>
__device__ void rf_item(const DataArgs& d, const LevelArgs& lv, const ItemArgs& it, int f, uint32_t i, const NodeCtx& nc,
                        TeamLds<MAXC, TASK, TEAM, IPT>& L, const LnCtx& lx, uint32_t* s_boff, int team_index) {
  constexpr int CAP = TeamLds<MAXC, TASK, TEAM, IPT>::CAP;
  const int t = team_tid<TEAM>();
  const int m = lv.m[i];
  const int64_t st = lv.start[i];
  const uint32_t* Kf = d.K + (int64_t)f * d.N;
  if (t == 0) {
    for (int c = 0; c < MAXC; ++c) L.cr.run[c] = 0;
    L.cr.run_n = 0;
    L.cr.run_s = 0;
    L.cr.last_key = 0;
    L.cr.best_g = 0.0;
    L.cr.best_key = kNone;
    L.cr.best_prev = kNone;
  }
  if (m <= CAP) {
    // Gather: independent loads first (rows, payloads), then the keys.
    uint32_t kmin = kNone, kmax = 0;
    uint32_t rows[IPT];
#pragma unroll
    for (int q = 0; q < IPT; ++q) {
      const int e = t + q * TEAM;
      if (e < m) {
        rows[q] = lv.row[st + e];
        L.aux[e] = lv.aux[st + e];
        if (TASK == 1) L.y[e] = lv.ey[st + e];
      }
    }
#pragma unroll
    for (int q = 0; q < IPT; ++q) {
      const int e = t + q * TEAM;
      if (e < m) {
        const uint32_t key = Kf[rows[q]];
        L.u.srt.k[0][e] = key;
        L.u.srt.v[0][e] = (uint16_t)e;
        kmin = min(kmin, key);
        kmax = max(kmax, key);
      }
    }
    kmin = team_reduce_min<TEAM>(kmin, L.w);
    kmax = team_reduce_max<TEAM>(kmax, L.w);
    for (int e = t; e < m; e += TEAM) L.u.srt.k[0][e] -= kmin;
    team_sync<TEAM>();
    sort_and_eval<MAXC, TASK, TEAM, IPT>(m, bit_length(kmax - kmin), kmin, nc, L, lx);
    return;
  }
  if constexpr (TEAM == 256 && BIG) {
    // Big node: bucket by dense key into this block's scratch slab, then
    // pack consecutive buckets into LDS-sized chunks.
    const uint32_t S = d.S[f];
    const int words = (TASK == 0) ? d.C : 3;
    int wbits = 0;
    while ((1 << (wbits + 1)) * words <= kHistWords) ++wbits;
    while (((S + (1u << wbits) - 1) >> wbits) > (uint32_t)kMaxBuckets) ++wbits;
    const int W = 1 << wbits;
    const int NB = (int)((S + (uint32_t)W - 1) >> wbits);
    uint32_t* skey = it.scr_key + (int64_t)team_index * it.scr_stride;
    uint32_t* saux = it.scr_aux + (int64_t)team_index * it.scr_stride;
    int64_t* sy = TASK == 1 ? it.scr_y + (int64_t)team_index * it.scr_stride : nullptr;
    for (int b = t; b <= NB; b += 256) s_boff[b] = 0;
    __syncthreads();
    constexpr int U = 8;
    for (int e0 = 0; e0 < m; e0 += 256 * U) {
      uint32_t rows[U];
#pragma unroll
      for (int q = 0; q < U; ++q) {
        const int e = e0 + t + q * 256;
        rows[q] = e < m ? lv.row[st + e] : 0u;
      }
#pragma unroll
      for (int q = 0; q < U; ++q) {
        const int e = e0 + t + q * 256;
        if (e < m) atomicAdd(&s_boff[Kf[rows[q]] >> wbits], 1u);
      }
    }
    __syncthreads();
    {
      const int per = (NB + 255) / 256;
      const int b0 = t * per, b1 = min(b0 + per, NB);
      uint32_t loc = 0;
      for (int b = b0; b < b1; ++b) loc += s_boff[b];
      uint32_t base = team_excl_scan<256>(loc, L.w, nullptr);
      for (int b = b0; b < b1; ++b) {
        const uint32_t c = s_boff[b];
        s_boff[b] = base;
        base += c;
      }
      if (t == 0) s_boff[NB] = (uint32_t)m;
      __syncthreads();
      for (int b = t; b < NB; b += 256) L.u.bpos[b] = s_boff[b];
      __syncthreads();
    }
    for (int e0 = 0; e0 < m; e0 += 256 * U) {
      uint32_t rows[U], auxv[U], keys[U];
      int64_t yv[U];
#pragma unroll
      for (int q = 0; q < U; ++q) {
        const int e = e0 + t + q * 256;
        if (e < m) {
          rows[q] = lv.row[st + e];
          auxv[q] = lv.aux[st + e];
          if (TASK == 1) yv[q] = lv.ey[st + e];
        }
      }
#pragma unroll
      for (int q = 0; q < U; ++q) {
        const int e = e0 + t + q * 256;
        if (e < m) keys[q] = Kf[rows[q]];
      }
#pragma unroll
      for (int q = 0; q < U; ++q) {
        const int e = e0 + t + q * 256;
        if (e < m) {
          const uint32_t pos = atomicAdd(&L.u.bpos[keys[q] >> wbits], 1u);
          skey[pos] = keys[q];
          saux[pos] = auxv[q];
          if (TASK == 1) sy[pos] = yv[q];
        }
      }
    }
    __syncthreads();
    int b = 0;
    while (b < NB) {
      const uint32_t cb = s_boff[b + 1] - s_boff[b];
      if (cb == 0) {
        ++b;
        continue;
      }
      if (cb > (uint32_t)CAP) {
        eval_hist_bucket<MAXC, TASK, IPT>(skey + s_boff[b], saux + s_boff[b], TASK == 1 ? sy + s_boff[b] : nullptr,
                                          cb, (uint32_t)b << wbits, W, nc, L, lx);
        ++b;
        continue;
      }
      int be = b;
      uint32_t tot = 0;
      while (be < NB) {
        const uint32_t c = s_boff[be + 1] - s_boff[be];
        if (c > (uint32_t)CAP || tot + c > (uint32_t)CAP) break;
        tot += c;
        ++be;
      }
      const uint32_t e0 = s_boff[b];
      const uint32_t kbase = (uint32_t)b << wbits;
      for (int e = t; e < (int)tot; e += 256) {
        L.u.srt.k[0][e] = skey[e0 + e] - kbase;
        L.u.srt.v[0][e] = (uint16_t)e;
        L.aux[e] = saux[e0 + e];
        if (TASK == 1) L.y[e] = sy[e0 + e];
      }
      __syncthreads();
      sort_and_eval<MAXC, TASK, 256, IPT>((int)tot, bit_length((uint32_t)(be - b) * (uint32_t)W - 1u), kbase, nc,
                                          L, lx);
      b = be;
    }
  }
}

template <int MAXC, int TASK, int TEAM, int IPT>
__device__ __forceinline__ void rf_write_result(const DataArgs& d, const ItemArgs& it, int64_t ridx, int f,
                                                const Carry<MAXC, TASK>& cr) {
  float score = 0.0f, thr = 0.0f;
  if (cr.best_key != kNone) {
    const float* sv = d.SV + d.SVoff[f];
    score = (float)(TASK == 0 ? cr.best_g : ldexp_d(cr.best_g, -2 * d.shift));
    thr = midpoint(sv[cr.best_prev], sv[cr.best_key]);
  }
  it.res_score[ridx] = score;
  it.res_thr[ridx] = thr;
  it.res_feat[ridx] = f;
}

// ------------------------------------------------- RF split search ------
// One team per (node, feature draw) item, grid-stride over items.  TEAM=64:
// four independent wave teams per workgroup for medium nodes (up to 64*IPT
// entries).  TEAM=256: whole workgroup; nodes up to 256*IPT entries are
// sorted in LDS, larger ones are bucketed by dense key into a scratch slab
// first.  Every boundary between distinct values is scored.
template <int MAXC, int TASK, int TEAM, bool BIG>
__global__ void __launch_bounds__(kBlock) k_search_rf(DataArgs d, LevelArgs lv, ItemArgs it) {
  constexpr int IPT = TEAM == kWave ? kTeamIpt : kIpt;
  constexpr int NT = kBlock / TEAM;
  resolve_level(lv, d.C);
  resolve_items(it);
  if ((int64_t)blockIdx.x * NT >= it.n_groups * it.per_group) return;
  __shared__ double s_inv[128], s_logc[128];
  __shared__ TeamLds<MAXC, TASK, TEAM, IPT> s_team[NT];
  __shared__ uint32_t s_boff[BIG ? kMaxBuckets + 1 : 1];
  load_tables(s_inv, s_logc);
  __syncthreads();
  const LnCtx lx{s_inv, s_logc, d.T, d.Tmax, d.igF};
  const int team = TEAM == 256 ? 0 : (int)(threadIdx.x / kWave);
  TeamLds<MAXC, TASK, TEAM, IPT>& L = s_team[team];
  const int t = team_tid<TEAM>();
  unsigned long long my_bytes = 0;
  const int64_t n_items = it.n_groups * it.per_group;
  const int team_global = (int)blockIdx.x * NT + team;
  for (int64_t item = team_global; item < n_items; item += (int64_t)gridDim.x * NT) {
    const int64_t grp = it.node_major ? item % it.n_groups : item / it.per_group;
    const int j = it.j0 + (int)(it.node_major ? item / it.n_groups : item % it.per_group);
    const uint32_t o = it.list ? it.list[grp] : (uint32_t)grp;
    const uint32_t i = lv.open_idx[o];
    const int f = (int)(rand_u32(d.seed, kStreamFeat, (uint32_t)lv.tree_global[lv.tree[i]], lv.id[i], (uint64_t)j) %
                        (uint32_t)d.F);
    const int64_t ridx = (int64_t)o * it.D + j;
    NodeCtx nc;
    const bool pure = node_context<MAXC, TASK, TEAM, IPT>(d, lv, i, nc, L, lx);
    if (pure) {
      // Every boundary of a pure node has gain exactly 0: no positive split.
      if (t == 0) {
        it.res_score[ridx] = 0.0f;
        it.res_thr[ridx] = 0.0f;
        it.res_feat[ridx] = f;
      }
      team_sync<TEAM>();
      continue;
    }
    if (t == 0) my_bytes += (unsigned long long)lv.m[i] * (TASK == 0 ? 12ull : 20ull);
    rf_item<MAXC, TASK, TEAM, IPT, BIG>(d, lv, it, f, i, nc, L, lx, s_boff, (int)blockIdx.x);
    team_sync<TEAM>();
    if (t == 0) rf_write_result<MAXC, TASK, TEAM, IPT>(d, it, ridx, f, L.cr);
    team_sync<TEAM>();
  }
  if (t == 0 && my_bytes && it.bytes) atomicAdd(it.bytes, my_bytes);
}

// Wave-team (TEAM = 64) and workgroup (TEAM = 256) tiers, node-major: one
// team per node walks the node's draws in order.  The node's statistics, row
// ids and payloads are read once; the key gather of the next draw is in
// flight while the current draw is sorted and walked.  Same results as
// k_search_rf<.., TEAM, false> item by item.
template <int MAXC, int TASK, int TEAM>
__global__ void __launch_bounds__(kBlock) k_search_rf_team(DataArgs d, LevelArgs lv, ItemArgs it) {
  constexpr int IPT = TEAM == kWave ? kTeamIpt : kIpt;
  constexpr int NT = kBlock / TEAM;
  resolve_level(lv, d.C);
  resolve_items(it);
  if ((int64_t)blockIdx.x * NT >= it.n_groups) return;  // no node for this workgroup
  // LDS copy of T's head for wave teams (still 2 workgroups per CU); the
  // workgroup teams keep their registers for occupancy instead
  constexpr int kTL = (TASK == 0 && TEAM == kWave) ? 1024 : 0;
  __shared__ double s_inv[128], s_logc[128];
  __shared__ int64_t s_T[kTL > 0 ? kTL : 1];
  __shared__ TeamLds<MAXC, TASK, TEAM, IPT> s_team[NT];
  load_tables(s_inv, s_logc);
  const int64_t nTL = kTL > 0 ? load_t_prefix<kTL>(s_T, d.T, d.Tmax) : 0;
  __syncthreads();
  const auto lx = [&]() {
    if constexpr (kTL > 0)
      return LnCtxL{s_inv, s_logc, d.T, d.Tmax, d.igF, s_T, nTL};
    else
      return LnCtx{s_inv, s_logc, d.T, d.Tmax, d.igF};
  }();
  const int team = TEAM == kWave ? (int)(threadIdx.x / kWave) : 0;
  TeamLds<MAXC, TASK, TEAM, IPT>& L = s_team[team];
  const int t = team_tid<TEAM>();
  unsigned long long my_bytes = 0;
  for (int64_t grp = (int64_t)blockIdx.x * NT + team; grp < it.n_groups; grp += (int64_t)gridDim.x * NT) {
    const uint32_t o = it.list ? it.list[grp] : (uint32_t)grp;
    const uint32_t i = lv.open_idx[o];
    const uint32_t tg = (uint32_t)lv.tree_global[lv.tree[i]];
    const uint32_t nid = lv.id[i];
    NodeCtx nc;
    const bool pure = node_context<MAXC, TASK, TEAM, IPT>(d, lv, i, nc, L, lx);
    if (pure) {
      // Every boundary of a pure node has gain exactly 0: no positive split.
      if (t == 0)
        for (int jj = 0; jj < it.per_group; ++jj) {
          const int j = it.j0 + jj;
          const int64_t ridx = (int64_t)o * it.D + j;
          it.res_score[ridx] = 0.0f;
          it.res_thr[ridx] = 0.0f;
          it.res_feat[ridx] = (int)(rand_u32(d.seed, kStreamFeat, tg, nid, (uint64_t)j) % (uint32_t)d.F);
        }
      team_sync<TEAM>();
      continue;
    }
    const int m = lv.m[i];
    const int64_t st = lv.start[i];
    uint32_t rows[IPT], kn[IPT];
#pragma unroll
    for (int q = 0; q < IPT; ++q) {
      const int e = t + q * TEAM;
      if (e < m) {
        rows[q] = lv.row[st + e];
        L.aux[e] = lv.aux[st + e];
        if (TASK == 1) L.y[e] = lv.ey[st + e];
      }
    }
    int f_nx = (int)(rand_u32(d.seed, kStreamFeat, tg, nid, (uint64_t)it.j0) % (uint32_t)d.F);
#pragma unroll
    for (int q = 0; q < IPT; ++q)
      if (t + q * TEAM < m) kn[q] = d.K[(int64_t)f_nx * d.N + rows[q]];
    for (int jj = 0; jj < it.per_group; ++jj) {
      const int j = it.j0 + jj;
      const int64_t ridx = (int64_t)o * it.D + j;
      const int f = f_nx;
      uint32_t key[IPT];
#pragma unroll
      for (int q = 0; q < IPT; ++q) key[q] = kn[q];
      if (jj + 1 < it.per_group) {
        f_nx = (int)(rand_u32(d.seed, kStreamFeat, tg, nid, (uint64_t)(j + 1)) % (uint32_t)d.F);
#pragma unroll
        for (int q = 0; q < IPT; ++q)
          if (t + q * TEAM < m) kn[q] = d.K[(int64_t)f_nx * d.N + rows[q]];
      }
      if (t == 0) {
        for (int c = 0; c < MAXC; ++c) L.cr.run[c] = 0;
        L.cr.run_n = 0;
        L.cr.run_s = 0;
        L.cr.last_key = 0;
        L.cr.best_g = 0.0;
        L.cr.best_key = kNone;
        L.cr.best_prev = kNone;
      }
      uint32_t kmin = kNone, kmax = 0;
#pragma unroll
      for (int q = 0; q < IPT; ++q) {
        const int e = t + q * TEAM;
        if (e < m) {
          kmin = min(kmin, key[q]);
          kmax = max(kmax, key[q]);
        }
      }
      kmin = team_reduce_min<TEAM>(kmin, L.w);
      kmax = team_reduce_max<TEAM>(kmax, L.w);
#pragma unroll
      for (int q = 0; q < IPT; ++q) {
        const int e = t + q * TEAM;
        if (e < m) {
          L.u.srt.k[0][e] = key[q] - kmin;
          L.u.srt.v[0][e] = (uint16_t)e;
        }
      }
      team_sync<TEAM>();
      sort_and_eval<MAXC, TASK, TEAM, IPT, true>(m, bit_length(kmax - kmin), kmin, nc, L, lx);
      team_sync<TEAM>();
      if (t == 0) rf_write_result<MAXC, TASK, TEAM, IPT>(d, it, ridx, f, L.cr);
      team_sync<TEAM>();
    }
    if (t == 0) my_bytes += (unsigned long long)m * it.per_group * (TASK == 0 ? 12ull : 20ull);
  }
  if (t == 0 && my_bytes && it.bytes) atomicAdd(it.bytes, my_bytes);
}

// ------------------------------------- chunk-parallel big-node search ---
// Items (node, feature draw) above the workgroup's LDS capacity.  The keys of
// an item (dense ranks of the drawn feature) are split into nb coarse buckets
// of consecutive keys, bucket = key >> shift, sized so that a bucket holds
// about kV2Target entries: one wave-sized chunk.  Every bucket of every item
// of a level is then evaluated in parallel:
//   1. k_v2_hist     per (item, tile): bucket counts and max key; the
//                    gathered keys are kept in entry order for the scatter
//   2. k_v2_scan     per item: bucket offsets, last key before each bucket,
//                    task lists
//   3. k_v2_scatter  per (item, tile): entries into bucket order, staged in
//                    LDS so that each bucket's run is written contiguously
//   4. k_v2_ctot     per bucket: class totals (regression: weight and target
//                    sums) of the bucket's entries
//   5. k_v2_carry    per item: totals before each bucket
//   6. k_v2_eval     per bucket, one wave each: LDS sort + scored walk from
//                    the carry
//   7. k_v2_spill    per overfull bucket (ties, clustered keys): fine
//                    sub-buckets walked in order by one workgroup
//   8. k_v2_reduce   per item: best bucket -> (score, threshold)
// Entry order inside a bucket does not matter: the walk sorts by key and
// scores only boundaries between distinct keys from integer prefix sums.
constexpr int kV2Tile = 16384;   // entries per histogram tile
constexpr int kXcds = 8;          // XCDs of an MI355X, one L2 each
constexpr int kScatterTile = 4096;  // entries per scatter tile (classification)
constexpr int kV2MaxB = 4096;    // buckets per item
constexpr int kV2Ipt = kIpt;     // entries per lane of the bucket evaluator
constexpr int kV2Target = kDefBucketTarget;  // entries per bucket aimed at (default)
constexpr int kV2WaveCap = kWave * kV2Ipt;  // bucket capacity of a wave (512)

struct V2Item {
  int32_t f;
  int32_t m;
  int64_t start;  // first entry of the node
  int64_t soff;   // first scratch slot of the item
  int32_t boff;   // first bucket slot of the item
  int32_t nb;     // buckets
  uint32_t width;  // bucket = key / width (v2_bucket)
  uint32_t magic;  // floor(2^32 / width)
  uint32_t node;  // level node index
  int32_t sgrp;   // buckets per super-bucket of the two-pass scatter (0: one pass)
  int64_t ridx;   // result slot (open position * D + draw)
  // node context (as node_context computes it)
  int64_t n_w;    // sum of weights
  int64_t P;      // classification: T(n) - sum_c T(n_c)
  int64_t ssum;   // regression: fixed-point target sum
  double inv;     // 1 / (n ln 2)
};

// key / width from the high word of key * floor(2^32 / width), which is the
// quotient or one below it, and one exact correction.
__device__ __forceinline__ uint32_t v2_bucket(uint32_t key, const V2Item& it) {
  uint32_t q = __umulhi(key, it.magic);
  if ((uint64_t)(q + 1u) * it.width <= key) ++q;
  return q;
}

struct V2Tile {
  uint32_t item;
  uint32_t e0;
};

struct V2Task {
  uint32_t item;
  uint32_t b;
};

struct V2Args {
  const V2Item* items;
  int64_t n_items;
  const V2Tile* tiles;
  // per bucket
  uint32_t* bcount;
  uint32_t* bmax;
  uint32_t* bcls;   // [bucket * C + c] (classification)
  uint32_t* bwn;    // weight sum (regression)
  int64_t* bsum;    // target sum (regression)
  uint32_t* bbase;  // scratch offset within the item
  uint32_t* bcur;
  uint32_t* bsup;   // two-pass scatter: write cursor of each super-bucket [boff + s]
  uint32_t* carry_cls;
  int64_t* carry_n;
  int64_t* carry_s;
  uint32_t* carry_last;
  double* res_g;
  uint32_t* res_key;
  uint32_t* res_prev;
  // scratch, item-major, bucket order
  uint32_t* skey;
  uint32_t* saux;
  int64_t* sy;
  uint32_t* skey2;  // overfull buckets: fine sub-bucket order
  uint32_t* gkey;   // keys in entry order (histogram -> scatter; aliases skey2)
  uint32_t* saux2;
  int64_t* sy2;
  uint32_t* tkey;   // two-pass scatter: keys in super-bucket order (aux, y in saux2, sy2)
  // task lists
  V2Task* tasks;
  V2Task* spill;
  unsigned int* counters;  // [0] tasks, [1] spill
};

// Exclusive max-scan over a team (one wave, or the 256-thread workgroup with
// kWaves words in s_w).
template <int TEAM>
__device__ __forceinline__ uint32_t team_excl_max(uint32_t v, uint32_t* s_w) {
  const int lane = threadIdx.x & (kWave - 1);
  uint32_t inc = v;
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    const uint32_t y = __shfl_up(inc, off, kWave);
    if (lane >= off) inc = max(inc, y);
  }
  uint32_t exc = __shfl_up(inc, 1, kWave);
  if (lane == 0) exc = 0;
  if constexpr (TEAM == kWave) {
    return exc;
  } else {
    const int w = threadIdx.x / kWave;
    if (lane == kWave - 1) s_w[w] = inc;
    __syncthreads();
    uint32_t pre = 0;
#pragma unroll
    for (int i = 0; i < kWaves; ++i) pre = (i < w) ? max(pre, s_w[i]) : pre;
    __syncthreads();
    return max(pre, exc);
  }
}

// Items with at most kV2SmallB buckets take one wave in the per-item passes
// (offsets, carries); larger ones a workgroup.
constexpr int kV2SmallB = 256;

// The item of a per-item team: list[team], teams of TEAM threads.
template <int TEAM>
__device__ __forceinline__ int64_t v2_team_item(const uint32_t* list, int64_t n_list) {
  const int64_t k = TEAM == kWave ? (int64_t)blockIdx.x * kWaves + threadIdx.x / kWave : (int64_t)blockIdx.x;
  return k < n_list ? (int64_t)list[k] : -1;
}

// Tiles are issued feature-major and split in eighths across the XCDs (see
// big_pipeline): a larger grid of smaller bucket tables was slower, since
// more tiles in flight gather from more columns of K at once.
template <int TASK>
__global__ void __launch_bounds__(kBlock) k_v2_hist(DataArgs d, LevelArgs lv, V2Args a) {
  __shared__ uint32_t s_cnt[kV2MaxB];
  __shared__ uint32_t s_max[kV2MaxB];
  const V2Tile tl = a.tiles[blockIdx.x];
  const V2Item it = a.items[tl.item];
  const int t = threadIdx.x;
  for (int b = t; b < it.nb; b += kBlock) {
    s_cnt[b] = 0;
    s_max[b] = 0;
  }
  __syncthreads();
  const uint32_t* Kf = d.K + (int64_t)it.f * d.N;
  const int n = min(kV2Tile, it.m - (int)tl.e0);
  const int64_t e0 = it.start + tl.e0;
  uint32_t* gkey = a.gkey + it.soff + tl.e0;
  constexpr int U = 8;
  for (int base = 0; base < n; base += kBlock * U) {
    uint32_t rows[U];
#pragma unroll
    for (int q = 0; q < U; ++q) {
      const int e = base + t + q * kBlock;
      if (e < n) rows[q] = lv.row[e0 + e];
    }
#pragma unroll
    for (int q = 0; q < U; ++q) {
      const int e = base + t + q * kBlock;
      if (e < n) {
        const uint32_t key = Kf[rows[q]];
        gkey[e] = key;
        const uint32_t b = v2_bucket(key, it);
        atomicAdd(&s_cnt[b], 1u);
        atomicMax(&s_max[b], key);
      }
    }
  }
  __syncthreads();
  for (int b = t; b < it.nb; b += kBlock) {
    const uint32_t c = s_cnt[b];
    if (!c) continue;
    const int64_t gb = (int64_t)it.boff + b;
    atomicAdd(&a.bcount[gb], c);
    atomicMax(&a.bmax[gb], s_max[b]);
  }
}

// One team per item: offsets, last key before each bucket, task lists.
template <int TEAM>
__global__ void __launch_bounds__(kBlock) k_v2_scan(V2Args a, const uint32_t* list, int64_t n_list) {
  __shared__ uint32_t s_w[kWaves];
  __shared__ uint32_t s_base[kWaves][2];
  const int64_t k = v2_team_item<TEAM>(list, n_list);
  if (k < 0) return;  // team-uniform; wave teams have no workgroup barrier
  const V2Item it = a.items[k];
  const int t = team_tid<TEAM>();
  uint32_t* sb = s_base[TEAM == kWave ? threadIdx.x / kWave : 0];
  constexpr int PER = (TEAM == kWave ? kV2SmallB : kV2MaxB) / TEAM;  // buckets per thread
  const int b0 = t * PER;
  const int64_t gb0 = (int64_t)it.boff + b0;
  uint32_t cnt[PER], mx[PER];
  uint32_t lcnt = 0, lmax = 0, n_norm = 0, n_spill = 0;
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const bool ok = b0 + q < it.nb;
    cnt[q] = ok ? a.bcount[gb0 + q] : 0u;
    mx[q] = ok ? a.bmax[gb0 + q] : 0u;
    lcnt += cnt[q];
    if (cnt[q]) {
      lmax = max(lmax, mx[q]);
      if (cnt[q] > (uint32_t)kV2WaveCap) ++n_spill;
      else ++n_norm;
    }
  }
  uint32_t off = team_excl_scan<TEAM>(lcnt, s_w, nullptr);
  uint32_t last = team_excl_max<TEAM>(lmax, s_w);
  uint32_t tot_norm = 0, tot_spill = 0;
  const uint32_t pn = team_excl_scan<TEAM>(n_norm, s_w, &tot_norm);
  const uint32_t ps = team_excl_scan<TEAM>(n_spill, s_w, &tot_spill);
  if (t == 0) {
    sb[0] = tot_norm ? atomicAdd(&a.counters[0], tot_norm) : 0u;
    sb[1] = tot_spill ? atomicAdd(&a.counters[1], tot_spill) : 0u;
  }
  team_sync<TEAM>();
  uint32_t in = sb[0] + pn, is = sb[1] + ps;
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    if (b0 + q >= it.nb) break;
    const int64_t gb = gb0 + q;
    a.bbase[gb] = off;
    a.bcur[gb] = off;
    if (it.sgrp > 0 && (b0 + q) % it.sgrp == 0) a.bsup[(int64_t)it.boff + (b0 + q) / it.sgrp] = off;
    a.carry_last[gb] = last;
    off += cnt[q];
    if (!cnt[q]) continue;
    last = max(last, mx[q]);
    const V2Task tk{(uint32_t)k, (uint32_t)(b0 + q)};
    if (cnt[q] > (uint32_t)kV2WaveCap)
      a.spill[is++] = tk;
    else
      a.tasks[in++] = tk;
  }
}

// Entries of one tile into bucket order.  Each entry takes a rank inside its
// bucket from an LDS counter, the tile reserves one run per bucket in the
// item's scratch, and the entries are staged in LDS in (bucket, rank) order
// so that consecutive threads write consecutive slots of a run.
// Items of at most kV2FewB (kV2MidB) buckets take the scatter variants with
// small bucket tables (less LDS: three workgroups per CU instead of two).
//
// Items with more than kV2MidB buckets hold about one entry per bucket in a
// tile, or a fraction of one, so a one-pass scatter writes runs of about one
// entry: every write touches its own cache line.  They go in two passes
// over the same tiles (MODE 1, then MODE 2): first into kV2Super
// super-buckets of sgrp consecutive buckets each (runs of tens of entries,
// into tkey / saux2 / sy2), then each tile of that super-bucket order into
// its buckets -- a tile now spans one or two super-buckets, so its runs are
// long again.
// Entry order inside a bucket is not used by the evaluation (it sorts).
constexpr int kV2FewB = 512;
constexpr int kV2MidB = 1024;
constexpr int kV2Super = 64;

template <int TASK, int MAXB, int MODE = 0>
__global__ void __launch_bounds__(kBlock) k_v2_scatter(DataArgs d, LevelArgs lv, V2Args a) {
  constexpr int TILE = TASK == 0 ? kScatterTile : 2048;
  constexpr int U = TILE / kBlock;
  __shared__ uint32_t s_cnt[MAXB];
  __shared__ uint32_t s_gpos[MAXB];
  __shared__ uint32_t s_key[TILE];
  __shared__ uint32_t s_aux[TILE];
  __shared__ uint16_t s_b[TILE];
  __shared__ int64_t s_y[TASK == 1 ? TILE : 1];
  __shared__ uint32_t s_w[kWaves];
  const V2Tile tl = a.tiles[blockIdx.x];
  const V2Item it = a.items[tl.item];
  const int t = threadIdx.x;
  const int n = min(TILE, it.m - (int)tl.e0);
  // input: keys in entry order (MODE 0 / 1) or in super-bucket order (MODE 2)
  const uint32_t* gkey = (MODE == 2 ? a.tkey : a.gkey) + it.soff + tl.e0;
  const uint32_t* gaux = MODE == 2 ? a.saux2 + it.soff + tl.e0 : lv.aux + it.start + tl.e0;
  const int64_t* gy = MODE == 2 ? a.sy2 + it.soff + tl.e0 : lv.ey + it.start + tl.e0;
  // table slots: bucket (MODE 0), super-bucket (MODE 1), bucket - b_lo (MODE 2)
  int b_lo = 0, n_slots = it.nb;
  if (MODE == 1) n_slots = (it.nb + it.sgrp - 1) / it.sgrp;
  if (MODE == 2) {
    const int s_first = (int)v2_bucket(gkey[0], it) / it.sgrp;
    const int s_last = (int)v2_bucket(gkey[n - 1], it) / it.sgrp;
    b_lo = s_first * it.sgrp;
    n_slots = min(it.nb, (s_last + 1) * it.sgrp) - b_lo;
  }
  auto slot_of = [&](uint32_t key) -> int {
    const int b = (int)v2_bucket(key, it);
    return MODE == 1 ? b / it.sgrp : b - b_lo;
  };
  for (int b = t; b < n_slots; b += kBlock) s_cnt[b] = 0;
  __syncthreads();
  uint32_t keys[U], auxv[U], rank[U];
  int64_t yv[U];
#pragma unroll
  for (int q = 0; q < U; ++q) {
    const int e = t + q * kBlock;
    if (e < n) {
      keys[q] = gkey[e];
      auxv[q] = gaux[e];
      if (TASK == 1) yv[q] = gy[e];
    }
  }
#pragma unroll
  for (int q = 0; q < U; ++q) {
    const int e = t + q * kBlock;
    if (e < n) rank[q] = atomicAdd(&s_cnt[slot_of(keys[q])], 1u);
  }
  __syncthreads();
  {
    // slots in chunks of PER per thread (several rounds for a MODE 2 table)
    constexpr int PER = MAXB / kBlock > 0 ? MAXB / kBlock : 1;
    uint32_t* cur = MODE == 1 ? a.bsup : a.bcur;
    const int64_t cbase = (int64_t)it.boff + (MODE == 2 ? b_lo : 0);
    const int b0 = t * PER;
    uint32_t c[PER], loc = 0;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      c[q] = b0 + q < n_slots ? s_cnt[b0 + q] : 0u;
      loc += c[q];
    }
    uint32_t off = team_excl_scan<256>(loc, s_w, nullptr);
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      if (b0 + q >= n_slots) break;
      s_cnt[b0 + q] = off;  // local offset of the slot's run
      s_gpos[b0 + q] = c[q] ? atomicAdd(&cur[cbase + b0 + q], c[q]) : 0u;
      off += c[q];
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < U; ++q) {
    const int e = t + q * kBlock;
    if (e < n) {
      const int b = slot_of(keys[q]);
      const uint32_t li = s_cnt[b] + rank[q];
      s_key[li] = keys[q];
      s_aux[li] = auxv[q];
      s_b[li] = (uint16_t)b;
      if (TASK == 1) s_y[li] = yv[q];
    }
  }
  __syncthreads();
  uint32_t* okey = MODE == 1 ? a.tkey : a.skey;
  uint32_t* oaux = MODE == 1 ? a.saux2 : a.saux;
  int64_t* oy = MODE == 1 ? a.sy2 : a.sy;
  for (int i = t; i < n; i += kBlock) {
    const uint32_t b = s_b[i];
    const int64_t pos = it.soff + s_gpos[b] + (uint32_t)i - s_cnt[b];
    okey[pos] = s_key[i];
    oaux[pos] = s_aux[i];
    if (TASK == 1) oy[pos] = s_y[i];
  }
}

// One workgroup per tile of kV2CTile buckets of an item, one wave per
// bucket: class totals (regression: weight and target sums) of each
// bucket's entries.
constexpr int kV2CTile = 64;

template <int MAXC, int TASK>
__global__ void __launch_bounds__(kBlock) k_v2_ctot(DataArgs d, V2Args a) {
  __shared__ uint32_t s_c[kWaves][MAXC];
  const int lane = threadIdx.x & (kWave - 1);
  const V2Tile tl = a.tiles[blockIdx.x];
  const V2Item it = a.items[tl.item];
  const int b_end = min(it.nb, (int)tl.e0 + kV2CTile);
  for (int b = (int)tl.e0 + threadIdx.x / kWave; b < b_end; b += kWaves) {
    const int64_t gb = (int64_t)it.boff + b;
    const uint32_t n = a.bcount[gb];
    const int64_t e0 = it.soff + a.bbase[gb];
    if constexpr (TASK == 0) {
      uint32_t* sc = s_c[threadIdx.x / kWave];
      if (lane < MAXC) sc[lane] = 0u;
      team_sync<kWave>();
      constexpr int U = 4;
      for (uint32_t e = lane; e < n; e += kWave * U) {
        uint32_t x[U];
#pragma unroll
        for (int q = 0; q < U; ++q) x[q] = e + q * kWave < n ? a.saux[e0 + e + q * kWave] : 0u;
#pragma unroll
        for (int q = 0; q < U; ++q)
          if (e + q * kWave < n) atomicAdd(&sc[x[q] >> 16], x[q] & 0xFFFFu);
      }
      team_sync<kWave>();
      if (lane < d.C) a.bcls[gb * d.C + lane] = sc[lane];
      team_sync<kWave>();
    } else {
      int64_t wn = 0, ws = 0;
      for (uint32_t e = lane; e < n; e += kWave) {
        const int64_t w = (int64_t)(a.saux[e0 + e] & 0xFFFFu);
        wn += w;
        ws += w * a.sy[e0 + e];
      }
#pragma unroll
      for (int off = kWave / 2; off > 0; off >>= 1) {
        wn += __shfl_xor(wn, off, kWave);
        ws += __shfl_xor(ws, off, kWave);
      }
      if (lane == 0) {
        a.bwn[gb] = (uint32_t)wn;
        a.bsum[gb] = ws;
      }
    }
  }
}

// One team per item: totals before each bucket.
template <int TASK, int TEAM>
__global__ void __launch_bounds__(kBlock) k_v2_carry(DataArgs d, V2Args a, const uint32_t* list, int64_t n_list) {
  __shared__ uint32_t s_w[8 * kWaves + kWaves];
  __shared__ int64_t s_w64[kWaves];
  const int64_t ki = v2_team_item<TEAM>(list, n_list);
  if (ki < 0) return;  // team-uniform; wave teams have no workgroup barrier
  const V2Item it = a.items[ki];
  const int t = team_tid<TEAM>();
  const int C = d.C;
  constexpr int PER = (TEAM == kWave ? kV2SmallB : kV2MaxB) / TEAM;
  const int b0 = t * PER;
  const int64_t gb0 = (int64_t)it.boff + b0;
  if constexpr (TASK == 0) {
    int64_t lw = 0;
    for (int q = 0; q < PER && b0 + q < it.nb; ++q)
      for (int c = 0; c < C; ++c) lw += a.bcls[(gb0 + q) * C + c];
    int64_t run_n = team_excl_scan_i64<TEAM>(lw, s_w64, nullptr);
    for (int q = 0; q < PER && b0 + q < it.nb; ++q) {
      a.carry_n[gb0 + q] = run_n;
      for (int c = 0; c < C; ++c) run_n += a.bcls[(gb0 + q) * C + c];
    }
    for (int c0 = 0; c0 < C; c0 += 8) {
      uint32_t v[8], tot[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        v[k] = 0;
        if (c0 + k < C)
          for (int q = 0; q < PER && b0 + q < it.nb; ++q) v[k] += a.bcls[(gb0 + q) * C + c0 + k];
      }
      team_excl_scan_vec<TEAM, 8>(v, s_w, tot);
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        if (c0 + k >= C) continue;
        uint32_t r = v[k];
        for (int q = 0; q < PER && b0 + q < it.nb; ++q) {
          const int64_t gb = gb0 + q;
          a.carry_cls[gb * C + c0 + k] = r;
          r += a.bcls[gb * C + c0 + k];
        }
      }
    }
  } else {
    int64_t lw = 0, ls = 0;
    for (int q = 0; q < PER && b0 + q < it.nb; ++q) {
      lw += a.bwn[gb0 + q];
      ls += a.bsum[gb0 + q];
    }
    int64_t run_n = team_excl_scan_i64<TEAM>(lw, s_w64, nullptr);
    int64_t run_s = team_excl_scan_i64<TEAM>(ls, s_w64, nullptr);
    for (int q = 0; q < PER && b0 + q < it.nb; ++q) {
      const int64_t gb = gb0 + q;
      a.carry_n[gb] = run_n;
      a.carry_s[gb] = run_s;
      run_n += a.bwn[gb];
      run_s += a.bsum[gb];
    }
  }
}

template <int MAXC, int TASK, int IPT>
__device__ __forceinline__ void v2_carry(const DataArgs& d, const V2Args& a, int64_t gb,
                                         TeamLds<MAXC, TASK, 256, IPT>& L) {
  if (threadIdx.x == 0) {
    for (int c = 0; c < MAXC; ++c) L.cr.run[c] = (TASK == 0 && c < d.C) ? a.carry_cls[gb * d.C + c] : 0u;
    L.cr.run_n = a.carry_n[gb];
    L.cr.run_s = TASK == 1 ? a.carry_s[gb] : 0;
    L.cr.last_key = a.carry_last[gb];
    L.cr.best_g = 0.0;
    L.cr.best_key = kNone;
    L.cr.best_prev = kNone;
  }
}

// Node context of an item from its precomputed fields; class totals to LDS
// (visible after the caller's next barrier).
template <int MAXC, int TASK, int IPT>
__device__ __forceinline__ NodeCtx v2_node(const DataArgs& d, const LevelArgs& lv, const V2Item& it,
                                           TeamLds<MAXC, TASK, 256, IPT>& L) {
  NodeCtx nc;
  nc.C = d.C;
  nc.n = it.n_w;
  nc.P = it.P;
  nc.ssum = it.ssum;
  nc.inv = it.inv;
  if (TASK == 0 && (int)threadIdx.x < d.C) L.ncls[threadIdx.x] = lv.cls[(int64_t)it.node * d.C + threadIdx.x];
  return nc;
}

// One wave per bucket of at most kV2WaveCap entries (four independent
// buckets per workgroup), grid-stride over the task list.  The loads of a
// wave's next bucket (task, item and bucket fields, carries, entries) are
// issued before its current bucket is sorted and walked.  Classification:
// k_v2_eval, three workgroups per CU (up to 16 classes: about 53 KB of LDS
// each, at most 168 VGPRs), bound by the latency of its x ln x reads;
// regression: k_v2_eval_reg.
template <int TASK, int IPT>
struct V2Pre {
  int64_t gb, e0;
  uint32_t node, kbase, bmax;
  int n;
  int64_t n_w, P, ssum;
  double inv;
  uint32_t carry_c;  // class carry of class `lane` (lanes < C)
  int64_t carry_n, carry_s;
  uint32_t carry_last;
  uint32_t ncls;     // node class total of class `lane`
  uint32_t key[IPT], aux[IPT];
  int64_t y[TASK == 1 ? IPT : 1];
};

template <int TASK, int IPT>
__device__ __forceinline__ void v2_prefetch(const DataArgs& d, const LevelArgs& lv, const V2Args& a, int64_t q,
                                            V2Pre<TASK, IPT>& P) {
  const int lane = threadIdx.x & (kWave - 1);
  const V2Task tk = a.tasks[q];
  const V2Item& it = a.items[tk.item];
  P.gb = (int64_t)it.boff + tk.b;
  P.kbase = tk.b * it.width;
  P.node = it.node;
  P.n_w = it.n_w;
  P.P = it.P;
  P.ssum = it.ssum;
  P.inv = it.inv;
  P.n = (int)a.bcount[P.gb];
  P.e0 = it.soff + a.bbase[P.gb];
  P.bmax = a.bmax[P.gb];
  P.carry_n = a.carry_n[P.gb];
  P.carry_s = TASK == 1 ? a.carry_s[P.gb] : 0;
  P.carry_last = a.carry_last[P.gb];
  P.carry_c = (TASK == 0 && lane < d.C) ? a.carry_cls[P.gb * d.C + lane] : 0u;
  P.ncls = (TASK == 0 && lane < d.C) ? lv.cls[(int64_t)P.node * d.C + lane] : 0u;
#pragma unroll
  for (int k = 0; k < IPT; ++k) {
    const int e = lane + k * kWave;
    if (e < P.n) {
      P.key[k] = a.skey[P.e0 + e];
      P.aux[k] = a.saux[P.e0 + e];
      if (TASK == 1) P.y[k] = a.sy[P.e0 + e];
    }
  }
}

template <int MAXC, int TASK>
__global__ void __launch_bounds__(kBlock, MAXC <= 16 ? 3 : 1) k_v2_eval(DataArgs d, LevelArgs lv, V2Args a, int64_t n_tasks) {
  static_assert(TASK == 0, "regression buckets go through k_v2_eval_reg");
  constexpr int IPT = kV2Ipt;
  __shared__ double s_inv[128], s_logc[128];
  __shared__ TeamLds<MAXC, TASK, kWave, IPT> s_team[kWaves];
  load_tables(s_inv, s_logc);
  __syncthreads();
  const LnCtx lx{s_inv, s_logc, d.T, d.Tmax, d.igF};
  const int t = threadIdx.x & (kWave - 1);
  TeamLds<MAXC, TASK, kWave, IPT>& L = s_team[threadIdx.x / kWave];
  const int64_t stride = (int64_t)gridDim.x * kWaves;
  int64_t q = (int64_t)blockIdx.x * kWaves + threadIdx.x / kWave;
  if (q >= n_tasks) return;  // wave-uniform; no workgroup barrier follows
  V2Pre<TASK, IPT> cur;
  v2_prefetch<TASK, IPT>(d, lv, a, q, cur);
  while (q < n_tasks) {
    NodeCtx nc;
    nc.C = d.C;
    nc.n = cur.n_w;
    nc.P = cur.P;
    nc.ssum = cur.ssum;
    nc.inv = cur.inv;
    if (t < d.C) {
      L.ncls[t] = cur.ncls;
      L.cr.run[t] = cur.carry_c;
    }
    if (t == 0) {
      for (int c = d.C; c < MAXC; ++c) L.cr.run[c] = 0u;
      L.cr.run_n = cur.carry_n;
      L.cr.run_s = cur.carry_s;
      L.cr.last_key = cur.carry_last;
      L.cr.best_g = 0.0;
      L.cr.best_key = kNone;
      L.cr.best_prev = kNone;
    }
    const int n = cur.n;
#pragma unroll
    for (int k = 0; k < IPT; ++k) {
      const int e = t + k * kWave;
      if (e < n) {
        L.u.srt.k[0][e] = cur.key[k] - cur.kbase;
        L.u.srt.v[0][e] = (uint16_t)e;
        L.aux[e] = cur.aux[k];
      }
    }
    const int64_t gb = cur.gb;
    const uint32_t kbase = cur.kbase;
    const int bits = bit_length(cur.bmax - kbase);
    team_sync<kWave>();
    const int64_t qn = q + stride;
    if (qn < n_tasks) v2_prefetch<TASK, IPT>(d, lv, a, qn, cur);
    sort_and_eval<MAXC, TASK, kWave, IPT>(n, bits, kbase, nc, L, lx);
    team_sync<kWave>();
    if (t == 0) {
      a.res_g[gb] = L.cr.best_g;
      a.res_key[gb] = L.cr.best_key;
      a.res_prev[gb] = L.cr.best_prev;
    }
    team_sync<kWave>();
    q = qn;
  }
}

// LDS of one wave of the regression bucket evaluator.  The payloads
// (weights, fixed-point targets) stay in registers through the sort and then
// go into the sort counters (targets) and the key buffer the sort left free
// (weights): about 10 KB per wave, three workgroups per CU.
template <int IPT>
struct RegWaveLds {
  static constexpr int CAP = kWave * IPT;
  static constexpr int SPAN = TeamLds<1, 1, kWave, IPT>::SPAN;
  uint32_t k[2][CAP];
  uint16_t v[2][CAP];
  union {
    uint16_t cnt[SPAN];
    int64_t y[CAP];
  } u;
  uint32_t w[8 * kWaves + kWaves];
  Carry<1, 1> cr;
  static_assert(sizeof(uint16_t) * SPAN >= sizeof(int64_t) * CAP, "targets fit the sort counters");
};

// Regression walk of a key-sorted bucket on one wave (the regression branch
// of eval_sorted_chunk with the payloads at wts / ys).
template <int IPT>
__device__ void eval_sorted_reg_wave(const uint32_t* keys, const uint16_t* vals, const uint32_t* wts,
                                     const int64_t* ys, int n, uint32_t kbase, const NodeCtx& nc, Carry<1, 1>& cr) {
  const int t = threadIdx.x & (kWave - 1);
  const int ipt = lane_span<kWave>(n);
  const int lo = t * ipt;
  const int hi = min(lo + ipt, n);
  BestCand tb{0.0, kNone, kNone};
  int64_t ln = 0, ls = 0;
  for (int p = lo; p < hi; ++p) {
    const uint16_t v = vals[p];
    const int64_t w = (int64_t)wts[v];
    ln += w;
    ls += w * ys[v];
  }
  int64_t tot_n, tot_s;
  int64_t pn = team_excl_scan_i64<kWave>(ln, nullptr, &tot_n);
  int64_t ps = team_excl_scan_i64<kWave>(ls, nullptr, &tot_s);
  pn += cr.run_n;
  ps += cr.run_s;
  if (lo < hi) {
    uint32_t prevk = (lo > 0) ? keys[lo - 1] + kbase : cr.last_key;
    for (int p = lo; p < hi; ++p) {
      const uint32_t k = keys[p] + kbase;
      if (pn > 0 && prevk < k) {
        const double g = variance_gain(pn, ps, nc.n - pn, nc.ssum - ps, 0);
        if (g > tb.g) tb = BestCand{g, k, prevk};
      }
      const uint16_t v = vals[p];
      const int64_t w = (int64_t)wts[v];
      pn += w;
      ps += w * ys[v];
      prevk = k;
    }
  }
  const BestCand b = wave_best(tb);
  if (t == 0) {
    carry_best(cr, b);
    cr.run_n += tot_n;
    cr.run_s += tot_s;
    if (n > 0) cr.last_key = keys[n - 1] + kbase;
  }
  team_sync<kWave>();
}

// k_v2_eval for regression: one wave per bucket, the same task order, sorts
// and walk, with the payloads written to LDS only once the bucket is sorted.
template <int MAXC, int TASK>
__global__ void __launch_bounds__(kBlock, 3) k_v2_eval_reg(DataArgs d, LevelArgs lv, V2Args a, int64_t n_tasks) {
  static_assert(TASK == 1, "regression only");
  constexpr int IPT = kV2Ipt;
  constexpr int kSpan = RegWaveLds<IPT>::SPAN;
  __shared__ RegWaveLds<IPT> s_wave[kWaves];
  const int t = threadIdx.x & (kWave - 1);
  RegWaveLds<IPT>& L = s_wave[threadIdx.x / kWave];
  const int64_t stride = (int64_t)gridDim.x * kWaves;
  int64_t q = (int64_t)blockIdx.x * kWaves + threadIdx.x / kWave;
  if (q >= n_tasks) return;  // wave-uniform; no workgroup barrier follows
  V2Pre<TASK, IPT> cur;
  v2_prefetch<TASK, IPT>(d, lv, a, q, cur);
  while (q < n_tasks) {
    NodeCtx nc;
    nc.C = d.C;
    nc.n = cur.n_w;
    nc.P = cur.P;
    nc.ssum = cur.ssum;
    nc.inv = cur.inv;
    if (t == 0) {
      L.cr.run[0] = 0u;
      L.cr.run_n = cur.carry_n;
      L.cr.run_s = cur.carry_s;
      L.cr.last_key = cur.carry_last;
      L.cr.best_g = 0.0;
      L.cr.best_key = kNone;
      L.cr.best_prev = kNone;
    }
    const int n = cur.n;
    uint32_t pw[IPT];
    int64_t py[IPT];
#pragma unroll
    for (int k = 0; k < IPT; ++k) {
      const int e = t + k * kWave;
      if (e < n) {
        L.k[0][e] = cur.key[k] - cur.kbase;
        L.v[0][e] = (uint16_t)e;
        pw[k] = cur.aux[k] & 0xFFFFu;
        py[k] = cur.y[k];
      }
    }
    const int64_t gb = cur.gb;
    const uint32_t kbase = cur.kbase;
    const int bits = bit_length(cur.bmax - kbase);
    team_sync<kWave>();
    const int64_t qn = q + stride;
    if (qn < n_tasks) v2_prefetch<TASK, IPT>(d, lv, a, qn, cur);
    int sorted = 0;  // buffer holding the sorted bucket
    if (bits > 4 && (1 << bits) <= kSpan) {
      lds_count_sort<IPT, kWave>(L.k[0], L.v[0], L.k[1], L.v[1], n, 1 << bits, reinterpret_cast<uint32_t*>(L.u.cnt),
                                 L.w);
      sorted = 1;
    } else {
      for (int sh = 0; sh < bits; sh += 4) {
        lds_sort_pass<IPT, kWave>(L.k[sorted], L.v[sorted], L.k[sorted ^ 1], L.v[sorted ^ 1], n, sh, L.u.cnt, L.w);
        sorted ^= 1;
      }
    }
    uint32_t* wts = L.k[sorted ^ 1];
#pragma unroll
    for (int k = 0; k < IPT; ++k) {
      const int e = t + k * kWave;
      if (e < n) {
        wts[e] = pw[k];
        L.u.y[e] = py[k];
      }
    }
    team_sync<kWave>();
    eval_sorted_reg_wave<IPT>(L.k[sorted], L.v[sorted], wts, L.u.y, n, kbase, nc, L.cr);
    if (t == 0) {
      a.res_g[gb] = L.cr.best_g;
      a.res_key[gb] = L.cr.best_key;
      a.res_prev[gb] = L.cr.best_prev;
    }
    team_sync<kWave>();
    q = qn;
  }
}

// One workgroup per overfull bucket: fine sub-buckets of the bucket's key
// range, packed into LDS-sized chunks or read through an LDS histogram
// (a single sub-bucket above the capacity: heavy ties), walked in order.
template <int MAXC, int TASK>
__global__ void __launch_bounds__(kBlock) k_v2_spill(DataArgs d, LevelArgs lv, V2Args a, int64_t n_spill) {
  constexpr int IPT = kIpt;
  constexpr int CAP = kCap;
  __shared__ double s_inv[128], s_logc[128];
  __shared__ TeamLds<MAXC, TASK, 256, IPT> L;
  __shared__ uint32_t s_boff[kMaxBuckets + 1];
  load_tables(s_inv, s_logc);
  __syncthreads();
  const LnCtx lx{s_inv, s_logc, d.T, d.Tmax, d.igF};
  const int t = threadIdx.x;
  for (int64_t q = blockIdx.x; q < n_spill; q += gridDim.x) {
    const V2Task tk = a.spill[q];
    const V2Item it = a.items[tk.item];
    const int64_t gb = (int64_t)it.boff + tk.b;
    const int n = (int)a.bcount[gb];
    const int64_t e0 = it.soff + a.bbase[gb];
    const uint32_t klo = tk.b * it.width;
    const uint32_t span = a.bmax[gb] - klo + 1u;
    const NodeCtx nc = v2_node(d, lv, it, L);
    v2_carry<MAXC, TASK, IPT>(d, a, gb, L);
    __syncthreads();
    if (n <= CAP) {
      // within the workgroup's LDS chunk: one sort and walk
      for (int e = t; e < n; e += 256) {
        L.u.srt.k[0][e] = a.skey[e0 + e] - klo;
        L.u.srt.v[0][e] = (uint16_t)e;
        L.aux[e] = a.saux[e0 + e];
        if (TASK == 1) L.y[e] = a.sy[e0 + e];
      }
      __syncthreads();
      sort_and_eval<MAXC, TASK, 256, IPT>(n, bit_length(span - 1u), klo, nc, L, lx);
      __syncthreads();
      if (t == 0) {
        a.res_g[gb] = L.cr.best_g;
        a.res_key[gb] = L.cr.best_key;
        a.res_prev[gb] = L.cr.best_prev;
      }
      __syncthreads();
      continue;
    }
    const int words = (TASK == 0) ? d.C : 3;
    int wbits = 0;
    while ((1 << (wbits + 1)) * words <= kHistWords) ++wbits;
    while (((span + (1u << wbits) - 1) >> wbits) > (uint32_t)kMaxBuckets) ++wbits;
    const int W = 1 << wbits;
    const int NB = (int)((span + (uint32_t)W - 1) >> wbits);
    const uint32_t* skey = a.skey + e0;
    const uint32_t* saux = a.saux + e0;
    const int64_t* sy = TASK == 1 ? a.sy + e0 : nullptr;
    uint32_t* dkey = a.skey2 + e0;
    uint32_t* daux = a.saux2 + e0;
    int64_t* dy = TASK == 1 ? a.sy2 + e0 : nullptr;
    for (int b = t; b <= NB; b += 256) s_boff[b] = 0;
    __syncthreads();
    for (int e = t; e < n; e += 256) atomicAdd(&s_boff[(skey[e] - klo) >> wbits], 1u);
    __syncthreads();
    {
      const int per = (NB + 255) / 256;
      const int b0 = t * per, b1 = min(b0 + per, NB);
      uint32_t loc = 0;
      for (int b = b0; b < b1; ++b) loc += s_boff[b];
      uint32_t base = team_excl_scan<256>(loc, L.w, nullptr);
      for (int b = b0; b < b1; ++b) {
        const uint32_t c = s_boff[b];
        s_boff[b] = base;
        base += c;
      }
      if (t == 0) s_boff[NB] = (uint32_t)n;
      __syncthreads();
      for (int b = t; b < NB; b += 256) L.u.bpos[b] = s_boff[b];
      __syncthreads();
    }
    for (int e = t; e < n; e += 256) {
      const uint32_t k = skey[e];
      const uint32_t pos = atomicAdd(&L.u.bpos[(k - klo) >> wbits], 1u);
      dkey[pos] = k;
      daux[pos] = saux[e];
      if (TASK == 1) dy[pos] = sy[e];
    }
    __syncthreads();
    int b = 0;
    while (b < NB) {
      const uint32_t cb = s_boff[b + 1] - s_boff[b];
      if (cb == 0) {
        ++b;
        continue;
      }
      if (cb > (uint32_t)CAP) {
        eval_hist_bucket<MAXC, TASK, IPT>(dkey + s_boff[b], daux + s_boff[b], TASK == 1 ? dy + s_boff[b] : nullptr,
                                          cb, klo + ((uint32_t)b << wbits), W, nc, L, lx);
        ++b;
        continue;
      }
      int be = b;
      uint32_t tot = 0;
      while (be < NB) {
        const uint32_t c = s_boff[be + 1] - s_boff[be];
        if (c > (uint32_t)CAP || tot + c > (uint32_t)CAP) break;
        tot += c;
        ++be;
      }
      const uint32_t c0 = s_boff[b];
      const uint32_t kbase = klo + ((uint32_t)b << wbits);
      for (int e = t; e < (int)tot; e += 256) {
        L.u.srt.k[0][e] = dkey[c0 + e] - kbase;
        L.u.srt.v[0][e] = (uint16_t)e;
        L.aux[e] = daux[c0 + e];
        if (TASK == 1) L.y[e] = dy[c0 + e];
      }
      __syncthreads();
      sort_and_eval<MAXC, TASK, 256, IPT>((int)tot, bit_length((uint32_t)(be - b) * (uint32_t)W - 1u), kbase, nc, L,
                                          lx);
      b = be;
    }
    __syncthreads();
    if (t == 0) {
      a.res_g[gb] = L.cr.best_g;
      a.res_key[gb] = L.cr.best_key;
      a.res_prev[gb] = L.cr.best_prev;
    }
    __syncthreads();
  }
}

// One wave per item: the best bucket (ties to the lower key).
template <int TASK>
__global__ void __launch_bounds__(kBlock) k_v2_reduce(DataArgs d, V2Args a, ItemArgs it) {
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t k = (int64_t)blockIdx.x * kWaves + threadIdx.x / kWave;
  if (k >= a.n_items) return;
  const V2Item bi = a.items[k];
  BestCand b{0.0, kNone, kNone};
  for (int q = lane; q < bi.nb; q += kWave) {
    const int64_t gb = (int64_t)bi.boff + q;
    if (!a.bcount[gb]) continue;
    const uint32_t key = a.res_key[gb];
    if (key != kNone && cand_better(a.res_g[gb], key, b.g, b.key)) b = BestCand{a.res_g[gb], key, a.res_prev[gb]};
  }
  b = wave_best(b);
  if (lane == 0) {
    float score = 0.0f, thr = 0.0f;
    if (b.key != kNone) {
      const float* sv = d.SV + d.SVoff[bi.f];
      score = (float)(TASK == 0 ? b.g : ldexp_d(b.g, -2 * d.shift));
      thr = midpoint(sv[b.prev], sv[b.key]);
    }
    it.res_score[bi.ridx] = score;
    it.res_thr[bi.ridx] = thr;
    it.res_feat[bi.ridx] = bi.f;
  }
}

// Item table of a phase: one item per (big group, draw).
__global__ void k_v2_items(DataArgs d, LevelArgs lv, const uint32_t* groups, int64_t n_groups, int per_group, int j0,
                           int D, int nb_cap, int target, V2Item* items) {
  const int64_t n = n_groups * per_group;
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t o = groups[k / per_group];
    const int j = j0 + (int)(k % per_group);
    const uint32_t i = lv.open_idx[o];
    V2Item b;
    b.f = (int)(rand_u32(d.seed, kStreamFeat, (uint32_t)lv.tree_global[lv.tree[i]], lv.id[i], (uint64_t)j) %
                (uint32_t)d.F);
    b.m = lv.m[i];
    b.start = lv.start[i];
    // Buckets of `width` consecutive keys of the feature's dense range
    // [0, S), so that a bucket holds about `target` of the node's m entries
    // (m * width / S), and at most nb_cap buckets.
    const uint32_t S = max(d.S[b.f], 1u);
    const uint64_t w_target = (uint64_t)S * (uint64_t)target / (uint64_t)max(b.m, 1);
    const uint32_t w_min = (uint32_t)(((uint64_t)S + nb_cap - 1) / nb_cap);
    b.width = (uint32_t)max<uint64_t>(max<uint64_t>(w_target, (uint64_t)w_min), 1);
    b.magic = b.width == 1 ? 0xFFFFFFFFu : (uint32_t)(0x100000000ull / b.width);
    b.nb = (int)((S + b.width - 1) / b.width);
    b.soff = 0;
    b.boff = 0;
    b.node = i;
    b.sgrp = 0;
    b.ridx = (int64_t)o * D + j;
    b.n_w = lv.count[i];
    b.P = 0;
    b.ssum = 0;
    b.inv = 0.0;
    if (d.C > 0) {
      const LnCtx lx{g_ln_table.inv, g_ln_table.logc, d.T, d.Tmax, d.igF};
      int64_t tsum = 0;
      for (int c = 0; c < d.C; ++c) tsum += lx.t(lv.cls[(int64_t)i * d.C + c]);
      b.P = lx.t((uint64_t)b.n_w) - tsum;
      b.inv = 1.0 / ((double)b.n_w * kLn2);
    } else {
      b.ssum = lv.sum[i];
    }
    items[k] = b;
  }
}

// ------------------------------------------------ ERT split search ------
// One block per item: threshold = fp32 mean of 10 node values drawn with
// replacement (gpuert.cu:167-183, summed in draw order), then a 2 x C
// histogram (gpuert.cu:287-301) or per-side target sums (gpuert.cu:348-361).
template <int MAXC, int TASK>
__device__ __forceinline__ void ert_block_body(DataArgs d, LevelArgs lv, ItemArgs it, int64_t bid, int64_t nblk,
                                               double* s_inv, double* s_logc) {
  resolve_level(lv, d.C);
  resolve_items(it);
  if ((int64_t)bid >= it.n_groups * it.per_group) return;
  __shared__ uint32_t s_left[MAXC];
  __shared__ unsigned long long s_nl;
  __shared__ unsigned long long s_sl;
  __shared__ float s_thr;
  load_ln_table(s_inv, s_logc);
  __syncthreads();
  const int t = threadIdx.x;
  unsigned long long my_bytes = 0;
  const int64_t n_items = it.n_groups * it.per_group;
  for (int64_t item = bid; item < n_items; item += nblk) {
    const int64_t grp = it.node_major ? item % it.n_groups : item / it.per_group;
    const int q = (int)(it.node_major ? item / it.n_groups : item % it.per_group);
    const uint32_t o = it.list ? it.list[grp] : (uint32_t)grp;
    const uint32_t i = lv.open_idx[o];
    const int tr = lv.tree[i];
    const DataArgs dd = forest_data(d, tr);
    const int j = item_draw(d, dd, it, q);
    if (j < 0) continue;  // past this forest's draws (joint build); block-uniform
    const uint32_t tg = (uint32_t)lv.tree_global[tr];
    const uint32_t nid = lv.id[i];
    const int f = (int)(rand_u32(dd.seed, kStreamFeat, tg, nid, (uint64_t)j) % (uint32_t)dd.F);
    const int m = lv.m[i];
    const int64_t st = lv.start[i];
    const float* Xf = dd.X + (int64_t)f * dd.N;
    if (t < MAXC) s_left[t] = 0;
    if (t == 0) {
      s_nl = 0;
      s_sl = 0;
      float acc = 0.0f;
      for (int s = 0; s < 10; ++s) {
        const uint32_t pos = rand_u32(dd.seed, kStreamThr, tg, nid, (uint64_t)j * 10 + s) % (uint32_t)m;
        acc = acc + Xf[lv.row[st + pos]];
      }
      s_thr = acc / 10.0f;
      my_bytes += (unsigned long long)m * (TASK == 0 ? 12ull : 20ull);
    }
    __syncthreads();
    const float thr = s_thr;
    uint32_t loc[MAXC];
#pragma unroll
    for (int c = 0; c < MAXC; ++c) loc[c] = 0;
    uint64_t nl = 0;
    int64_t sl = 0;
    for (int e = t; e < m; e += kBlock) {
      const uint32_t row = lv.row[st + e];
      if (side_of(Xf[row], thr) == 0) {
        const uint32_t a = lv.aux[st + e];
        const uint32_t w = a & 0xFFFFu;
        nl += w;
        if (TASK == 0) {
          const int lab = (int)(a >> 16);
#pragma unroll
          for (int c = 0; c < MAXC; ++c) loc[c] += (c == lab) ? w : 0u;
        } else {
          sl += (int64_t)w * lv.ey[st + e];
        }
      }
    }
    if (TASK == 0) {
#pragma unroll
      for (int c = 0; c < MAXC; ++c) {
        uint32_t v = loc[c];
#pragma unroll
        for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
        if ((t & (kWave - 1)) == 0 && v) atomicAdd(&s_left[c], v);
      }
    }
    {
      uint64_t v = nl;
      int64_t sm = sl;
#pragma unroll
      for (int off = kWave / 2; off > 0; off >>= 1) {
        v += __shfl_xor(v, off, kWave);
        sm += __shfl_xor(sm, off, kWave);
      }
      if ((t & (kWave - 1)) == 0) {
        atomicAdd(&s_nl, (unsigned long long)v);
        if (TASK == 1) atomicAdd(&s_sl, (unsigned long long)sm);
      }
    }
    __syncthreads();
    if (t == 0) {
      const int64_t n = lv.count[i];
      const int64_t nlt = (int64_t)s_nl;
      float score;
      if (nlt == 0 || nlt == n) {
        score = -FLT_MAX;
      } else if (TASK == 0) {
        const uint32_t* nclsp = lv.cls + (int64_t)i * d.C;
        const int F = dd.igF;
        const double inv = 1.0 / ((double)n * kLn2);
        int64_t P = xlnx_fixed((uint64_t)n, F, s_inv, s_logc);
        int64_t sls = 0, srs = 0;
        for (int c = 0; c < d.C; ++c) {
          P -= xlnx_fixed(nclsp[c], F, s_inv, s_logc);
          sls += xlnx_fixed(s_left[c], F, s_inv, s_logc);
          srs += xlnx_fixed(nclsp[c] - s_left[c], F, s_inv, s_logc);
        }
        const int64_t Q = (sls - xlnx_fixed((uint64_t)nlt, F, s_inv, s_logc)) +
                          (srs - xlnx_fixed((uint64_t)(n - nlt), F, s_inv, s_logc));
        score = (float)ig_from_fixed(P + Q, F, inv);
      } else {
        const int64_t slt = (int64_t)s_sl;
        score = (float)variance_gain(nlt, slt, n - nlt, lv.sum[i] - slt, dd.shift);
      }
      const int64_t ridx = (int64_t)o * it.D + j;
      it.res_score[ridx] = score;
      it.res_thr[ridx] = thr;
      it.res_feat[ridx] = f;
    }
    __syncthreads();
  }
  if (t == 0 && my_bytes && it.bytes) atomicAdd(it.bytes, my_bytes);
}
template <int MAXC, int TASK>
__global__ void __launch_bounds__(kBlock) k_search_ert(DataArgs d, LevelArgs lv, ItemArgs it) {
  __shared__ double s_inv[128], s_logc[128];
  ert_block_body<MAXC, TASK>(d, lv, it, blockIdx.x, gridDim.x, s_inv, s_logc);
}

// ------------------------------------------ wave-per-item split search --
// Small nodes: one wave64 per (node, draw) item, several items per block, no
// block barriers.  RF: each lane holds one entry (m <= 64); the rank of its
// key is counted over the wave, ds_permute moves the entries into key order,
// and the per-class running sums of the fixed-point x ln x terms become
// wave exclusive scans of per-lane deltas.
__device__ __forceinline__ uint32_t ds_permute_u32(uint32_t dst_lane, uint32_t v) {
  return (uint32_t)__builtin_amdgcn_ds_permute((int)(dst_lane << 2), (int)v);
}

__device__ __forceinline__ int64_t wave_excl_i64(int64_t v) { return wave_incl_scan_i64(v) - v; }

__device__ __forceinline__ int64_t wave_sum_i64(int64_t v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v = max(v, (uint32_t)__shfl_xor(v, off, kWave));
  return v;
}


template <int MAXC, int TASK>
__global__ void __launch_bounds__(kBlock) k_search_rf_wave(DataArgs d, LevelArgs lv, ItemArgs it) {
  resolve_level(lv, d.C);
  resolve_items(it);
  if ((int64_t)blockIdx.x * kWaves >= it.n_groups * it.per_group) return;
  __shared__ double s_inv[128], s_logc[128];
  load_tables(s_inv, s_logc);
  __syncthreads();
  const LnCtx lx{s_inv, s_logc, d.T, d.Tmax, d.igF};
  const int lane = threadIdx.x & (kWave - 1);
  const int wv = threadIdx.x / kWave;
  unsigned long long my_bytes = 0;
  const int64_t n_items = it.n_groups * it.per_group;
  for (int64_t item = (int64_t)blockIdx.x * kWaves + wv; item < n_items; item += (int64_t)gridDim.x * kWaves) {
    const int64_t grp = item / it.per_group;
    const int j = it.j0 + (int)(item % it.per_group);
    const uint32_t o = it.list ? it.list[grp] : (uint32_t)grp;
    const uint32_t i = lv.open_idx[o];
    const uint32_t nid = lv.id[i];
    const int f = (int)(rand_u32(d.seed, kStreamFeat, (uint32_t)lv.tree_global[lv.tree[i]], nid, (uint64_t)j) %
                        (uint32_t)d.F);
    const int m = lv.m[i];
    const int64_t st = lv.start[i];
    const int64_t n = lv.count[i];
    const int64_t ridx = (int64_t)o * it.D + j;
    // node statistics
    uint32_t ncls_lane = 0;
    int F = 0;
    int64_t P = 0, SR0 = 0;
    double inv = 0.0;
    if (TASK == 0) {
      ncls_lane = lane < d.C ? lv.cls[(int64_t)i * d.C + lane] : 0u;
      if ((int64_t)wave_max_u32(ncls_lane) == n) {
        if (lane == 0) {
          it.res_score[ridx] = 0.0f;
          it.res_thr[ridx] = 0.0f;
          it.res_feat[ridx] = f;
        }
        continue;
      }
      F = d.igF;
      inv = 1.0 / ((double)n * kLn2);
      SR0 = wave_sum_i64(lane < d.C ? lx.t(ncls_lane) : 0);
      P = lx.t((uint64_t)n) - SR0;
    }
    if (lane == 0) my_bytes += (unsigned long long)m * (TASK == 0 ? 12ull : 20ull);
    const bool valid = lane < m;
    uint32_t key = kNone, pay = 0;
    int64_t yv = 0;
    if (valid) {
      const uint32_t row = lv.row[st + lane];
      pay = lv.aux[st + lane];
      key = d.K[(int64_t)f * d.N + row];
      if (TASK == 1) yv = lv.ey[st + lane];
    }
    uint32_t cnt = 0;
    for (int r = 0; r < m; ++r) {
      const uint32_t kr = __shfl(key, r, kWave);
      cnt += (kr < key || (kr == key && r < lane)) ? 1u : 0u;
    }
    const uint32_t rank = valid ? cnt : (uint32_t)lane;
    key = ds_permute_u32(rank, key);
    pay = ds_permute_u32(rank, pay);
    if (TASK == 1) {
      const uint32_t lo = ds_permute_u32(rank, (uint32_t)(uint64_t)yv);
      const uint32_t hi = ds_permute_u32(rank, (uint32_t)((uint64_t)yv >> 32));
      yv = (int64_t)(((uint64_t)hi << 32) | lo);
    }
    const uint32_t w = valid ? (pay & 0xFFFFu) : 0u;
    const uint32_t prevk = __shfl_up(key, 1, kWave);
    const bool cand = valid && lane >= 1 && prevk < key;
    double g = 0.0;
    if (TASK == 0) {
      const uint32_t lab = pay >> 16;
      uint32_t pre = 0;
      for (int r = 0; r < m; ++r) {
        const uint32_t pr = __shfl(pay, r, kWave);
        pre += (r < lane && (pr >> 16) == lab) ? (pr & 0xFFFFu) : 0u;
      }
      const uint32_t nc = __shfl(ncls_lane, (int)lab, kWave);
      int64_t dl = 0, dr = 0;
      if (valid) {
        dl = lx.t(pre + w) - lx.t(pre);
        dr = lx.t(nc - pre - w) - lx.t(nc - pre);
      }
      const int64_t sl = wave_excl_i64(dl);
      const int64_t sr = SR0 + wave_excl_i64(dr);
      const int64_t nl = wave_excl_i64((int64_t)w);
      if (cand) {
        const int64_t q = (sl - lx.t((uint64_t)nl)) + (sr - lx.t((uint64_t)(n - nl)));
        g = ig_from_fixed(P + q, F, inv);
      }
    } else {
      const int64_t pn = wave_excl_i64((int64_t)w);
      const int64_t ps = wave_excl_i64((int64_t)w * yv);
      if (cand) g = variance_gain(pn, ps, n - pn, lv.sum[i] - ps, 0);
    }
    BestCand b = (cand && g > 0.0) ? BestCand{g, key, prevk} : BestCand{0.0, kNone, kNone};
    b = wave_best(b);
    if (lane == 0) {
      float score = 0.0f, thr = 0.0f;
      if (b.key != kNone) {
        const float* sv = d.SV + d.SVoff[f];
        score = (float)(TASK == 0 ? b.g : ldexp_d(b.g, -2 * d.shift));
        thr = midpoint(sv[b.prev], sv[b.key]);
      }
      it.res_score[ridx] = score;
      it.res_thr[ridx] = thr;
      it.res_feat[ridx] = f;
    }
  }
  if (lane == 0 && my_bytes && it.bytes) atomicAdd(it.bytes, my_bytes);
}

// Segment-per-node RF search for tiny nodes: a wave holds 64/SEG nodes, one
// per SEG-lane segment (one entry per lane, m <= SEG, C <= SEG), and loops
// over the node's draws, so node statistics and entry loads are shared by
// all draws.  Shuffles, scans and reductions stay inside the segment.
template <int SEG>
__device__ __forceinline__ int64_t seg_excl_i64(int64_t v) {
  const int sl = threadIdx.x & (SEG - 1);
  int64_t x = v;
#pragma unroll
  for (int off = 1; off < SEG; off <<= 1) {
    const int64_t y = __shfl_up(x, off, SEG);
    if (sl >= off) x += y;
  }
  return x - v;
}

template <int SEG>
__device__ __forceinline__ int64_t seg_sum_i64(int64_t v) {
#pragma unroll
  for (int off = SEG / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, SEG);
  return v;
}

template <int SEG>
__device__ __forceinline__ uint32_t seg_max_u32(uint32_t v) {
#pragma unroll
  for (int off = SEG / 2; off > 0; off >>= 1) v = max(v, (uint32_t)__shfl_xor(v, off, SEG));
  return v;
}

template <int SEG>
__device__ __forceinline__ BestCand seg_best(BestCand b) {
#pragma unroll
  for (int off = SEG / 2; off > 0; off >>= 1) {
    const double g = __shfl_xor(b.g, off, SEG);
    const uint32_t k = __shfl_xor(b.key, off, SEG);
    const uint32_t p = __shfl_xor(b.prev, off, SEG);
    if (cand_better(g, k, b.g, b.key)) {
      b.g = g;
      b.key = k;
      b.prev = p;
    }
  }
  return b;
}

template <int MAXC, int TASK, int SEG>
__global__ void __launch_bounds__(kBlock, SEG == 64 ? 4 : 1) k_search_rf_seg(DataArgs d, LevelArgs lv, ItemArgs it) {
  constexpr int NPW = kWave / SEG;  // nodes per wave
  resolve_level(lv, d.C);
  resolve_items(it);
  if ((int64_t)blockIdx.x * kWaves * NPW >= it.n_groups) return;
  constexpr int kTL = 2048;         // LDS copy of T's head
  __shared__ double s_inv[128], s_logc[128];
  __shared__ int64_t s_T[TASK == 0 ? kTL : 1];
  load_tables(s_inv, s_logc);
  const int64_t nTL = TASK == 0 ? load_t_prefix<kTL>(s_T, d.T, d.Tmax) : 0;
  __syncthreads();
  const LnCtxL lx{s_inv, s_logc, d.T, d.Tmax, d.igF, s_T, nTL};
  const int sl = threadIdx.x & (SEG - 1);
  const int seg = (threadIdx.x & (kWave - 1)) / SEG;
  unsigned long long my_bytes = 0;
  const int64_t wave_id = (int64_t)blockIdx.x * kWaves + threadIdx.x / kWave;
  const int64_t n_waves = (int64_t)gridDim.x * kWaves;
  for (int64_t g0 = wave_id * NPW; g0 < it.n_groups; g0 += n_waves * NPW) {
    const int64_t g = g0 + seg;
    const bool node_ok = g < it.n_groups;
    uint32_t o = 0, i = 0, tg = 0, nid = 0;
    int m = 0;
    int64_t st = 0, n = 0, ssum = 0;
    if (node_ok) {
      o = it.list ? it.list[g] : (uint32_t)g;
      i = lv.open_idx[o];
      tg = (uint32_t)lv.tree_global[lv.tree[i]];
      nid = lv.id[i];
      m = lv.m[i];
      st = lv.start[i];
      n = lv.count[i];
      if (TASK == 1) ssum = lv.sum[i];
    }
    // node statistics (shared by all draws)
    uint32_t ncls_lane = 0;
    int64_t P = 0, SR0 = 0;
    double inv = 0.0;
    bool pure = false;
    if (TASK == 0) {
      ncls_lane = (node_ok && sl < d.C) ? lv.cls[(int64_t)i * d.C + sl] : 0u;
      pure = (int64_t)seg_max_u32<SEG>(ncls_lane) == n;
      SR0 = seg_sum_i64<SEG>(sl < d.C ? lx.t(ncls_lane) : 0);
      P = lx.t((uint64_t)n) - SR0;
      inv = 1.0 / ((double)n * kLn2);
    }
    const bool valid = node_ok && sl < m;
    uint32_t row = 0, pay0 = 0;
    int64_t y0 = 0;
    if (valid) {
      row = lv.row[st + sl];
      pay0 = lv.aux[st + sl];
      if (TASK == 1) y0 = lv.ey[st + sl];
    }
    if (sl == 0 && node_ok && !pure) my_bytes += (unsigned long long)m * it.per_group * (TASK == 0 ? 12ull : 20ull);
    // Draw pipeline: the key gather (and value offset) of draw jj + 1 and the
    // threshold values of draw jj - 1 are in flight while draw jj is ranked
    // and scored.
    int f_nx = 0;
    uint32_t key_nx = kNone;
    int64_t svo_nx = 0;
    if (!pure && it.per_group > 0) {
      f_nx = (int)(rand_u32(d.seed, kStreamFeat, tg, nid, (uint64_t)it.j0) % (uint32_t)d.F);
      key_nx = valid ? d.K[(int64_t)f_nx * d.N + row] : kNone;
      svo_nx = d.SVoff[f_nx];
    }
    bool pend = false, p_has = false;
    int64_t p_ridx = 0;
    int p_f = 0;
    float p_score = 0.0f, p_lo = 0.0f, p_hi = 0.0f;
    for (int jj = 0; jj < it.per_group; ++jj) {
      const int j = it.j0 + jj;
      const int64_t ridx = (int64_t)o * it.D + j;
      if (pure) {
        if (sl == 0 && node_ok) {
          it.res_score[ridx] = 0.0f;
          it.res_thr[ridx] = 0.0f;
          it.res_feat[ridx] = (int)(rand_u32(d.seed, kStreamFeat, tg, nid, (uint64_t)j) % (uint32_t)d.F);
        }
        continue;
      }
      const int f = f_nx;
      const int64_t svo = svo_nx;
      uint32_t key = key_nx;
      if (jj + 1 < it.per_group) {
        f_nx = (int)(rand_u32(d.seed, kStreamFeat, tg, nid, (uint64_t)(j + 1)) % (uint32_t)d.F);
        key_nx = valid ? d.K[(int64_t)f_nx * d.N + row] : kNone;
        svo_nx = d.SVoff[f_nx];
      }
      // rank inside the segment (ties by lane), then move entries into key order
      uint32_t cnt = 0;
#pragma unroll
      for (int r = 0; r < SEG; ++r) {
        const uint32_t kr = __shfl(key, r, SEG);
        cnt += (r < m && (kr < key || (kr == key && r < sl))) ? 1u : 0u;
      }
      const uint32_t rank = valid ? cnt : (uint32_t)sl;
      const uint32_t dst = (uint32_t)(threadIdx.x & (kWave - 1) & ~(SEG - 1)) + rank;
      key = ds_permute_u32(dst, key);
      const uint32_t pay = ds_permute_u32(dst, pay0);
      int64_t yv = 0;
      if (TASK == 1) {
        const uint32_t lo = ds_permute_u32(dst, (uint32_t)(uint64_t)y0);
        const uint32_t hi = ds_permute_u32(dst, (uint32_t)((uint64_t)y0 >> 32));
        yv = (int64_t)(((uint64_t)hi << 32) | lo);
      }
      const uint32_t w = valid ? (pay & 0xFFFFu) : 0u;
      const uint32_t prevk = __shfl_up(key, 1, SEG);
      const bool cand = valid && sl >= 1 && prevk < key;
      double gval = 0.0;
      if (TASK == 0) {
        const uint32_t lab = pay >> 16;
        uint32_t pre = 0;
#pragma unroll
        for (int r = 0; r < SEG; ++r) {
          const uint32_t pr = __shfl(pay, r, SEG);
          pre += (r < sl && (pr >> 16) == lab) ? (pr & 0xFFFFu) : 0u;
        }
        const uint32_t ncl = __shfl(ncls_lane, (int)lab, SEG);
        int64_t dl = 0, dr = 0;
        if (valid) {
          dl = lx.t(pre + w) - lx.t(pre);
          dr = lx.t(ncl - pre - w) - lx.t(ncl - pre);
        }
        const int64_t slv = seg_excl_i64<SEG>(dl);
        const int64_t srv = SR0 + seg_excl_i64<SEG>(dr);
        const int64_t nl = seg_excl_i64<SEG>((int64_t)w);
        if (cand) {
          const int64_t q = (slv - lx.t((uint64_t)nl)) + (srv - lx.t((uint64_t)(n - nl)));
          gval = ig_from_fixed(P + q, lx.F, inv);
        }
      } else {
        const int64_t pn = seg_excl_i64<SEG>((int64_t)w);
        const int64_t ps = seg_excl_i64<SEG>((int64_t)w * yv);
        if (cand) gval = variance_gain(pn, ps, n - pn, ssum - ps, 0);
      }
      BestCand b = (cand && gval > 0.0) ? BestCand{gval, key, prevk} : BestCand{0.0, kNone, kNone};
      b = seg_best<SEG>(b);
      if (pend) {
        it.res_score[p_ridx] = p_score;
        it.res_thr[p_ridx] = p_has ? midpoint(p_lo, p_hi) : 0.0f;
        it.res_feat[p_ridx] = p_f;
        pend = false;
      }
      if (sl == 0 && node_ok) {
        pend = true;
        p_ridx = ridx;
        p_f = f;
        p_has = b.key != kNone;
        p_score = 0.0f;
        if (p_has) {
          const float* sv = d.SV + svo;
          p_score = (float)(TASK == 0 ? b.g : ldexp_d(b.g, -2 * d.shift));
          p_lo = sv[b.prev];
          p_hi = sv[b.key];
        }
      }
    }
    if (pend) {
      it.res_score[p_ridx] = p_score;
      it.res_thr[p_ridx] = p_has ? midpoint(p_lo, p_hi) : 0.0f;
      it.res_feat[p_ridx] = p_f;
    }
  }
  if (sl == 0 && my_bytes && it.bytes) atomicAdd(it.bytes, my_bytes);
}

// ERT, one wave per item (any node size; lanes stride over the entries).
template <int MAXC, int TASK>
__device__ __forceinline__ void ert_wave_body(DataArgs d, LevelArgs lv, ItemArgs it, int64_t bid, int64_t nblk,
                                              double* s_inv, double* s_logc) {
  resolve_level(lv, d.C);
  resolve_items(it);
  if ((int64_t)bid * kWaves >= it.n_groups * it.per_group) return;
  load_ln_table(s_inv, s_logc);
  __syncthreads();
  const int lane = threadIdx.x & (kWave - 1);
  const int wv = threadIdx.x / kWave;
  unsigned long long my_bytes = 0;
  const int64_t n_items = it.n_groups * it.per_group;
  for (int64_t item = (int64_t)bid * kWaves + wv; item < n_items; item += (int64_t)nblk * kWaves) {
    const int64_t grp = item / it.per_group;
    const uint32_t o = it.list ? it.list[grp] : (uint32_t)grp;
    const uint32_t i = lv.open_idx[o];
    const int tri = lv.tree[i];
    const DataArgs dd = forest_data(d, tri);
    const int j = item_draw(d, dd, it, (int)(item % it.per_group));
    if (j < 0) continue;  // past this forest's draws (joint build); wave-uniform
    const uint32_t tg = (uint32_t)lv.tree_global[tri];
    const uint32_t nid = lv.id[i];
    const int f = (int)(rand_u32(dd.seed, kStreamFeat, tg, nid, (uint64_t)j) % (uint32_t)dd.F);
    const int m = lv.m[i];
    const int64_t st = lv.start[i];
    const int64_t n = lv.count[i];
    const float* Xf = dd.X + (int64_t)f * dd.N;
    float xs = 0.0f;
    if (lane < 10) {
      const uint32_t pos = rand_u32(dd.seed, kStreamThr, tg, nid, (uint64_t)j * 10 + lane) % (uint32_t)m;
      xs = Xf[lv.row[st + pos]];
    }
    float acc = 0.0f;
    for (int s = 0; s < 10; ++s) acc = acc + __shfl(xs, s, kWave);
    const float thr = acc / 10.0f;
    // Left-side counts.  Extra-trees entries weigh 1 (every row once), so a
    // chunk's class counts are popcounts of ballots: lane c accumulates
    // class c in lc, the left total is wave-uniform in nl_u.  A chunk with
    // another weight takes per-lane accumulators reduced at the end.
    uint32_t loc[MAXC];
#pragma unroll
    for (int c = 0; c < MAXC; ++c) loc[c] = 0;
    uint32_t lc = 0;
    int64_t nl_u = 0, nl_l = 0, sl = 0;
    bool any_loc = false;  // wave-uniform
    for (int e0 = 0; e0 < m; e0 += kWave) {
      const int e = e0 + lane;
      bool left = false;
      uint32_t a = 0;
      if (e < m) {
        const uint32_t row = lv.row[st + e];
        left = side_of(Xf[row], thr) == 0;
        if (left) a = lv.aux[st + e];
      }
      const uint32_t w = a & 0xFFFFu;
      if (TASK == 0) {
        const int lab = (int)(a >> 16);
        if (__ballot(left && w != 1u) == 0) {
          nl_u += __popcll(__ballot(left));
#pragma unroll
          for (int c = 0; c < MAXC; ++c)
            if (c < d.C) {
              const uint32_t cnt = (uint32_t)__popcll(__ballot(left && lab == c));
              if (lane == c) lc += cnt;
            }
        } else {
          any_loc = true;
          if (left) {
            nl_l += w;
#pragma unroll
            for (int c = 0; c < MAXC; ++c) loc[c] += (c == lab) ? w : 0u;
          }
        }
      } else if (left) {
        nl_l += w;
        sl += (int64_t)w * lv.ey[st + e];
      }
    }
    const int64_t nl = nl_u + ((TASK == 1 || any_loc) ? wave_sum_i64(nl_l) : 0);
    if (lane == 0) my_bytes += (unsigned long long)m * (TASK == 0 ? 12ull : 20ull);
    float score;
    if (nl == 0 || nl == n) {
      score = -FLT_MAX;
    } else if (TASK == 0) {
      // lane c keeps class c's left count
      if (any_loc) {
#pragma unroll
        for (int c = 0; c < MAXC; ++c) {
          uint32_t v = loc[c];
#pragma unroll
          for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
          lc += (lane == c) ? v : 0u;
        }
      }
      const int F = dd.igF;
      const double inv = 1.0 / ((double)n * kLn2);
      int64_t tn = 0, tl = 0, tr = 0;
      if (lane < d.C) {
        const uint32_t nc = lv.cls[(int64_t)i * d.C + lane];
        tn = xlnx_fixed(nc, F, s_inv, s_logc);
        tl = xlnx_fixed(lc, F, s_inv, s_logc);
        tr = xlnx_fixed(nc - lc, F, s_inv, s_logc);
      }
      const int64_t P = xlnx_fixed((uint64_t)n, F, s_inv, s_logc) - wave_sum_i64(tn);
      const int64_t Q = (wave_sum_i64(tl) - xlnx_fixed((uint64_t)nl, F, s_inv, s_logc)) +
                        (wave_sum_i64(tr) - xlnx_fixed((uint64_t)(n - nl), F, s_inv, s_logc));
      score = (float)ig_from_fixed(P + Q, F, inv);
    } else {
      sl = wave_sum_i64(sl);
      score = (float)variance_gain(nl, sl, n - nl, lv.sum[i] - sl, dd.shift);
    }
    if (lane == 0) {
      const int64_t ridx = (int64_t)o * it.D + j;
      it.res_score[ridx] = score;
      it.res_thr[ridx] = thr;
      it.res_feat[ridx] = f;
    }
  }
  if (lane == 0 && my_bytes && it.bytes) atomicAdd(it.bytes, my_bytes);
}
template <int MAXC, int TASK>
__global__ void __launch_bounds__(kBlock) k_search_ert_wave(DataArgs d, LevelArgs lv, ItemArgs it) {
  __shared__ double s_inv[128], s_logc[128];
  ert_wave_body<MAXC, TASK>(d, lv, it, blockIdx.x, gridDim.x, s_inv, s_logc);
}

// ERT, one node per SEG-lane segment (nodes of at most SEG entries and at
// most SEG classes; kWave / SEG nodes per wave).  A node's entries sit one
// per lane and stay in registers across its draws: a draw takes its ten
// threshold samples from those lanes by shuffles, so its two gathers (the
// samples and the entries' values of the drawn feature) are independent,
// and the gathers of draw q + 1 are issued before draw q is counted and
// scored.  The draw-independent entropy term of the node is computed once.
// Every draw's (score, threshold, feature) equals k_search_ert_wave's bit
// for bit: the same Philox draws, the same sequential fp32 threshold sum
// (gpuert.cu:167-183), exact integer side counts (gpuert.cu:287-301) or
// target sums (gpuert.cu:348-361) and the same fixed-point entropy sums.
template <int MAXC, int TASK, int SEG>
__device__ __forceinline__ void ert_seg_body(DataArgs d, LevelArgs lv, ItemArgs it, int64_t bid, int64_t nblk,
                                             double* s_inv, double* s_logc) {
  constexpr int NPW = kWave / SEG;  // nodes per wave
  resolve_level(lv, d.C);
  resolve_items(it);
  if ((int64_t)bid * kWaves * NPW >= it.n_groups) return;
  load_ln_table(s_inv, s_logc);
  __syncthreads();
  const int lane = threadIdx.x & (kWave - 1);
  const int sl = lane & (SEG - 1);
  const int base = lane & ~(SEG - 1);
  const unsigned long long segmask = SEG == kWave ? ~0ull : (((1ull << (SEG & 63)) - 1ull) << base);
  unsigned long long my_bytes = 0;
  const int64_t wave_id = (int64_t)bid * kWaves + threadIdx.x / kWave;
  const int64_t n_waves = (int64_t)nblk * kWaves;
  for (int64_t g0 = wave_id * NPW; g0 < it.n_groups; g0 += n_waves * NPW) {
    const int64_t g = g0 + lane / SEG;
    const bool node_ok = g < it.n_groups;
    uint32_t o = 0, i = 0, tg = 0, nid = 0;
    int m = 0;
    int64_t st = 0, n = 0, ssum = 0;
    int tri = 0;
    if (node_ok) {
      o = it.list ? it.list[g] : (uint32_t)g;
      i = lv.open_idx[o];
      tri = lv.tree[i];
      tg = (uint32_t)lv.tree_global[tri];
      nid = lv.id[i];
      m = lv.m[i];
      st = lv.start[i];
      n = lv.count[i];
      if (TASK == 1) ssum = lv.sum[i];
    }
    const DataArgs dd = forest_data(d, tri);
    const bool valid = node_ok && sl < m;
    uint32_t row = 0, a = 0;
    int64_t ey = 0;
    if (valid) {
      row = lv.row[st + sl];
      a = lv.aux[st + sl];
      if (TASK == 1) ey = lv.ey[st + sl];
    }
    const uint32_t w = a & 0xFFFFu;
    const int lab = (int)(a >> 16);
    // Extra-trees entries weigh 1 (every row once): counts are popcounts of
    // ballots.  A wave holding another weight sums weights by shuffles.
    const bool unit_w = __ballot(valid && w != 1u) == 0;  // wave-uniform
    const int F = dd.igF;
    uint32_t nc = 0;
    int64_t P = 0;
    double inv = 0.0;
    if (TASK == 0) {
      if (node_ok && sl < d.C) nc = lv.cls[(int64_t)i * d.C + sl];
      P = xlnx_fixed((uint64_t)n, F, s_inv, s_logc) -
          seg_sum_i64<SEG>(sl < d.C ? xlnx_fixed(nc, F, s_inv, s_logc) : 0);
      inv = 1.0 / ((double)n * kLn2);
    }
    const uint32_t mm = m > 0 ? (uint32_t)m : 1u;
    // The gathers of draw j (< 0: none): the drawn feature, this lane's
    // threshold sample (lanes sl < 10) and its entry's value.
    auto gather = [&](int j, int& f, float& xs, float& xe) {
      const uint32_t jj = j >= 0 ? (uint32_t)j : 0u;
      f = (int)(rand_u32(dd.seed, kStreamFeat, tg, nid, (uint64_t)jj) % (uint32_t)dd.F);
      const uint32_t pos =
          rand_u32(dd.seed, kStreamThr, tg, nid, (uint64_t)jj * 10 + (uint32_t)(sl < 10 ? sl : 0)) % mm;
      const uint32_t rp = (uint32_t)__shfl((int)row, base + (int)pos, kWave);
      const float* Xf = dd.X + (int64_t)f * dd.N;
      xs = (j >= 0 && node_ok && sl < 10) ? Xf[rp] : 0.0f;
      xe = (j >= 0 && valid) ? Xf[row] : 0.0f;
    };
    auto draw_of = [&](int q) { return (node_ok && q < it.per_group) ? item_draw(d, dd, it, q) : -1; };
    int j_c = draw_of(0), f_c = 0;
    float xs_c = 0.0f, xe_c = 0.0f;
    gather(j_c, f_c, xs_c, xe_c);
    for (int q = 0; q < it.per_group; ++q) {
      const int j_n = draw_of(q + 1);
      int f_n = 0;
      float xs_n = 0.0f, xe_n = 0.0f;
      if (q + 1 < it.per_group) gather(j_n, f_n, xs_n, xe_n);
      float acc = 0.0f;
      for (int s = 0; s < 10; ++s) acc = acc + __shfl(xs_c, base + s, kWave);
      const float thr = acc / 10.0f;
      const bool left = valid && j_c >= 0 && side_of(xe_c, thr) == 0;
      int64_t nl, sls = 0;
      uint32_t lc = 0;
      if (unit_w) {
        nl = __popcll(__ballot(left) & segmask);
      } else {
        nl = seg_sum_i64<SEG>(left ? (int64_t)w : 0);
      }
      if (TASK == 0) {
#pragma unroll
        for (int c = 0; c < MAXC; ++c)
          if (c < d.C) {
            uint32_t cnt;
            if (unit_w)
              cnt = (uint32_t)__popcll(__ballot(left && lab == c) & segmask);
            else
              cnt = (uint32_t)seg_sum_i64<SEG>((left && lab == c) ? (int64_t)w : 0);
            if (sl == c) lc = cnt;
          }
      } else {
        sls = seg_sum_i64<SEG>(left ? (int64_t)w * ey : 0);
      }
      float score;
      if (TASK == 0) {
        int64_t tl = 0, tr = 0;
        if (sl < d.C) {
          tl = xlnx_fixed(lc, F, s_inv, s_logc);
          tr = xlnx_fixed(nc - lc, F, s_inv, s_logc);
        }
        const int64_t Q = (seg_sum_i64<SEG>(tl) - xlnx_fixed((uint64_t)nl, F, s_inv, s_logc)) +
                          (seg_sum_i64<SEG>(tr) - xlnx_fixed((uint64_t)(n - nl), F, s_inv, s_logc));
        score = (float)ig_from_fixed(P + Q, F, inv);
      } else {
        score = (float)variance_gain(nl, sls, n - nl, ssum - sls, dd.shift);
      }
      if (nl == 0 || nl == n) score = -FLT_MAX;
      if (sl == 0 && j_c >= 0) {
        const int64_t ridx = (int64_t)o * it.D + j_c;
        it.res_score[ridx] = score;
        it.res_thr[ridx] = thr;
        it.res_feat[ridx] = f_c;
        my_bytes += (unsigned long long)m * (TASK == 0 ? 12ull : 20ull);
      }
      j_c = j_n;
      f_c = f_n;
      xs_c = xs_n;
      xe_c = xe_n;
    }
  }
  if (my_bytes && it.bytes) atomicAdd(it.bytes, my_bytes);
}
// Four waves per SIMD (128 VGPRs, no spills; three at the unbounded 131).
template <int MAXC, int TASK, int SEG>
__global__ void __launch_bounds__(kBlock, 4) k_search_ert_seg(DataArgs d, LevelArgs lv, ItemArgs it) {
  __shared__ double s_inv[128], s_logc[128];
  ert_seg_body<MAXC, TASK, SEG>(d, lv, it, blockIdx.x, gridDim.x, s_inv, s_logc);
}

// ERT, one wave per node of at most kErtTeamIpt * 64 entries, all draws in
// order.  The node's rows and payloads are read once into registers (slot
// q of lane l holds entry l + 64 q); a draw gathers its feature's values
// of all entries at once, and the next draw's feature, threshold samples
// and gathers are in flight while the current draw is counted and scored.
// Results equal k_search_ert_wave's bit for bit (the same draws, the same
// sequential fp32 threshold sum, exact integer counts and fixed-point
// entropy sums; gpuert.cu:144-230, 287-301, 348-361).
constexpr int kErtTeamIpt = 8;

template <int MAXC, int TASK>
__device__ __forceinline__ void ert_team_body(DataArgs d, LevelArgs lv, ItemArgs it, int64_t bid, int64_t nblk,
                                              double* s_inv, double* s_logc) {
  constexpr int IPT = kErtTeamIpt;
  resolve_level(lv, d.C);
  resolve_items(it);
  if ((int64_t)bid * kWaves >= it.n_groups) return;
  load_ln_table(s_inv, s_logc);
  __syncthreads();
  const int lane = threadIdx.x & (kWave - 1);
  unsigned long long my_bytes = 0;
  const int64_t wave_id = (int64_t)bid * kWaves + threadIdx.x / kWave;
  const int64_t n_waves = (int64_t)nblk * kWaves;
  for (int64_t g = wave_id; g < it.n_groups; g += n_waves) {
    const uint32_t o = it.list ? it.list[g] : (uint32_t)g;
    const uint32_t i = lv.open_idx[o];
    const int tri = lv.tree[i];
    const DataArgs dd = forest_data(d, tri);
    const uint32_t tg = (uint32_t)lv.tree_global[tri];
    const uint32_t nid = lv.id[i];
    const int m = lv.m[i];
    const int64_t st = lv.start[i];
    const int64_t n = lv.count[i];
    const int64_t ssum = TASK == 1 ? lv.sum[i] : 0;
    uint32_t rows[IPT], aux[IPT];
    int64_t eys[IPT];
    bool anyw = false;
#pragma unroll
    for (int q = 0; q < IPT; ++q) {
      const int e = lane + q * kWave;
      rows[q] = 0;
      aux[q] = 0;
      eys[q] = 0;
      if (e < m) {
        rows[q] = lv.row[st + e];
        aux[q] = lv.aux[st + e];
        if (TASK == 1) eys[q] = lv.ey[st + e];
        anyw |= (aux[q] & 0xFFFFu) != 1u;
      }
    }
    // Extra-trees entries weigh 1 (every row once): counts are popcounts of
    // ballots.  A node holding another weight sums weights by shuffles.
    const bool unit_w = __ballot(anyw) == 0;  // wave-uniform
    const int F = dd.igF;
    uint32_t nc = 0;
    int64_t P = 0;
    double inv = 0.0;
    if (TASK == 0) {
      if (lane < d.C) nc = lv.cls[(int64_t)i * d.C + lane];
      P = xlnx_fixed((uint64_t)n, F, s_inv, s_logc) - wave_sum_i64(lane < d.C ? xlnx_fixed(nc, F, s_inv, s_logc) : 0);
      inv = 1.0 / ((double)n * kLn2);
    }
    // Gathers of draw j: the drawn feature, this lane's threshold sample
    // (lanes < 10) and every entry's value.
    auto gather = [&](int j, int& f, float& xs, float* xe) {
      const uint32_t jj = j >= 0 ? (uint32_t)j : 0u;
      f = (int)(rand_u32(dd.seed, kStreamFeat, tg, nid, (uint64_t)jj) % (uint32_t)dd.F);
      const float* Xf = dd.X + (int64_t)f * dd.N;
      xs = 0.0f;
      if (j >= 0 && lane < 10) {
        const uint32_t pos = rand_u32(dd.seed, kStreamThr, tg, nid, (uint64_t)jj * 10 + lane) % (uint32_t)m;
        xs = Xf[lv.row[st + pos]];
      }
#pragma unroll
      for (int q = 0; q < IPT; ++q) xe[q] = (j >= 0 && lane + q * kWave < m) ? Xf[rows[q]] : 0.0f;
    };
    int j_c = item_draw(d, dd, it, 0), f_c = 0;
    float xs_c = 0.0f, xe_c[IPT];
    gather(j_c, f_c, xs_c, xe_c);
    for (int qd = 0; qd < it.per_group; ++qd) {
      const int j_n = qd + 1 < it.per_group ? item_draw(d, dd, it, qd + 1) : -1;
      int f_n = 0;
      float xs_n = 0.0f, xe_n[IPT];
      if (qd + 1 < it.per_group) gather(j_n, f_n, xs_n, xe_n);
      if (j_c >= 0) {  // wave-uniform
        float acc = 0.0f;
        for (int s2 = 0; s2 < 10; ++s2) acc = acc + __shfl(xs_c, s2, kWave);
        const float thr = acc / 10.0f;
        bool left[IPT];
#pragma unroll
        for (int q = 0; q < IPT; ++q) left[q] = lane + q * kWave < m && side_of(xe_c[q], thr) == 0;
        int64_t nl = 0, sls = 0;
        uint32_t lc = 0;
        if (unit_w) {
#pragma unroll
          for (int q = 0; q < IPT; ++q) nl += __popcll(__ballot(left[q]));
        } else {
          int64_t v = 0;
#pragma unroll
          for (int q = 0; q < IPT; ++q) v += left[q] ? (int64_t)(aux[q] & 0xFFFFu) : 0;
          nl = wave_sum_i64(v);
        }
        if (TASK == 0) {
          // one class at a time (lane c keeps class c's left count)
#pragma unroll 1
          for (int c = 0; c < d.C; ++c) {
            uint32_t cnt = 0;
            if (unit_w) {
#pragma unroll
              for (int q = 0; q < IPT; ++q) cnt += (uint32_t)__popcll(__ballot(left[q] && (int)(aux[q] >> 16) == c));
            } else {
              int64_t v = 0;
#pragma unroll
              for (int q = 0; q < IPT; ++q)
                v += (left[q] && (int)(aux[q] >> 16) == c) ? (int64_t)(aux[q] & 0xFFFFu) : 0;
              cnt = (uint32_t)wave_sum_i64(v);
            }
            if (lane == c) lc = cnt;
          }
        } else {
          int64_t v = 0;
#pragma unroll
          for (int q = 0; q < IPT; ++q) v += left[q] ? (int64_t)(aux[q] & 0xFFFFu) * eys[q] : 0;
          sls = wave_sum_i64(v);
        }
        float score;
        if (nl == 0 || nl == n) {
          score = -FLT_MAX;
        } else if (TASK == 0) {
          int64_t tl = 0, tr = 0;
          if (lane < d.C) {
            tl = xlnx_fixed(lc, F, s_inv, s_logc);
            tr = xlnx_fixed(nc - lc, F, s_inv, s_logc);
          }
          const int64_t Q = (wave_sum_i64(tl) - xlnx_fixed((uint64_t)nl, F, s_inv, s_logc)) +
                            (wave_sum_i64(tr) - xlnx_fixed((uint64_t)(n - nl), F, s_inv, s_logc));
          score = (float)ig_from_fixed(P + Q, F, inv);
        } else {
          score = (float)variance_gain(nl, sls, n - nl, ssum - sls, dd.shift);
        }
        if (lane == 0) {
          const int64_t ridx = (int64_t)o * it.D + j_c;
          it.res_score[ridx] = score;
          it.res_thr[ridx] = thr;
          it.res_feat[ridx] = f_c;
          my_bytes += (unsigned long long)m * (TASK == 0 ? 12ull : 20ull);
        }
      }
      j_c = j_n;
      f_c = f_n;
      xs_c = xs_n;
#pragma unroll
      for (int q = 0; q < IPT; ++q) xe_c[q] = xe_n[q];
    }
  }
  if (lane == 0 && my_bytes && it.bytes) atomicAdd(it.bytes, my_bytes);
}
template <int MAXC, int TASK>
__global__ void __launch_bounds__(kBlock) k_search_ert_team(DataArgs d, LevelArgs lv, ItemArgs it) {
  __shared__ double s_inv[128], s_logc[128];
  ert_team_body<MAXC, TASK>(d, lv, it, blockIdx.x, gridDim.x, s_inv, s_logc);
}

// All extra-trees tiers of a phase in one launch: workgroups [first[k],
// first[k + 1]) run tier k's code over tier k's list (each tier exactly as
// its own kernel would, with its share of the grid).  Small level steps
// are latency-bound, tier after tier; in one launch the tiers run side by
// side without a join between them.
constexpr int64_t kErtMergeEntries = (int64_t)8 << 20;

struct ErtTiers {
  int64_t first[7];
  const uint32_t* list[6];
  const unsigned int* n_dev[6];
  unsigned long long* bytes[6];
};

// Three waves per SIMD (168 VGPRs, no spills; two at the unbounded 171).
template <int MAXC, int TASK>
__global__ void __launch_bounds__(kBlock, 3) k_search_ert_all(DataArgs d, LevelArgs lv, ItemArgs it, ErtTiers ts) {
  __shared__ double s_inv[128], s_logc[128];
  const int64_t b = blockIdx.x;
  int k = 0;
  while (k < 5 && b >= ts.first[k + 1]) ++k;
  ItemArgs a = it;
  a.list = ts.list[k];
  a.n_groups = 0;
  a.n_dev = ts.n_dev[k];
  a.bytes = ts.bytes[k];
  const int64_t bid = b - ts.first[k], nblk = ts.first[k + 1] - ts.first[k];
  switch (k) {
    case 0: ert_seg_body<MAXC, TASK, 16>(d, lv, a, bid, nblk, s_inv, s_logc); break;
    case 1: ert_seg_body<MAXC, TASK, 32>(d, lv, a, bid, nblk, s_inv, s_logc); break;
    case 2: ert_seg_body<MAXC, TASK, 64>(d, lv, a, bid, nblk, s_inv, s_logc); break;
    case 3: ert_team_body<MAXC, TASK>(d, lv, a, bid, nblk, s_inv, s_logc); break;
    case 4: ert_wave_body<MAXC, TASK>(d, lv, a, bid, nblk, s_inv, s_logc); break;
    default: ert_block_body<MAXC, TASK>(d, lv, a, bid, nblk, s_inv, s_logc); break;
  }
}

template <int MAXC, int TASK>
void launch_search_ert_all(int64_t grid, hipStream_t s, const DataArgs& d, const LevelArgs& lv, const ItemArgs& it,
                           const ErtTiers& ts) {
  hipLaunchKernelGGL((k_search_ert_all<MAXC, TASK>), dim3((unsigned)grid), dim3(kBlock), 0, s, d, lv, it, ts);
}
using ErtAllFn = void (*)(int64_t, hipStream_t, const DataArgs&, const LevelArgs&, const ItemArgs&, const ErtTiers&);
ErtAllFn pick_ert_all(int task, int C) {
  if (task == 1) return launch_search_ert_all<1, 1>;
  if (C <= 2) return launch_search_ert_all<2, 0>;
  if (C <= 4) return launch_search_ert_all<4, 0>;
  if (C <= 8) return launch_search_ert_all<8, 0>;
  if (C <= 16) return launch_search_ert_all<16, 0>;
  return launch_search_ert_all<32, 0>;
}

// Split a list of open positions (or all of them) by node size into the
// wave tier and the block tier.  Order inside a tier is irrelevant: every
// item writes its own result slot.
// Wave-aggregated: one global atomic per wave and tier.  Tier k holds the
// groups with cuts[k-1] < m <= cuts[k].
struct Cuts {
  int c[kMaxTiers];  // ascending; the last tier takes everything above
  int n;             // number of tiers
};
struct TierLists {
  uint32_t* l[kMaxTiers];
};

__device__ __forceinline__ int tier_of(int m, const Cuts& cuts) {
  int tier = cuts.n - 1;
#pragma unroll
  for (int k = 0; k < kMaxTiers - 1; ++k)
    if (k < cuts.n - 1 && m <= cuts.c[k]) {
      tier = k;
      break;
    }
  return tier;
}

// Each wave owns a contiguous range of the list: it counts its tiers, takes
// one slot range per tier with a single atomic, then writes its entries.
// Workgroups of a tier split: every wave adds its tier counts with one
// global atomic per tier, so the grid bounds the atomics on those few words.
constexpr int kTierGrid = 256;
__global__ void k_tier_split(const uint32_t* list, int64_t n, const uint32_t* open_idx, const int32_t* node_m, Cuts cuts,
                             TierLists out, unsigned int* counts, const unsigned int* n_dev, const int64_t* off_dev) {
  if (n_dev) n = *n_dev;            // device-described level
  if (off_dev) node_m += *off_dev;
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t wave = (int64_t)blockIdx.x * kWaves + threadIdx.x / kWave;
  const int64_t n_waves = (int64_t)gridDim.x * kWaves;
  const int64_t per = (((n + n_waves - 1) / n_waves) + kWave - 1) & ~(int64_t)(kWave - 1);
  const int64_t lo = wave * per, hi = min(lo + per, n);
  if (lo >= hi) return;  // wave-uniform
  unsigned int cnt[kMaxTiers];
#pragma unroll
  for (int k = 0; k < kMaxTiers; ++k) cnt[k] = 0;
  for (int64_t b = lo; b < hi; b += kWave) {
    const int64_t g = b + lane;
    const int tier = g < hi ? tier_of(node_m[open_idx[list ? list[g] : (uint32_t)g]], cuts) : -1;
#pragma unroll
    for (int k = 0; k < kMaxTiers; ++k)
      if (k < cuts.n) cnt[k] += (unsigned int)__popcll(__ballot(tier == k));
  }
  unsigned int base[kMaxTiers];
#pragma unroll
  for (int k = 0; k < kMaxTiers; ++k) {
    unsigned int b0 = 0;
    if (lane == 0 && k < cuts.n && cnt[k]) b0 = atomicAdd(&counts[k], cnt[k]);
    base[k] = __shfl(b0, 0, kWave);
  }
  const uint64_t below = (1ull << lane) - 1ull;
  for (int64_t b = lo; b < hi; b += kWave) {
    const int64_t g = b + lane;
    uint32_t o = 0;
    int tier = -1;
    if (g < hi) {
      o = list ? list[g] : (uint32_t)g;
      tier = tier_of(node_m[open_idx[o]], cuts);
    }
#pragma unroll
    for (int k = 0; k < kMaxTiers; ++k) {
      if (k >= cuts.n) break;
      const uint64_t mask = __ballot(tier == k);
      if (tier == k) out.l[k][base[k] + (unsigned int)__popcll(mask & below)] = o;
      base[k] += (unsigned int)__popcll(mask);
    }
  }
}

// ------------------------------------------------------ decide ----------
__device__ __forceinline__ bool split_better(float sa, int fa, float ta, float sb, int fb, float tb) {
  if (sa != sb) return sa > sb;
  if (fa != fb) return fa < fb;
  return ta < tb;
}

// Phase 1: best of the mandatory draws (T1 ties); decide whether the
// retry draws (gpurf.cu:177-182, 242) are needed.  Pure classification
// nodes score 0 on every draw, so their result is the lowest feature index
// among all draws, without evaluating the retries.
struct DecideArgs {
  int64_t n_open;
  int D0, R, D;
  const float* res_score;
  const float* res_thr;
  const int32_t* res_feat;
  float* ch_score;
  float* ch_thr;
  int32_t* ch_feat;
  uint32_t* retry_flag;
  const uint32_t* retry_list;
  int64_t n_retry;
  const unsigned int* n_open_dev;   // device-described level
  const unsigned int* n_retry_dev;
};

__global__ void k_decide1(DataArgs d, LevelArgs lv, DecideArgs a, int algorithm) {
  resolve_level(lv, d.C);
  if (a.n_open_dev) a.n_open = *a.n_open_dev;
  for (int64_t o = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; o < a.n_open;
       o += (int64_t)gridDim.x * blockDim.x) {
    const float* rs = a.res_score + o * a.D;
    const float* rt = a.res_thr + o * a.D;
    const int32_t* rf = a.res_feat + o * a.D;
    int D0 = a.D0, R = a.R;
    if (d.forests) {  // joint build: the node's own forest's draws
      const DataArgs dd = forest_data(d, lv.tree[lv.open_idx[o]]);
      D0 = dd.D0;
      R = dd.R;
    }
    float bs = rs[0], bt = rt[0];
    int bf = rf[0];
    for (int j = 1; j < D0; ++j)
      if (split_better(rs[j], rf[j], rt[j], bs, bf, bt)) {
        bs = rs[j];
        bf = rf[j];
        bt = rt[j];
      }
    uint32_t need = 0;
    if (R > 0 && (double)bs <= 1e-3) {
      const uint32_t i = lv.open_idx[o];
      bool pure = false;
      if (algorithm == 0 && d.C > 0 && lv.cls) {
        uint32_t mx = 0;
        for (int c = 0; c < d.C; ++c) mx = max(mx, lv.cls[(int64_t)i * d.C + c]);
        pure = (int64_t)mx == (int64_t)lv.count[i];
      }
      if (pure) {
        const uint32_t tg = (uint32_t)lv.tree_global[lv.tree[i]];
        for (int r = 0; r < R; ++r) {
          const int f = (int)(rand_u32(d.seed, kStreamFeat, tg, lv.id[i], (uint64_t)(D0 + r)) % (uint32_t)d.F);
          if (split_better(0.0f, f, 0.0f, bs, bf, bt)) {
            bs = 0.0f;
            bf = f;
            bt = 0.0f;
          }
        }
      } else {
        need = 1;
      }
    }
    a.ch_score[o] = bs;
    a.ch_thr[o] = bt;
    a.ch_feat[o] = bf;
    a.retry_flag[o] = need;
  }
}

// Phase 2: continue the sequential retry rule on the nodes that needed it.
__global__ void k_decide2(DataArgs d, LevelArgs lv, DecideArgs a) {
  resolve_level(lv, d.C);
  if (a.n_retry_dev) a.n_retry = *a.n_retry_dev;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < a.n_retry;
       q += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t o = a.retry_list[q];
    float bs = a.ch_score[o], bt = a.ch_thr[o];
    int bf = a.ch_feat[o];
    const float* rs = a.res_score + (int64_t)o * a.D;
    const float* rt = a.res_thr + (int64_t)o * a.D;
    const int32_t* rf = a.res_feat + (int64_t)o * a.D;
    int D0 = a.D0, R = a.R;
    if (d.forests) {  // joint build: the node's own forest's draws
      const DataArgs dd = forest_data(d, lv.tree[lv.open_idx[o]]);
      D0 = dd.D0;
      R = dd.R;
    }
    for (int r = 0; r < R; ++r) {
      if ((double)bs > 1e-3) break;
      const int j = D0 + r;
      if (split_better(rs[j], rf[j], rt[j], bs, bf, bt)) {
        bs = rs[j];
        bf = rf[j];
        bt = rt[j];
      }
    }
    a.ch_score[o] = bs;
    a.ch_thr[o] = bt;
    a.ch_feat[o] = bf;
  }
}

// ------------------------------------------------------ partition -------
// One block per open node: count the left side of the chosen split, then a
// stable two-way compaction into the next entry buffer (FIX of the
// reference's atomic slot order, gpudte.cu:205-232).  Children occupy the
// parent's range: left [start, start+mL), right [start+mL, start+m).
struct PartArgs {
  int64_t n_open;
  const int32_t* ch_feat;
  const float* ch_thr;
  uint32_t* nrow;
  uint32_t* naux;
  int64_t* ney;
  // left child statistics per open node
  int32_t* l_m;
  int32_t* l_count;
  uint32_t* l_cls;
  int64_t* l_sum;
  uint64_t* l_hash;
  uint8_t* split_ok;
  unsigned long long* bytes;
  const uint32_t* list;  // optional: which open positions to process
  const unsigned int* n_dev;  // device-described level: the launch's node count
};

// Side bits of the first pass kept for the scatter pass (nodes of up to
// 64 Ki entries), which then reads no X.
constexpr int kSideWords = 1024;

template <int MAXC, int TASK>
__device__ __forceinline__ void part_block_body(DataArgs d, LevelArgs lv, PartArgs pa, int64_t bid, int64_t nblk) {
  resolve_level(lv, d.C);
  if (pa.n_dev) pa.n_open = *pa.n_dev;
  if ((int64_t)bid >= pa.n_open) return;
  __shared__ unsigned long long s_side[kSideWords];
  __shared__ uint32_t s_w[kWaves];
  __shared__ uint32_t s_cls[MAXC];
  __shared__ unsigned long long s_cnt, s_sum, s_hash;
  __shared__ uint32_t s_ml;
  const int t = threadIdx.x;
  unsigned long long my_bytes = 0;
  for (int64_t q = bid; q < pa.n_open; q += nblk) {
    const int64_t o = pa.list ? pa.list[q] : q;
    const uint32_t i = lv.open_idx[o];
    const int m = lv.m[i];
    const int64_t st = lv.start[i];
    const int f = pa.ch_feat[o];
    const float thr = pa.ch_thr[o];
    const DataArgs dd = forest_data(d, lv.tree[i]);
    const float* Xf = dd.X + (int64_t)f * dd.N;
    if (t < MAXC) s_cls[t] = 0;
    if (t == 0) {
      s_cnt = 0;
      s_sum = 0;
      s_hash = 0;
      s_ml = 0;
    }
    __syncthreads();
    uint32_t loc[MAXC];
#pragma unroll
    for (int c = 0; c < MAXC; ++c) loc[c] = 0;
    uint64_t cnt = 0, hsh = 0;
    int64_t sm = 0;
    uint32_t ml = 0;
    const bool keep = m <= kSideWords * kWave;
    for (int e = t; e < m; e += kBlock) {
      const uint32_t row = lv.row[st + e];
      const bool isl = side_of(Xf[row], thr) == 0;
      if (keep) {
        const unsigned long long bits = __ballot(isl);
        if ((t & (kWave - 1)) == 0) s_side[e / kWave] = bits;
      }
      if (isl) {
        const uint32_t a = lv.aux[st + e];
        const uint32_t w = a & 0xFFFFu;
        ++ml;
        cnt += w;
        hsh += (uint64_t)w * mix64(row);
        if (TASK == 0) {
          const int lab = (int)(a >> 16);
#pragma unroll
          for (int c = 0; c < MAXC; ++c) loc[c] += (c == lab) ? w : 0u;
        } else {
          sm += (int64_t)w * lv.ey[st + e];
        }
      }
    }
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
      cnt += __shfl_xor(cnt, off, kWave);
      hsh += __shfl_xor(hsh, off, kWave);
      sm += __shfl_xor(sm, off, kWave);
      ml += __shfl_xor(ml, off, kWave);
    }
    if (TASK == 0) {
#pragma unroll
      for (int c = 0; c < MAXC; ++c) {
        uint32_t v = loc[c];
#pragma unroll
        for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
        if ((t & (kWave - 1)) == 0 && v) atomicAdd(&s_cls[c], v);
      }
    }
    if ((t & (kWave - 1)) == 0) {
      atomicAdd(&s_cnt, (unsigned long long)cnt);
      atomicAdd(&s_hash, (unsigned long long)hsh);
      atomicAdd(&s_sum, (unsigned long long)sm);
      atomicAdd(&s_ml, ml);
    }
    __syncthreads();
    const int mL = (int)s_ml;
    const bool ok = mL > 0 && mL < m;
    if (t == 0) {
      my_bytes += (unsigned long long)m * (TASK == 0 ? 12ull : 20ull);
      pa.split_ok[o] = ok ? 1 : 0;
      pa.l_m[o] = mL;
      pa.l_count[o] = (int32_t)s_cnt;
      pa.l_hash[o] = (uint64_t)s_hash;
      if (TASK == 1) pa.l_sum[o] = (int64_t)s_sum;
    }
    if (TASK == 0 && t < d.C) pa.l_cls[o * d.C + t] = s_cls[t];
    if (ok) {
      // Stable scatter, tile by tile.
      uint32_t lbase = 0, rbase = (uint32_t)mL;
      for (int e0 = 0; e0 < m; e0 += kBlock) {
        const int e = e0 + t;
        uint32_t row = 0, isl = 0;
        if (e < m) {
          row = lv.row[st + e];
          isl = keep ? (uint32_t)(s_side[e / kWave] >> (e & (kWave - 1))) & 1u
                     : (side_of(Xf[row], thr) == 0 ? 1u : 0u);
        }
        uint32_t tot;
        const uint32_t pl = block_excl_scan(isl, s_w, &tot);
        if (e < m) {
          const uint32_t pos = isl ? lbase + pl : rbase + (uint32_t)(e - e0) - pl;
          pa.nrow[st + pos] = row;
          pa.naux[st + pos] = lv.aux[st + e];
          if (TASK == 1) pa.ney[st + pos] = lv.ey[st + e];
        }
        lbase += tot;
        rbase += (uint32_t)min(kBlock, m - e0) - tot;
      }
    }
    __syncthreads();
  }
  if (t == 0 && my_bytes && pa.bytes) atomicAdd(pa.bytes, my_bytes);
}
template <int MAXC, int TASK>
__global__ void __launch_bounds__(kBlock) k_partition(DataArgs d, LevelArgs lv, PartArgs pa) {
  part_block_body<MAXC, TASK>(d, lv, pa, blockIdx.x, gridDim.x);
}

// Nodes above kPartBig entries (the top levels, where a workgroup per node
// leaves most CUs idle): the same stable partition over chunks of
// kPartChunk entries, each chunk a workgroup.
//   1. k_part_chunk_stats    per chunk: left count, left statistics
//                            (atomics into the node's accumulators)
//   2. k_part_chunk_finish   per node: the node's results, left offset of
//                            every chunk (exclusive scan of the chunks'
//                            left counts)
//   3. k_part_chunk_scatter  per chunk: stable two-way compaction at the
//                            chunk's offsets (lefts before it on the left,
//                            rights before it after the node's mL lefts)
// Sums of integers, so the statistics equal k_partition's bit for bit, and
// entries land where k_partition puts them (a stable compaction in place of
// the reference's atomic slot order, gpudte.cu:205-232).
constexpr int kPartBig = 32768;
constexpr int kPartChunk = 8192;

struct PartChunks {
  const uint32_t* list;  // big-node open positions
  int64_t n_nodes;
  const uint32_t* coff;  // first chunk of each big node (n_nodes + 1)
  uint32_t* chunk_ml;    // left entries of each chunk, then their exclusive scan per node
  uint32_t* acc_ml;      // per big node
  unsigned long long* acc_cnt;
  unsigned long long* acc_hash;
  unsigned long long* acc_sum;
  uint32_t* acc_cls;     // [node * C + c]
};

__global__ void k_part_nchunks(const uint32_t* list, int64_t n, const uint32_t* open_idx, const int32_t* node_m,
                               uint32_t* nch) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
    nch[q] = (uint32_t)((node_m[open_idx[list[q]]] + kPartChunk - 1) / kPartChunk);
}

// Big node of chunk k: the last q with coff[q] <= k.
__device__ __forceinline__ int64_t part_chunk_node(const PartChunks& pc, int64_t k) {
  int64_t lo = 0, hi = pc.n_nodes - 1;
  while (lo < hi) {
    const int64_t mid = (lo + hi + 1) >> 1;
    if ((int64_t)pc.coff[mid] <= k) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <int MAXC, int TASK>
__global__ void __launch_bounds__(kBlock) k_part_chunk_stats(DataArgs d, LevelArgs lv, PartArgs pa, PartChunks pc) {
  __shared__ uint32_t s_cls[MAXC];
  __shared__ unsigned long long s_cnt, s_sum, s_hash;
  __shared__ uint32_t s_ml;
  const int t = threadIdx.x;
  const int64_t k = blockIdx.x;
  const int64_t q = part_chunk_node(pc, k);
  const uint32_t o = pc.list[q];
  const uint32_t i = lv.open_idx[o];
  const int m = lv.m[i];
  const int e0 = (int)(k - pc.coff[q]) * kPartChunk;
  const int e1 = min(m, e0 + kPartChunk);
  const int64_t st = lv.start[i];
  const DataArgs dd = forest_data(d, lv.tree[i]);
  const float* Xf = dd.X + (int64_t)pa.ch_feat[o] * dd.N;
  const float thr = pa.ch_thr[o];
  if (t < MAXC) s_cls[t] = 0;
  if (t == 0) {
    s_cnt = 0;
    s_sum = 0;
    s_hash = 0;
    s_ml = 0;
  }
  __syncthreads();
  uint32_t loc[MAXC];
#pragma unroll
  for (int c = 0; c < MAXC; ++c) loc[c] = 0;
  uint64_t cnt = 0, hsh = 0;
  int64_t sm = 0;
  uint32_t ml = 0;
  for (int e = e0 + t; e < e1; e += kBlock) {
    const uint32_t row = lv.row[st + e];
    const uint32_t a = lv.aux[st + e];
    if (side_of(Xf[row], thr) == 0) {
      const uint32_t w = a & 0xFFFFu;
      ++ml;
      cnt += w;
      hsh += (uint64_t)w * mix64(row);
      if (TASK == 0) {
        const int lab = (int)(a >> 16);
#pragma unroll
        for (int c = 0; c < MAXC; ++c) loc[c] += (c == lab) ? w : 0u;
      } else {
        sm += (int64_t)w * lv.ey[st + e];
      }
    }
  }
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) {
    cnt += __shfl_xor(cnt, off, kWave);
    hsh += __shfl_xor(hsh, off, kWave);
    sm += __shfl_xor(sm, off, kWave);
    ml += __shfl_xor(ml, off, kWave);
  }
  if (TASK == 0) {
#pragma unroll
    for (int c = 0; c < MAXC; ++c) {
      uint32_t v = loc[c];
#pragma unroll
      for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
      if ((t & (kWave - 1)) == 0 && v) atomicAdd(&s_cls[c], v);
    }
  }
  if ((t & (kWave - 1)) == 0) {
    atomicAdd(&s_cnt, (unsigned long long)cnt);
    atomicAdd(&s_hash, (unsigned long long)hsh);
    atomicAdd(&s_sum, (unsigned long long)sm);
    atomicAdd(&s_ml, ml);
  }
  __syncthreads();
  if (t == 0) {
    pc.chunk_ml[k] = s_ml;
    if (s_ml) atomicAdd(&pc.acc_ml[q], s_ml);
    if (s_cnt) atomicAdd(&pc.acc_cnt[q], s_cnt);
    if (s_hash) atomicAdd(&pc.acc_hash[q], s_hash);
    if (TASK == 1 && s_sum) atomicAdd(&pc.acc_sum[q], s_sum);
  }
  if (TASK == 0 && t < d.C && s_cls[t]) atomicAdd(&pc.acc_cls[q * d.C + t], s_cls[t]);
}

// One workgroup per big node.
template <int TASK>
__global__ void __launch_bounds__(kBlock) k_part_chunk_finish(DataArgs d, LevelArgs lv, PartArgs pa, PartChunks pc) {
  __shared__ uint32_t s_w[kWaves];
  const int64_t q = blockIdx.x;
  const int t = threadIdx.x;
  const uint32_t o = pc.list[q];
  const uint32_t i = lv.open_idx[o];
  const int m = lv.m[i];
  const int mL = (int)pc.acc_ml[q];
  const bool ok = mL > 0 && mL < m;
  if (t == 0) {
    pa.split_ok[o] = ok ? 1 : 0;
    pa.l_m[o] = mL;
    pa.l_count[o] = (int32_t)pc.acc_cnt[q];
    pa.l_hash[o] = (uint64_t)pc.acc_hash[q];
    if (TASK == 1) pa.l_sum[o] = (int64_t)pc.acc_sum[q];
  }
  if (TASK == 0 && t < d.C) pa.l_cls[(int64_t)o * d.C + t] = pc.acc_cls[q * d.C + t];
  // exclusive scan of the chunks' left counts, in place
  const uint32_t c0 = pc.coff[q], c1 = pc.coff[q + 1];
  uint32_t base = 0;
  for (uint32_t k0 = c0; k0 < c1; k0 += kBlock) {
    const uint32_t k = k0 + t;
    const uint32_t v = k < c1 ? pc.chunk_ml[k] : 0u;
    uint32_t tot;
    const uint32_t ex = block_excl_scan(v, s_w, &tot);
    if (k < c1) pc.chunk_ml[k] = base + ex;
    base += tot;
    __syncthreads();
  }
  if (t == 0 && pa.bytes) atomicAdd(pa.bytes, (unsigned long long)m * (TASK == 0 ? 12ull : 20ull));
}

template <int TASK>
__global__ void __launch_bounds__(kBlock) k_part_chunk_scatter(DataArgs d, LevelArgs lv, PartArgs pa, PartChunks pc) {
  __shared__ uint32_t s_w[kWaves];
  const int t = threadIdx.x;
  const int64_t k = blockIdx.x;
  const int64_t q = part_chunk_node(pc, k);
  const uint32_t o = pc.list[q];
  if (!pa.split_ok[o]) return;  // workgroup-uniform
  const uint32_t i = lv.open_idx[o];
  const int m = lv.m[i];
  const int mL = (int)pc.acc_ml[q];
  const int e0 = (int)(k - pc.coff[q]) * kPartChunk;
  const int e1 = min(m, e0 + kPartChunk);
  const int64_t st = lv.start[i];
  const DataArgs dd = forest_data(d, lv.tree[i]);
  const float* Xf = dd.X + (int64_t)pa.ch_feat[o] * dd.N;
  const float thr = pa.ch_thr[o];
  uint32_t lbase = pc.chunk_ml[k];                           // lefts before this chunk
  uint32_t rbase = (uint32_t)mL + (uint32_t)e0 - lbase;      // rights before it follow every left
  for (int b = e0; b < e1; b += kBlock) {
    const int e = b + t;
    uint32_t row = 0, isl = 0;
    if (e < e1) {
      row = lv.row[st + e];
      isl = side_of(Xf[row], thr) == 0 ? 1u : 0u;
    }
    uint32_t tot;
    const uint32_t pl = block_excl_scan(isl, s_w, &tot);
    if (e < e1) {
      const uint32_t pos = isl ? lbase + pl : rbase + (uint32_t)(e - b) - pl;
      pa.nrow[st + pos] = row;
      pa.naux[st + pos] = lv.aux[st + e];
      if (TASK == 1) pa.ney[st + pos] = lv.ey[st + e];
    }
    lbase += tot;
    rbase += (uint32_t)min(kBlock, e1 - b) - tot;
  }
}

// Small nodes: one wave per node; stable compaction with ballots, tile by
// tile (64 entries), statistics reduced across the wave.
template <int MAXC, int TASK>
__device__ __forceinline__ void part_wave_body(DataArgs d, LevelArgs lv, PartArgs pa, const uint32_t* list,
                                               int64_t n_list, int64_t bid, int64_t nblk) {
  resolve_level(lv, d.C);
  if (pa.n_dev) n_list = *pa.n_dev;
  const int lane = threadIdx.x & (kWave - 1);
  const uint64_t lt_mask = (1ull << lane) - 1ull;
  unsigned long long my_bytes = 0;
  for (int64_t q = (int64_t)bid * kWaves + threadIdx.x / kWave; q < n_list;
       q += (int64_t)nblk * kWaves) {
    const uint32_t o = list[q];
    const uint32_t i = lv.open_idx[o];
    const int m = lv.m[i];
    const int64_t st = lv.start[i];
    const float thr = pa.ch_thr[o];
    const DataArgs dd = forest_data(d, lv.tree[i]);
    const float* Xf = dd.X + (int64_t)pa.ch_feat[o] * dd.N;
    uint32_t loc[MAXC];
#pragma unroll
    for (int c = 0; c < MAXC; ++c) loc[c] = 0;
    uint64_t cnt = 0, hsh = 0;
    int64_t sm = 0;
    uint32_t ml = 0;
    uint32_t side = 0;  // bit k: entry k * 64 + lane goes left (nodes of up to 2048 entries)
    const bool keep = m <= 32 * kWave;
    for (int e0 = 0; e0 < m; e0 += kWave) {
      const int e = e0 + lane;
      bool isl = false;
      if (e < m) {
        const uint32_t row = lv.row[st + e];
        isl = side_of(Xf[row], thr) == 0;
        if (isl && keep) side |= 1u << (e0 / kWave);
        if (isl) {
          const uint32_t a = lv.aux[st + e];
          const uint32_t w = a & 0xFFFFu;
          cnt += w;
          hsh += (uint64_t)w * mix64(row);
          if (TASK == 0) {
            const int lab = (int)(a >> 16);
#pragma unroll
            for (int c = 0; c < MAXC; ++c) loc[c] += (c == lab) ? w : 0u;
          } else {
            sm += (int64_t)w * lv.ey[st + e];
          }
        }
      }
      ml += (uint32_t)__popcll(__ballot(isl));
    }
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
      cnt += __shfl_xor(cnt, off, kWave);
      hsh += __shfl_xor(hsh, off, kWave);
      sm += __shfl_xor(sm, off, kWave);
    }
    const int mL = (int)ml;
    const bool ok = mL > 0 && mL < m;
    if (TASK == 0) {
#pragma unroll
      for (int c = 0; c < MAXC; ++c) {
        uint32_t v = loc[c];
#pragma unroll
        for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
        if (lane == c && c < d.C) pa.l_cls[(int64_t)o * d.C + c] = v;
      }
    }
    if (lane == 0) {
      my_bytes += (unsigned long long)m * (TASK == 0 ? 12ull : 20ull);
      pa.split_ok[o] = ok ? 1 : 0;
      pa.l_m[o] = mL;
      pa.l_count[o] = (int32_t)cnt;
      pa.l_hash[o] = hsh;
      if (TASK == 1) pa.l_sum[o] = sm;
    }
    if (ok) {
      uint32_t lbase = 0, rbase = (uint32_t)mL;
      for (int e0 = 0; e0 < m; e0 += kWave) {
        const int e = e0 + lane;
        uint32_t row = 0;
        bool isl = false;
        if (e < m) {
          row = lv.row[st + e];
          isl = keep ? ((side >> (e0 / kWave)) & 1u) != 0 : side_of(Xf[row], thr) == 0;
        }
        const uint64_t bl = __ballot(isl);
        const uint64_t br = __ballot(e < m && !isl);
        if (e < m) {
          const uint32_t pos = isl ? lbase + (uint32_t)__popcll(bl & lt_mask) : rbase + (uint32_t)__popcll(br & lt_mask);
          pa.nrow[st + pos] = row;
          pa.naux[st + pos] = lv.aux[st + e];
          if (TASK == 1) pa.ney[st + pos] = lv.ey[st + e];
        }
        lbase += (uint32_t)__popcll(bl);
        rbase += (uint32_t)__popcll(br);
      }
    }
  }
  if (lane == 0 && my_bytes && pa.bytes) atomicAdd(pa.bytes, my_bytes);
}
template <int MAXC, int TASK>
__global__ void __launch_bounds__(kBlock) k_partition_wave(DataArgs d, LevelArgs lv, PartArgs pa,
                                                           const uint32_t* list, int64_t n_list) {
  part_wave_body<MAXC, TASK>(d, lv, pa, list, n_list, blockIdx.x, gridDim.x);
}

// Nodes of at most SEG entries: one SEG-lane segment per node, one entry
// per lane, a single pass (side, left statistics, stable positions from the
// segment's ballots).  Same outputs as k_partition_wave.
template <int MAXC, int TASK, int SEG>
__device__ __forceinline__ void part_seg_body(DataArgs d, LevelArgs lv, PartArgs pa, const uint32_t* list,
                                              int64_t n_list, int64_t bid, int64_t nblk) {
  constexpr int NPW = kWave / SEG;
  resolve_level(lv, d.C);
  if (pa.n_dev) n_list = *pa.n_dev;
  __shared__ uint32_t s_cls[kWaves][NPW][TASK == 0 ? MAXC : 1];
  const int lane = threadIdx.x & (kWave - 1);
  const int sl = lane & (SEG - 1);
  const int seg = lane / SEG;
  const int wv = threadIdx.x / kWave;
  const uint64_t seg_mask = (SEG == kWave ? ~0ull : ((1ull << SEG) - 1ull)) << (seg * SEG);
  const uint64_t below = (1ull << lane) - 1ull;
  unsigned long long my_bytes = 0;
  const int64_t wave_id = (int64_t)bid * kWaves + wv;
  const int64_t n_waves = (int64_t)nblk * kWaves;
  for (int64_t q0 = wave_id * NPW; q0 < n_list; q0 += n_waves * NPW) {
    const int64_t q = q0 + seg;
    const bool node_ok = q < n_list;
    uint32_t o = 0;
    int m = 0;
    int64_t st = 0;
    float thr = 0.0f;
    const float* Xf = d.X;
    if (node_ok) {
      o = list[q];
      const uint32_t i = lv.open_idx[o];
      m = lv.m[i];
      st = lv.start[i];
      thr = pa.ch_thr[o];
      if (d.forests) {
        const DataArgs dd = forest_data(d, lv.tree[i]);
        Xf = dd.X + (int64_t)pa.ch_feat[o] * dd.N;
      } else {
        Xf = d.X + (int64_t)pa.ch_feat[o] * d.N;
      }
    }
    const bool valid = node_ok && sl < m;
    uint32_t row = 0, a = 0;
    int64_t ey = 0;
    bool isl = false;
    if (valid) {
      row = lv.row[st + sl];
      a = lv.aux[st + sl];
      if (TASK == 1) ey = lv.ey[st + sl];
      isl = side_of(Xf[row], thr) == 0;
    }
    const uint64_t bl = __ballot(isl) & seg_mask;
    const uint64_t bv = __ballot(valid) & seg_mask;
    const int mL = __popcll(bl);
    const bool ok = mL > 0 && mL < m;
    const uint32_t w = isl ? (a & 0xFFFFu) : 0u;
    const int64_t cnt = seg_sum_i64<SEG>((int64_t)w);
    const int64_t hsh = seg_sum_i64<SEG>(isl ? (int64_t)((uint64_t)w * mix64(row)) : 0);
    const int64_t sm = TASK == 1 ? seg_sum_i64<SEG>(isl ? (int64_t)w * ey : 0) : 0;
    if constexpr (TASK == 0) {
      uint32_t* sc = s_cls[wv][seg];
      for (int c = sl; c < MAXC; c += SEG) sc[c] = 0u;
      team_sync<kWave>();
      if (isl) atomicAdd(&sc[a >> 16], w);
      team_sync<kWave>();
      if (node_ok)
        for (int c = sl; c < d.C; c += SEG) pa.l_cls[(int64_t)o * d.C + c] = sc[c];
      team_sync<kWave>();
    }
    if (node_ok && sl == 0) {
      my_bytes += (unsigned long long)m * (TASK == 0 ? 12ull : 20ull);
      pa.split_ok[o] = ok ? 1 : 0;
      pa.l_m[o] = mL;
      pa.l_count[o] = (int32_t)cnt;
      pa.l_hash[o] = (uint64_t)hsh;
      if (TASK == 1) pa.l_sum[o] = sm;
    }
    if (ok && valid) {
      const uint32_t pos = isl ? (uint32_t)__popcll(bl & below)
                               : (uint32_t)mL + (uint32_t)__popcll(bv & ~bl & below);
      pa.nrow[st + pos] = row;
      pa.naux[st + pos] = a;
      if (TASK == 1) pa.ney[st + pos] = ey;
    }
  }
  if (sl == 0 && my_bytes && pa.bytes) atomicAdd(pa.bytes, my_bytes);
}
template <int MAXC, int TASK, int SEG>
__global__ void __launch_bounds__(kBlock) k_partition_seg(DataArgs d, LevelArgs lv, PartArgs pa,
                                                          const uint32_t* list, int64_t n_list) {
  part_seg_body<MAXC, TASK, SEG>(d, lv, pa, list, n_list, blockIdx.x, gridDim.x);
}

// All partition tiers of a level step in one launch (small data, see
// k_search_ert_all): workgroups [first[k], first[k + 1]) run tier k
// (16-lane segments, 32-lane segments, a wave, a workgroup per node).
struct PartTiers {
  int64_t first[5];
  const uint32_t* list[4];
  const unsigned int* n_dev[4];
  int64_t n_bound;  // list capacity (the device counts hold the sizes)
};

template <int MAXC, int TASK>
__global__ void __launch_bounds__(kBlock) k_partition_all(DataArgs d, LevelArgs lv, PartArgs pa, PartTiers pt) {
  const int64_t b = blockIdx.x;
  int k = 0;
  while (k < 3 && b >= pt.first[k + 1]) ++k;
  pa.n_dev = pt.n_dev[k];
  const int64_t bid = b - pt.first[k], nblk = pt.first[k + 1] - pt.first[k];
  switch (k) {
    case 0: part_seg_body<MAXC, TASK, 16>(d, lv, pa, pt.list[0], pt.n_bound, bid, nblk); break;
    case 1: part_seg_body<MAXC, TASK, 32>(d, lv, pa, pt.list[1], pt.n_bound, bid, nblk); break;
    case 2: part_wave_body<MAXC, TASK>(d, lv, pa, pt.list[2], pt.n_bound, bid, nblk); break;
    default:
      pa.list = pt.list[3];
      part_block_body<MAXC, TASK>(d, lv, pa, bid, nblk);
      break;
  }
}

template <int MAXC, int TASK>
void launch_partition_all(int64_t grid, hipStream_t s, const DataArgs& d, const LevelArgs& lv, const PartArgs& pa,
                          const PartTiers& pt) {
  hipLaunchKernelGGL((k_partition_all<MAXC, TASK>), dim3((unsigned)grid), dim3(kBlock), 0, s, d, lv, pa, pt);
}
using PartAllFn = void (*)(int64_t, hipStream_t, const DataArgs&, const LevelArgs&, const PartArgs&, const PartTiers&);
PartAllFn pick_part_all(int task, int C) {
  if (task == 1) return launch_partition_all<1, 1>;
  if (C <= 2) return launch_partition_all<2, 0>;
  if (C <= 4) return launch_partition_all<4, 0>;
  if (C <= 8) return launch_partition_all<8, 0>;
  if (C <= 16) return launch_partition_all<16, 0>;
  return launch_partition_all<32, 0>;
}

// ------------------------------------------------------ children --------
struct ChildArgs {
  int64_t n_level;         // nodes at this level
  int64_t off_level;       // level-major offset of this level
  int64_t off_next;        // level-major offset of the next level
  int depth;
  const uint8_t* open;     // per level node
  const uint32_t* open_pos;  // level index -> open position (valid if open)
  const uint8_t* split_ok;   // per open position
  const int32_t* ch_feat;
  const float* ch_thr;
  const float* ch_score;
  const int32_t* l_m;
  const int32_t* l_count;
  const uint32_t* l_cls;
  const int64_t* l_sum;
  const uint64_t* l_hash;
  uint32_t* nchild;        // per level node (scan input)
  const uint32_t* cpos;    // exclusive scan of nchild (n_level + 1)
  const uint32_t* lf;      // first level index of each tree (n_trees + 1)
  const uint32_t* base;    // tree-local id of each tree's first node at this level
  // output arrays (level-major)
  int32_t* o_tree;
  uint32_t* o_id;
  int32_t* o_depth;
  int32_t* o_count;
  int32_t* o_m;
  uint32_t* o_cls;
  int64_t* o_sum;
  uint64_t* o_hash;
  float* o_value;
  int32_t* o_feature;
  float* o_thr;
  float* o_score;
  int64_t* o_left;
  // next level working arrays
  int64_t* n_start;
  uint8_t* n_open;
  const int64_t* start;    // current level starts
  const DevLevel* lvl;     // device-described level: n_level, offsets, depth
};

__device__ __forceinline__ void resolve_children(ChildArgs& a) {
  if (a.lvl) {
    a.n_level = a.lvl->n;
    a.off_level = a.lvl->off;
    a.off_next = a.off_level + a.n_level;
    a.depth = (int)a.lvl->depth;
  }
}

__global__ void k_children_count(ChildArgs a) {
  resolve_children(a);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < a.n_level;
       i += (int64_t)gridDim.x * blockDim.x) {
    uint32_t nc = 0;
    if (a.open[i]) {
      const uint32_t o = a.open_pos[i];
      nc = a.split_ok[o] ? 2u : 0u;
    }
    a.nchild[i] = nc;
  }
}

__global__ void k_children_write(DataArgs d, ChildArgs a) {
  resolve_children(a);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < a.n_level;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t gi = a.off_level + i;
    if (a.nchild[i] == 0) {
      a.o_feature[gi] = -1;
      a.o_thr[gi] = -1.0f;
      a.o_score[gi] = 0.0f;
      a.o_left[gi] = -1;
      continue;
    }
    const uint32_t o = a.open_pos[i];
    const int tr = a.o_tree[gi];
    const DataArgs dd = forest_data(d, tr);
    a.o_feature[gi] = a.ch_feat[o];
    a.o_thr[gi] = a.ch_thr[o];
    a.o_score[gi] = a.ch_score[o];
    const uint32_t p = a.cpos[i];
    const int64_t gl = a.off_next + p;
    a.o_left[gi] = gl;
    const uint32_t first = a.cpos[a.lf[tr]];
    const uint32_t idl = a.base[tr] + (a.lf[tr + 1] - a.lf[tr]) + (p - first);
    const int m = a.o_m[gi];
    const int mL = a.l_m[o];
    const int cnt = a.o_count[gi];
    const int cL = a.l_count[o];
    const int64_t st = a.start[i];
    for (int s = 0; s < 2; ++s) {
      const int64_t g = gl + s;
      a.o_tree[g] = tr;
      a.o_id[g] = idl + s;
      a.o_depth[g] = a.depth + 1;
      const int cm = s == 0 ? mL : m - mL;
      const int cc = s == 0 ? cL : cnt - cL;
      a.o_m[g] = cm;
      a.o_count[g] = cc;
      a.o_hash[g] = s == 0 ? a.l_hash[o] : a.o_hash[gi] - a.l_hash[o];
      if (d.C > 0 && a.o_cls) {
        for (int c = 0; c < d.C; ++c) {
          const uint32_t lc = a.l_cls[(int64_t)o * d.C + c];
          a.o_cls[g * d.C + c] = s == 0 ? lc : a.o_cls[gi * d.C + c] - lc;
        }
        a.o_value[g] = 0.0f;
      } else {
        const int64_t sm = s == 0 ? a.l_sum[o] : a.o_sum[gi] - a.l_sum[o];
        a.o_sum[g] = sm;
        a.o_value[g] = (float)ldexp_d((double)sm / (double)cc, -dd.shift);
      }
      a.n_start[p + s] = s == 0 ? st : st + mL;
      a.n_open[p + s] = (cc > dd.min_node && a.depth + 1 < dd.max_depth) ? 1 : 0;
    }
  }
}

// Per-tree bookkeeping for the next level: first index and base id.
__global__ void k_tree_update(int n_trees, const uint32_t* cpos, const uint32_t* lf, const uint32_t* base,
                              uint32_t* lf_next, uint32_t* base_next) {
  const int tr = blockIdx.x * blockDim.x + threadIdx.x;
  if (tr > n_trees) return;
  lf_next[tr] = cpos[lf[tr]];
  if (tr < n_trees) base_next[tr] = base[tr] + (lf[tr + 1] - lf[tr]);
}

// ------------------------------------------------------ scans -----------
constexpr int kScanTile = kBlock * 16;

__global__ void k_scan_tiles(const uint32_t* in, int64_t n, uint32_t* tile_sums) {
  __shared__ uint32_t s_w[kWaves];
  const int64_t t0 = (int64_t)blockIdx.x * kScanTile;
  uint32_t s = 0;
  for (int k = 0; k < 16; ++k) {
    const int64_t idx = t0 + (int64_t)threadIdx.x * 16 + k;
    if (idx < n) s += in[idx];
  }
  uint32_t tot;
  block_excl_scan(s, s_w, &tot);
  if (threadIdx.x == 0) tile_sums[blockIdx.x] = tot;
}

__global__ void k_scan_apply(const uint32_t* in, int64_t n, const uint32_t* tile_off, uint32_t* out) {
  __shared__ uint32_t s_w[kWaves];
  const int64_t t0 = (int64_t)blockIdx.x * kScanTile;
  uint32_t v[16];
  uint32_t s = 0;
  for (int k = 0; k < 16; ++k) {
    const int64_t idx = t0 + (int64_t)threadIdx.x * 16 + k;
    v[k] = idx < n ? in[idx] : 0u;
    s += v[k];
  }
  uint32_t base = block_excl_scan(s, s_w, nullptr) + tile_off[blockIdx.x];
  for (int k = 0; k < 16; ++k) {
    const int64_t idx = t0 + (int64_t)threadIdx.x * 16 + k;
    if (idx < n) out[idx] = base;
    base += v[k];
  }
  const int64_t last = (int64_t)gridDim.x * kScanTile;
  (void)last;
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == kBlock - 1) out[n] = base;
}

// ---------------------------------------------------- bootstrap ---------
// Multiplicity of every row in the bootstrap of every tree: draw i of tree t
// is rand(seed, BOOT, t, 0, i) % N (gpurf.cu:101-113 keyed by (tree, i)).
__global__ void k_boot_count(uint32_t* mult, int64_t N, int n_trees, const int32_t* tree_global,
                             uint64_t seed) {
  const int tr = blockIdx.y;
  const uint32_t tg = (uint32_t)tree_global[tr];
  uint32_t* mt = mult + (int64_t)tr * N;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t r = rand_u32(seed, kStreamBoot, tg, 0, (uint64_t)i) % (uint32_t)N;
    atomicAdd(&mt[r], 1u);
  }
  (void)n_trees;
}

__global__ void k_nonzero(const uint32_t* mult, int64_t n, uint32_t* flag) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    flag[i] = mult[i] ? 1u : 0u;
}

// Compact in-bag rows into entries (ascending within each tree) and
// accumulate each root's statistics.  A few workgroups per tree (grid-stride
// over its rows), class counts in per-thread LDS columns: the root's
// counters then take a handful of global atomics per workgroup instead of
// one per row range, and no two lanes add into one LDS word.
constexpr int kBootCompactBlocks = 32;  // workgroups per tree
__global__ void __launch_bounds__(kBlock) k_boot_compact(const uint32_t* mult, const uint32_t* pos, int64_t N,
                                                         const uint32_t* label, const int64_t* yfix, int C,
                                                         uint32_t* row, uint32_t* aux, int64_t* ey,
                                                         uint32_t* root_cls, int64_t* root_sum, uint64_t* root_hash,
                                                         uint32_t* overflow) {
  __shared__ uint32_t s_cls[32 * kBlock];  // [class * kBlock + thread]
  const int tr = blockIdx.y;
  const uint32_t* mt = mult + (int64_t)tr * N;
  const uint32_t* pt = pos + (int64_t)tr * N;
  const int t = threadIdx.x;
  if (!yfix)
    for (int c = 0; c < C; ++c) s_cls[c * kBlock + t] = 0;
  uint64_t hsh = 0;
  int64_t sm = 0;
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < N; r += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t w = mt[r];
    if (!w) continue;
    if (w > 0xFFFFu) atomicOr(overflow, 1u);
    const uint32_t p = pt[r];
    const uint32_t lab = label ? label[r] : 0u;
    row[p] = (uint32_t)r;
    aux[p] = (lab << 16) | (w & 0xFFFFu);
    if (yfix) {
      ey[p] = yfix[r];
      sm += (int64_t)w * yfix[r];
    } else {
      s_cls[lab * kBlock + t] += w;
    }
    hsh += (uint64_t)w * mix64((uint64_t)r);
  }
  __syncthreads();
  if (!yfix && t < C) {
    uint32_t v = 0;
    for (int i = 0; i < kBlock; ++i) v += s_cls[t * kBlock + i];
    if (v) atomicAdd(&root_cls[(int64_t)tr * C + t], v);
  }
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) {
    hsh += __shfl_xor(hsh, off, kWave);
    sm += __shfl_xor(sm, off, kWave);
  }
  if ((threadIdx.x & (kWave - 1)) == 0) {
    atomicAdd((unsigned long long*)&root_hash[tr], (unsigned long long)hsh);
    if (yfix) atomicAdd((unsigned long long*)&root_sum[tr], (unsigned long long)sm);
  }
}

// Extra trees: every tree starts from all rows with weight 1 (gpuert.cu:97-102).
__global__ void k_identity_entries(int64_t N, int n_trees, const uint32_t* label, const int64_t* yfix,
                                   uint32_t* row, uint32_t* aux, int64_t* ey) {
  const int tr = blockIdx.y;
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < N; r += (int64_t)gridDim.x * blockDim.x) {
    const int64_t p = (int64_t)tr * N + r;
    row[p] = (uint32_t)r;
    aux[p] = ((label ? label[r] : 0u) << 16) | 1u;
    if (yfix) ey[p] = yfix[r];
  }
  (void)n_trees;
}

__global__ void k_root_stats_identity(int64_t N, const uint32_t* label, const int64_t* yfix, int C,
                                      uint32_t* cls1, int64_t* sum1, uint64_t* hash1) {
  __shared__ uint32_t s_cls[32];  // C <= 32: counted in LDS, one global atomic per class and block
  if (threadIdx.x < 32) s_cls[threadIdx.x] = 0u;
  __syncthreads();
  uint64_t hsh = 0;
  int64_t sm = 0;
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < N; r += (int64_t)gridDim.x * blockDim.x) {
    hsh += mix64((uint64_t)r);
    if (yfix) sm += yfix[r];
    else atomicAdd(&s_cls[label[r]], 1u);
  }
  __syncthreads();
  if (!yfix && (int)threadIdx.x < C && s_cls[threadIdx.x]) atomicAdd(&cls1[threadIdx.x], s_cls[threadIdx.x]);
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) {
    hsh += __shfl_xor(hsh, off, kWave);
    sm += __shfl_xor(sm, off, kWave);
  }
  if ((threadIdx.x & (kWave - 1)) == 0) {
    atomicAdd((unsigned long long*)hash1, (unsigned long long)hsh);
    if (yfix) atomicAdd((unsigned long long*)sum1, (unsigned long long)sm);
  }
}

__global__ void k_canon(float* X, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    if (X[i] == 0.0f) X[i] = 0.0f;
}

// ------------------------------------------------------ finalize --------
struct FinalArgs {
  int64_t n_nodes;
  int C;
  const int32_t* tree;
  const uint32_t* id;
  const int64_t* tree_start;
  const int32_t* feature;
  const float* thr;
  const float* score;
  const int64_t* left;
  const int32_t* depth;
  const int32_t* count;
  const int32_t* m;
  const uint64_t* hash;
  const uint32_t* cls;
  const int64_t* sum;
  const float* value;
  int32_t* f_feature;
  float* f_thr;
  float* f_score;
  int32_t* f_left;
  int32_t* f_depth;
  int32_t* f_count;
  int32_t* f_m;
  uint64_t* f_hash;
  uint32_t* f_cls;
  int64_t* f_sum;
  float* f_value;
};

__global__ void k_finalize(FinalArgs a) {
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < a.n_nodes;
       g += (int64_t)gridDim.x * blockDim.x) {
    const int64_t pos = a.tree_start[a.tree[g]] + a.id[g];
    a.f_feature[pos] = a.feature[g];
    a.f_thr[pos] = a.thr[g];
    a.f_score[pos] = a.score[g];
    const int64_t l = a.left[g];
    a.f_left[pos] = l < 0 ? -1 : (int32_t)(a.tree_start[a.tree[l]] + a.id[l]);
    a.f_depth[pos] = a.depth[g];
    a.f_count[pos] = a.count[g];
    a.f_m[pos] = a.m[g];
    a.f_hash[pos] = a.hash[g];
    if (a.cls)
      for (int c = 0; c < a.C; ++c) a.f_cls[pos * a.C + c] = a.cls[g * a.C + c];
    a.f_sum[pos] = a.sum ? a.sum[g] : 0;
    a.f_value[pos] = a.value[g];
  }
}

__global__ void k_predict_nodes(int64_t n, int C, int task, const int32_t* feature, const float* thr,
                                const int32_t* left, const uint32_t* cls, const float* value, PNode* out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    PNode p;
    p.feature = feature[i];
    p.pad = 0;
    if (p.feature >= 0) {
      p.thr = thr[i];
      p.left = left[i];
    } else if (task == 0) {
      // vote = argmax, strict > from 0: lowest class on ties (gpudte.cu:341-351)
      int best = 0;
      uint32_t bv = 0;
      for (int c = 0; c < C; ++c)
        if (cls[i * C + c] > bv) {
          bv = cls[i * C + c];
          best = c;
        }
      p.thr = 0.0f;
      p.left = best;
    } else {
      p.thr = value[i];
      p.left = 0;
    }
    out[i] = p;
  }
}

// ---------------------------------------- device-described level step --
// Partition tiers: 16- and 32-lane segments, a wave, a workgroup per node.
constexpr Cuts kPartCuts{{16, 32, 512, 0x7FFFFFFF, 0, 0}, 4};

__global__ void k_open_list_dev(const uint8_t* open, const uint32_t* pos, const DevLevel* L, uint32_t* list) {
  const int64_t n = L->n;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    if (open[i]) list[pos[i]] = (uint32_t)i;
}

__global__ void k_compact_list_dev(const uint32_t* flag, const uint32_t* pos, const unsigned int* n_dev,
                                   uint32_t* list) {
  const int64_t n = *n_dev;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    if (flag[i]) list[pos[i]] = (uint32_t)i;
}

// Exclusive scan of n u32 (n read from the device) into out[0..n], out[n] =
// total, also stored to *total: kDScanBlocks workgroups each own a
// contiguous range (per-range sums; then every range sums the sums before
// it for its offset and is rescanned in tiles of kScanTile).  Two kernels:
// a level step is a chain of short launches, each a fixed latency.
constexpr int kDScanBlocks = 512;

__device__ __forceinline__ int64_t dscan_n(const int64_t* n64, const unsigned int* n32) {
  return n64 ? *n64 : (int64_t)*n32;
}

// In: uint32_t counts, or uint8_t flags (a level's open flags, scanned as
// they are).
template <class In>
__global__ void __launch_bounds__(kBlock) k_dscan_sums(const In* in, const int64_t* n64,
                                                       const unsigned int* n32, uint32_t* sums) {
  __shared__ uint32_t s_w[kWaves];
  const int64_t n = dscan_n(n64, n32);
  const int64_t per = (n + kDScanBlocks - 1) / kDScanBlocks;
  const int64_t lo = blockIdx.x * per, hi = min(lo + per, n);
  uint32_t acc = 0;
  for (int64_t i = lo + threadIdx.x; i < hi; i += kBlock) acc += in[i];
  uint32_t tot;
  block_excl_scan(acc, s_w, &tot);
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

template <class In>
__global__ void __launch_bounds__(kBlock) k_dscan_apply(const In* in, const int64_t* n64,
                                                        const unsigned int* n32, const uint32_t* sums, uint32_t* out,
                                                        unsigned int* total) {
  __shared__ uint32_t s_w[kWaves];
  const int64_t n = dscan_n(n64, n32);
  const int64_t per = (n + kDScanBlocks - 1) / kDScanBlocks;
  const int64_t lo = blockIdx.x * per, hi = min(lo + per, n);
  uint32_t part = 0;
  for (int i = threadIdx.x; i < (int)blockIdx.x; i += kBlock) part += sums[i];
  uint32_t carry;
  (void)block_excl_scan(part, s_w, &carry);  // sum of the ranges before this one
  if (blockIdx.x == kDScanBlocks - 1 && threadIdx.x == 0) {
    const uint32_t tot = carry + sums[blockIdx.x];
    out[n] = tot;
    if (total) *total = tot;
  }
  for (int64_t t0 = lo; t0 < hi; t0 += kScanTile) {
    uint32_t v[16];
    uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const int64_t idx = t0 + (int64_t)threadIdx.x * 16 + k;
      v[k] = idx < hi ? (uint32_t)in[idx] : 0u;
      sum += v[k];
    }
    uint32_t tile_tot;
    uint32_t base = block_excl_scan(sum, s_w, &tile_tot) + carry;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const int64_t idx = t0 + (int64_t)threadIdx.x * 16 + k;
      if (idx < hi) out[idx] = base;
      base += v[k];
    }
    carry += tile_tot;
    __syncthreads();  // s_w is reused by the next tile's scan
  }
}

// End of a level step: the next level becomes current; its counters start
// at zero (a kernel, so that a captured step holds no memset node).
__device__ __forceinline__ void level_advance(DevLevel* L) {
  L->off += L->n;
  L->n = L->n_next;
  L->depth += 1;
  L->n_open = 0;
  L->n_retry = 0;
  L->n_next = 0;
  for (int k = 0; k < kMaxTiers; ++k) L->tier[k] = L->rtier[k] = L->ptier[k] = 0;
}

// Last kernel of a device level step: the per-tree cursors of the next level
// (k_tree_update's work) and, by one thread, the level advance.  Nothing in
// the kernel reads L, so the advance cannot race with the tree update.
__global__ void k_tree_update_advance(int n_trees, const uint32_t* cpos, const uint32_t* lf, const uint32_t* base,
                                      uint32_t* lf_next, uint32_t* base_next, DevLevel* L) {
  const int tr = blockIdx.x * blockDim.x + threadIdx.x;
  if (tr == 0) level_advance(L);
  if (tr > n_trees) return;
  lf_next[tr] = cpos[lf[tr]];
  if (tr < n_trees) base_next[tr] = base[tr] + (lf[tr + 1] - lf[tr]);
}

// ------------------------------------------------------ small helpers ---
__global__ void k_open_flags(const uint8_t* open, int64_t n, uint32_t* flag) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    flag[i] = open[i] ? 1u : 0u;
}

__global__ void k_open_list(const uint8_t* open, const uint32_t* pos, int64_t n, uint32_t* list) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    if (open[i]) list[pos[i]] = (uint32_t)i;
}

__global__ void k_compact_list(const uint32_t* flag, const uint32_t* pos, int64_t n, uint32_t* list) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    if (flag[i]) list[pos[i]] = (uint32_t)i;
}

__global__ void k_add_u64(unsigned long long* p, unsigned long long v) { *p += v; }

__global__ void k_iota(uint32_t* out, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = (uint32_t)i;
}

// Appends one tree batch to the merged canonical arrays: child links and
// tree starts shift by the batch's node offset.
__global__ void k_offset_links(const int32_t* left, int64_t n, int32_t off, int32_t* out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t l = left[i];
    out[i] = l < 0 ? l : l + off;
  }
}
__global__ void k_offset_i64(const int64_t* in, int64_t n, int64_t off, int64_t* out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = in[i] + off;
}

__global__ void k_strided_gather(const uint32_t* in, int64_t stride, int n, uint32_t* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[(int64_t)i * stride];
}

// Largest node of a level (out[0]) and the entries of its nodes above
// `cut` entries (out[1]).
__global__ void k_level_sizes(const int32_t* m, int64_t n, int cut, unsigned long long* out) {
  unsigned long long v = 0, big = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const unsigned long long x = (unsigned long long)m[i];
    v = max(v, x);
    big += x > (unsigned long long)cut ? x : 0ull;
  }
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) {
    v = max(v, (unsigned long long)__shfl_xor(v, off, kWave));
    big += __shfl_xor(big, off, kWave);
  }
  if ((threadIdx.x & (kWave - 1)) == 0) {
    atomicMax(&out[0], v);
    if (big) atomicAdd(&out[1], big);
  }
}

// ------------------------------------------------------ host side -------
int grid_for(int64_t n, int per_block = kBlock, int cap = 65535 * 4) {
  int64_t g = (n + per_block - 1) / per_block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (int)g;
}

// Exclusive scan of n u32 into out[0..n] (out[n] = total).
void scan_u32(const uint32_t* in, int64_t n, uint32_t* out, hipStream_t s, std::vector<DevBuf<uint32_t>>& tmp,
              int depth = 0) {
  const int64_t tiles = (n + kScanTile - 1) / kScanTile;
  // Sized once: a resize during the recursion would move the buffers that
  // outer levels still reference.
  if (tmp.size() < 16) tmp.resize(16);
  if (depth >= 7) fail(DTE_ERR_INTERNAL, "scan too deep");
  DevBuf<uint32_t>& sums = tmp[2 * depth];
  DevBuf<uint32_t>& offs = tmp[2 * depth + 1];
  sums.reserve((size_t)tiles + 1);
  offs.reserve((size_t)tiles + 1);
  if (n == 0) {
    DTE_HIP(hipMemsetAsync(out, 0, sizeof(uint32_t), s));
    return;
  }
  hipLaunchKernelGGL(k_scan_tiles, dim3((unsigned)tiles), dim3(kBlock), 0, s, in, n, sums.p);
  if (tiles == 1) {
    DTE_HIP(hipMemsetAsync(offs.p, 0, sizeof(uint32_t), s));
  } else {
    scan_u32(sums.p, tiles, offs.p, s, tmp, depth + 1);
  }
  hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)tiles), dim3(kBlock), 0, s, in, n, offs.p, out);
}

template <typename T>
T read_one(const T* p, hipStream_t s) {
  T v;
  DTE_HIP(hipMemcpyAsync(&v, p, sizeof(T), hipMemcpyDeviceToHost, s));
  DTE_HIP(hipStreamSynchronize(s));
  return v;
}


// Profile family of a split-search tier (one kernel per family).
const char* search_family(int algorithm, int tier) {
  static const char* rf[] = {"search_seg16", "search_seg32", "search_seg64", "search_team64", "search_block",
                             "search_chunked", "search_block"};
  static const char* ert[] = {"search_ert_seg16", "search_ert_seg32", "search_ert_seg64", "search_ert_team",
                              "search_ert_wave", "search_ert_block"};
  return algorithm == DTE_RANDOM_FOREST ? rf[tier] : ert[tier];
}

// RF tiers: 0/1/2 = one 16/32/64-lane segment per node (one entry per lane,
// draws looped in the kernel); 3 = wave teams (four items per workgroup);
// 4 = whole workgroup per item within the LDS capacity; 5 = the chunk
// pipeline (big_pipeline); 6 = whole workgroup per item of any size (the
// plain driver).  ERT tiers: 0/1/2 = one 16/32/64-lane segment per node
// (draws looped in the kernel), 3 = one wave per node (entries in
// registers), 4 = one wave per item, 5 = workgroup per item.
template <int MAXC, int TASK>
void launch_search(int algorithm, int tier, int grid, hipStream_t s, const DataArgs& d, const LevelArgs& lv,
                   const ItemArgs& it) {
  if (algorithm == DTE_RANDOM_FOREST) {
    switch (tier) {
      case 0: hipLaunchKernelGGL((k_search_rf_seg<MAXC, TASK, 16>), dim3(grid), dim3(kBlock), 0, s, d, lv, it); break;
      case 1: hipLaunchKernelGGL((k_search_rf_seg<MAXC, TASK, 32>), dim3(grid), dim3(kBlock), 0, s, d, lv, it); break;
      case 2: hipLaunchKernelGGL((k_search_rf_seg<MAXC, TASK, 64>), dim3(grid), dim3(kBlock), 0, s, d, lv, it); break;
      case 3: hipLaunchKernelGGL((k_search_rf_team<MAXC, TASK, 64>), dim3(grid), dim3(kBlock), 0, s, d, lv, it); break;
      case 4: hipLaunchKernelGGL((k_search_rf_team<MAXC, TASK, 256>), dim3(grid), dim3(kBlock), 0, s, d, lv, it); break;
      default: hipLaunchKernelGGL((k_search_rf<MAXC, TASK, 256, true>), dim3(grid), dim3(kBlock), 0, s, d, lv, it); break;
    }
  } else {
    switch (tier) {
      case 0: hipLaunchKernelGGL((k_search_ert_seg<MAXC, TASK, 16>), dim3(grid), dim3(kBlock), 0, s, d, lv, it); break;
      case 1: hipLaunchKernelGGL((k_search_ert_seg<MAXC, TASK, 32>), dim3(grid), dim3(kBlock), 0, s, d, lv, it); break;
      case 2: hipLaunchKernelGGL((k_search_ert_seg<MAXC, TASK, 64>), dim3(grid), dim3(kBlock), 0, s, d, lv, it); break;
      case 3: hipLaunchKernelGGL((k_search_ert_team<MAXC, TASK>), dim3(grid), dim3(kBlock), 0, s, d, lv, it); break;
      case 4: hipLaunchKernelGGL((k_search_ert_wave<MAXC, TASK>), dim3(grid), dim3(kBlock), 0, s, d, lv, it); break;
      default: hipLaunchKernelGGL((k_search_ert<MAXC, TASK>), dim3(grid), dim3(kBlock), 0, s, d, lv, it); break;
    }
  }
}

template <int MAXC, int TASK>
void launch_partition(int grid, hipStream_t s, const DataArgs& d, const LevelArgs& lv, const PartArgs& pa) {
  hipLaunchKernelGGL((k_partition<MAXC, TASK>), dim3(grid), dim3(kBlock), 0, s, d, lv, pa);
}

template <int MAXC, int TASK>
void launch_partition_wave(int grid, hipStream_t s, const DataArgs& d, const LevelArgs& lv, const PartArgs& pa,
                           const uint32_t* list, int64_t n) {
  hipLaunchKernelGGL((k_partition_wave<MAXC, TASK>), dim3(grid), dim3(kBlock), 0, s, d, lv, pa, list, n);
}

// Nodes of at most 16 (32) entries, 4 (2) per wave; the grid covers n nodes.
template <int MAXC, int TASK, int SEG>
void launch_partition_seg(int grid_cap, hipStream_t s, const DataArgs& d, const LevelArgs& lv, const PartArgs& pa,
                          const uint32_t* list, int64_t n) {
  constexpr int per_block = kWaves * (kWave / SEG);
  const int grid = (int)std::min<int64_t>((n + per_block - 1) / per_block, grid_cap);
  hipLaunchKernelGGL((k_partition_seg<MAXC, TASK, SEG>), dim3(grid), dim3(kBlock), 0, s, d, lv, pa, list, n);
}


template <int MAXC, int TASK>
void launch_partition_chunks(hipStream_t s, const DataArgs& d, const LevelArgs& lv, const PartArgs& pa,
                             const PartChunks& pc, int64_t n_chunks) {
  hipLaunchKernelGGL((k_part_chunk_stats<MAXC, TASK>), dim3((unsigned)n_chunks), dim3(kBlock), 0, s, d, lv, pa, pc);
  hipLaunchKernelGGL((k_part_chunk_finish<TASK>), dim3((unsigned)pc.n_nodes), dim3(kBlock), 0, s, d, lv, pa, pc);
  hipLaunchKernelGGL((k_part_chunk_scatter<TASK>), dim3((unsigned)n_chunks), dim3(kBlock), 0, s, d, lv, pa, pc);
}

// Scratch and bookkeeping of the big-node pipeline, reused across levels.
struct BigState {
  DevBuf<V2Item> items;
  DevBuf<V2Tile> htiles, stiles, ctiles;
  DevBuf<uint32_t> bcount, bmax, bcls, bwn, bbase, bcur, bsup, carry_cls, carry_last, res_key, res_prev;
  DevBuf<int64_t> bsum, carry_n, carry_s;
  DevBuf<double> res_g;
  DevBuf<uint32_t> skey, saux, skey2, saux2, tkey;
  DevBuf<int64_t> sy, sy2;
  DevBuf<V2Task> tasks, spill;
  DevBuf<unsigned int> ctr;
  DevBuf<uint32_t> lists;  // item indices: small items (wave teams), then the rest
  std::vector<V2Item> host_items;
  std::vector<V2Tile> host_htiles, host_stiles, host_ctiles, host_stiles_mid, host_stiles_many;
  std::vector<uint32_t> host_lists;
  std::vector<int64_t> host_fcount;   // histogram tiles: items per feature
  std::vector<uint32_t> host_forder;  // batch items in feature order
  std::vector<V2Tile> host_forder_tiles;
  int target = kV2Target;  // entries per bucket aimed at
};

// Chunk-parallel search of the big items of one phase (see k_v2_*): items in
// batches of at most kV2Batch scratch entries.  A batch's scratch is about
// 20 B (classification) to 36 B (regression) per entry, so 1.5 Gi entries
// hold a whole level of a 200-tree 1M-row fit in one batch at 32-58 GB of
// the 288 GB of HBM (fewer, larger launches and host round trips than
// 384 Mi-entry batches: headline fit 1231 -> 1221 ms, 4M x 64 regression
// fit 3.98 -> 3.76 s with the side stream of the other tiers).
constexpr int64_t kV2Batch = (int64_t)1536 << 20;

template <int MAXC, int TASK>
void big_pipeline(BigState& B, const DataArgs& d, const LevelArgs& lv, const ItemArgs& it, const uint32_t* groups,
                  int64_t n_groups, hipStream_t s, int num_cu) {
  const int64_t n_all = n_groups * it.per_group;
  if (n_all == 0) return;
  const int nb_cap = kV2MaxB;
  const int C = d.C;
  B.items.reserve(n_all);
  hipLaunchKernelGGL(k_v2_items, dim3(grid_for(n_all)), dim3(kBlock), 0, s, d, lv, groups, n_groups, it.per_group,
                     it.j0, it.D, nb_cap, B.target, B.items.p);
  B.host_items.resize(n_all);
  DTE_HIP(hipMemcpyAsync(B.host_items.data(), B.items.p, sizeof(V2Item) * n_all, hipMemcpyDeviceToHost, s));
  DTE_HIP(hipStreamSynchronize(s));
  constexpr int kSTile = TASK == 0 ? kScatterTile : 2048;
  int64_t k0 = 0;
  while (k0 < n_all) {
    int64_t k1 = k0, el = 0, nb = 0;
    B.host_htiles.clear();
    B.host_stiles.clear();
    B.host_stiles_mid.clear();
    B.host_stiles_many.clear();
    B.host_ctiles.clear();
    B.host_lists.clear();
    while (k1 < n_all && (k1 == k0 || el + B.host_items[k1].m <= kV2Batch)) {
      V2Item& bi = B.host_items[k1];
      bi.boff = (int32_t)nb;
      bi.soff = el;
      const uint32_t li = (uint32_t)(k1 - k0);
      auto& st = bi.nb <= kV2FewB ? B.host_stiles : (bi.nb <= kV2MidB ? B.host_stiles_mid : B.host_stiles_many);
      bi.sgrp = bi.nb > kV2MidB ? (bi.nb + kV2Super - 1) / kV2Super : 0;  // two-pass scatter
      for (int e0 = 0; e0 < bi.m; e0 += kSTile) st.push_back(V2Tile{li, (uint32_t)e0});
      for (int b0 = 0; b0 < bi.nb; b0 += kV2CTile) B.host_ctiles.push_back(V2Tile{li, (uint32_t)b0});
      el += bi.m;
      nb += bi.nb;
      ++k1;
    }
    const int64_t n_items = k1 - k0;
    // Histogram tiles feature-major: the tiles in flight gather from a few
    // feature columns of K at a time instead of from all of them, so the
    // columns stay in the caches.  The histogram's results do not depend on
    // the tile order (per-bucket counts and maxima, keys stored per entry).
    B.host_fcount.assign((size_t)d.F + 1, 0);
    for (int64_t k = k0; k < k1; ++k) ++B.host_fcount[(size_t)B.host_items[k].f + 1];
    for (int f = 0; f < d.F; ++f) B.host_fcount[(size_t)f + 1] += B.host_fcount[(size_t)f];
    B.host_forder.resize((size_t)n_items);
    for (int64_t k = k0; k < k1; ++k) B.host_forder[B.host_fcount[(size_t)B.host_items[k].f]++] = (uint32_t)(k - k0);
    for (const uint32_t li : B.host_forder)
      for (int e0 = 0; e0 < B.host_items[k0 + li].m; e0 += kV2Tile)
        B.host_forder_tiles.push_back(V2Tile{li, (uint32_t)e0});
    {
      // Workgroups go to the 8 XCDs round robin (workgroup w to XCD w % 8),
      // each XCD with its own L2: XCD x takes the x-th eighth of the
      // feature-ordered tiles, so each L2 holds its own few columns.
      const int64_t nt = (int64_t)B.host_forder_tiles.size();
      const int64_t q = (nt + kXcds - 1) / kXcds;
      B.host_htiles.reserve((size_t)nt);
      for (int64_t j = 0; j < q; ++j)
        for (int x = 0; x < kXcds; ++x) {
          const int64_t src = x * q + j;
          if (src < std::min<int64_t>((int64_t)(x + 1) * q, nt)) B.host_htiles.push_back(B.host_forder_tiles[src]);
        }
      B.host_forder_tiles.clear();
    }
    // scatter tiles of items with few buckets first, then the rest
    const int64_t n_sfew = (int64_t)B.host_stiles.size();
    B.host_stiles.insert(B.host_stiles.end(), B.host_stiles_mid.begin(), B.host_stiles_mid.end());
    const int64_t n_smid = (int64_t)B.host_stiles.size();
    B.host_stiles.insert(B.host_stiles.end(), B.host_stiles_many.begin(), B.host_stiles_many.end());
    const int64_t n_sall = (int64_t)B.host_stiles.size();
    for (int64_t k = k0; k < k1; ++k)
      if (B.host_items[k].nb <= kV2SmallB) B.host_lists.push_back((uint32_t)(k - k0));
    const int64_t n_small = (int64_t)B.host_lists.size();
    for (int64_t k = k0; k < k1; ++k)
      if (B.host_items[k].nb > kV2SmallB) B.host_lists.push_back((uint32_t)(k - k0));
    const int64_t n_large = n_items - n_small;
    B.lists.reserve(n_items);
    DTE_HIP(hipMemcpyAsync(B.lists.p, B.host_lists.data(), 4 * n_items, hipMemcpyHostToDevice, s));
    DTE_HIP(hipMemcpyAsync(B.items.p, B.host_items.data() + k0, sizeof(V2Item) * n_items, hipMemcpyHostToDevice, s));
    B.htiles.reserve(B.host_htiles.size());
    B.stiles.reserve(B.host_stiles.size());
    B.ctiles.reserve(B.host_ctiles.size());
    DTE_HIP(hipMemcpyAsync(B.ctiles.p, B.host_ctiles.data(), sizeof(V2Tile) * B.host_ctiles.size(),
                           hipMemcpyHostToDevice, s));
    DTE_HIP(hipMemcpyAsync(B.htiles.p, B.host_htiles.data(), sizeof(V2Tile) * B.host_htiles.size(),
                           hipMemcpyHostToDevice, s));
    DTE_HIP(hipMemcpyAsync(B.stiles.p, B.host_stiles.data(), sizeof(V2Tile) * B.host_stiles.size(),
                           hipMemcpyHostToDevice, s));
    B.bcount.reserve(nb);
    B.bmax.reserve(nb);
    B.bbase.reserve(nb);
    B.bcur.reserve(nb);
    B.bsup.reserve(nb);
    B.carry_n.reserve(nb);
    B.carry_last.reserve(nb);
    B.res_g.reserve(nb);
    B.res_key.reserve(nb);
    B.res_prev.reserve(nb);
    B.tasks.reserve(nb);
    B.spill.reserve(nb);
    B.ctr.reserve(2);
    if (TASK == 0) {
      B.bcls.reserve((size_t)nb * C);
      B.carry_cls.reserve((size_t)nb * C);
    } else {
      B.bwn.reserve(nb);
      B.bsum.reserve(nb);
      B.carry_s.reserve(nb);
    }
    B.skey.reserve(el);
    B.saux.reserve(el);
    B.skey2.reserve(el);
    B.saux2.reserve(el);
    if (n_sall > n_smid) B.tkey.reserve(el);
    if (TASK == 1) {
      B.sy.reserve(el);
      B.sy2.reserve(el);
    }
    DTE_HIP(hipMemsetAsync(B.bcount.p, 0, 4 * nb, s));
    DTE_HIP(hipMemsetAsync(B.bmax.p, 0, 4 * nb, s));
    DTE_HIP(hipMemsetAsync(B.ctr.p, 0, 8, s));
    V2Args a{};
    a.items = B.items.p;
    a.n_items = n_items;
    a.bcount = B.bcount.p;
    a.bmax = B.bmax.p;
    a.bcls = B.bcls.p;
    a.bwn = B.bwn.p;
    a.bsum = B.bsum.p;
    a.bbase = B.bbase.p;
    a.bcur = B.bcur.p;
    a.bsup = B.bsup.p;
    a.carry_cls = B.carry_cls.p;
    a.carry_n = B.carry_n.p;
    a.carry_s = B.carry_s.p;
    a.carry_last = B.carry_last.p;
    a.res_g = B.res_g.p;
    a.res_key = B.res_key.p;
    a.res_prev = B.res_prev.p;
    a.skey = B.skey.p;
    a.saux = B.saux.p;
    a.sy = B.sy.p;
    a.skey2 = B.skey2.p;
    a.gkey = B.skey2.p;
    a.saux2 = B.saux2.p;
    a.sy2 = B.sy2.p;
    a.tkey = B.tkey.p;
    a.tasks = B.tasks.p;
    a.spill = B.spill.p;
    a.counters = B.ctr.p;
    a.tiles = B.htiles.p;
    {
      Timer tm("chunk_hist", 0.0, s);
      hipLaunchKernelGGL(k_v2_hist<TASK>, dim3((unsigned)B.host_htiles.size()), dim3(kBlock), 0, s, d, lv, a);
    }
    {
      Timer tm("chunk_scan", 0.0, s);
      if (n_small)
        hipLaunchKernelGGL(k_v2_scan<kWave>, dim3((unsigned)((n_small + kWaves - 1) / kWaves)), dim3(kBlock), 0, s,
                           a, B.lists.p, n_small);
      if (n_large)
        hipLaunchKernelGGL(k_v2_scan<256>, dim3((unsigned)n_large), dim3(kBlock), 0, s, a, B.lists.p + n_small,
                           n_large);
    }
    a.tiles = B.stiles.p;
    {
      Timer tm("chunk_scatter", 0.0, s);
      if (n_sfew)
        hipLaunchKernelGGL((k_v2_scatter<TASK, kV2FewB>), dim3((unsigned)n_sfew), dim3(kBlock), 0, s, d, lv, a);
      a.tiles = B.stiles.p + n_sfew;
      if (n_smid > n_sfew)
        hipLaunchKernelGGL((k_v2_scatter<TASK, kV2MidB>), dim3((unsigned)(n_smid - n_sfew)), dim3(kBlock), 0, s, d,
                           lv, a);
      a.tiles = B.stiles.p + n_smid;
      if (n_sall > n_smid) {
        // many buckets: super-buckets, then buckets, over the same tiles
        hipLaunchKernelGGL((k_v2_scatter<TASK, kV2Super, 1>), dim3((unsigned)(n_sall - n_smid)), dim3(kBlock), 0, s,
                           d, lv, a);
        hipLaunchKernelGGL((k_v2_scatter<TASK, kV2MaxB, 2>), dim3((unsigned)(n_sall - n_smid)), dim3(kBlock), 0, s,
                           d, lv, a);
      }
    }
    a.tiles = B.ctiles.p;
    {
      Timer tm("chunk_ctot", 0.0, s);
      hipLaunchKernelGGL((k_v2_ctot<MAXC, TASK>), dim3((unsigned)B.host_ctiles.size()), dim3(kBlock), 0, s, d, a);
    }
    {
      Timer tm("chunk_carry", 0.0, s);
      if (n_small)
        hipLaunchKernelGGL((k_v2_carry<TASK, kWave>), dim3((unsigned)((n_small + kWaves - 1) / kWaves)), dim3(kBlock),
                           0, s, d, a, B.lists.p, n_small);
      if (n_large)
        hipLaunchKernelGGL((k_v2_carry<TASK, 256>), dim3((unsigned)n_large), dim3(kBlock), 0, s, d, a,
                           B.lists.p + n_small, n_large);
    }
    unsigned int nt[2];
    DTE_HIP(hipMemcpyAsync(nt, B.ctr.p, 8, hipMemcpyDeviceToHost, s));
    DTE_HIP(hipStreamSynchronize(s));
    // Algorithmic bytes of the batch's items (12 B per entry, 20 B for
    // regression) belong to the evaluation kernel, where every entry is read
    // and scored; the other pipeline kernels are the price of bucketing.
    if (nt[0]) {
      Timer tm("chunk_eval", profiling_enabled() ? (double)el * (TASK == 0 ? 12.0 : 20.0) : 0.0, s);
      // Two full rounds of the three workgroups a CU holds: grid-stride
      // over similar buckets, so no partial last round (8 per CU, 2.67
      // rounds, was 4.5 % slower).
      const dim3 grid((unsigned)std::min<int64_t>((nt[0] + kWaves - 1) / kWaves, (int64_t)num_cu * 6));
      if constexpr (TASK == 1)
        hipLaunchKernelGGL((k_v2_eval_reg<MAXC, TASK>), grid, dim3(kBlock), 0, s, d, lv, a, (int64_t)nt[0]);
      else
        hipLaunchKernelGGL((k_v2_eval<MAXC, TASK>), grid, dim3(kBlock), 0, s, d, lv, a, (int64_t)nt[0]);
    }
    if (nt[1]) {
      Timer tm("chunk_spill", 0.0, s);
      hipLaunchKernelGGL((k_v2_spill<MAXC, TASK>), dim3((unsigned)std::min<int64_t>(nt[1], (int64_t)num_cu * 4)),
                         dim3(kBlock), 0, s, d, lv, a, (int64_t)nt[1]);
    }
    hipLaunchKernelGGL(k_v2_reduce<TASK>, dim3((unsigned)((n_items + kWaves - 1) / kWaves)), dim3(kBlock), 0, s, d,
                       a, it);
    if (it.bytes) {
      // algorithmic bytes of the batch: 12 (20) bytes per entry
      const unsigned long long by = (unsigned long long)el * (TASK == 0 ? 12ull : 20ull);
      hipLaunchKernelGGL(k_add_u64, dim3(1), dim3(1), 0, s, it.bytes, by);
    }
    k0 = k1;
  }
}

using BigFn = void (*)(BigState&, const DataArgs&, const LevelArgs&, const ItemArgs&, const uint32_t*, int64_t,
                       hipStream_t, int);
BigFn pick_big(int task, int C) {
  if (task == 1) return big_pipeline<1, 1>;
  if (C <= 2) return big_pipeline<2, 0>;
  if (C <= 4) return big_pipeline<4, 0>;
  if (C <= 8) return big_pipeline<8, 0>;
  if (C <= 16) return big_pipeline<16, 0>;
  return big_pipeline<32, 0>;
}

using SearchFn = void (*)(int, int, int, hipStream_t, const DataArgs&, const LevelArgs&, const ItemArgs&);
using PartFn = void (*)(int, hipStream_t, const DataArgs&, const LevelArgs&, const PartArgs&);
using PartWaveFn = void (*)(int, hipStream_t, const DataArgs&, const LevelArgs&, const PartArgs&, const uint32_t*,
                            int64_t);

SearchFn pick_search(int task, int C) {
  if (task == 1) return launch_search<1, 1>;
  if (C <= 2) return launch_search<2, 0>;
  if (C <= 4) return launch_search<4, 0>;
  if (C <= 8) return launch_search<8, 0>;
  if (C <= 16) return launch_search<16, 0>;
  return launch_search<32, 0>;
}
template <int SEG>
PartWaveFn pick_part_seg(int task, int C) {
  if (task == 1) return launch_partition_seg<1, 1, SEG>;
  if (C <= 2) return launch_partition_seg<2, 0, SEG>;
  if (C <= 4) return launch_partition_seg<4, 0, SEG>;
  if (C <= 8) return launch_partition_seg<8, 0, SEG>;
  if (C <= 16) return launch_partition_seg<16, 0, SEG>;
  return launch_partition_seg<32, 0, SEG>;
}
using PartChunkFn = void (*)(hipStream_t, const DataArgs&, const LevelArgs&, const PartArgs&, const PartChunks&,
                             int64_t);
PartChunkFn pick_part_chunks(int task, int C) {
  if (task == 1) return launch_partition_chunks<1, 1>;
  if (C <= 2) return launch_partition_chunks<2, 0>;
  if (C <= 4) return launch_partition_chunks<4, 0>;
  if (C <= 8) return launch_partition_chunks<8, 0>;
  if (C <= 16) return launch_partition_chunks<16, 0>;
  return launch_partition_chunks<32, 0>;
}
PartWaveFn pick_part_wave(int task, int C) {
  if (task == 1) return launch_partition_wave<1, 1>;
  if (C <= 2) return launch_partition_wave<2, 0>;
  if (C <= 4) return launch_partition_wave<4, 0>;
  if (C <= 8) return launch_partition_wave<8, 0>;
  if (C <= 16) return launch_partition_wave<16, 0>;
  return launch_partition_wave<32, 0>;
}

PartFn pick_part(int task, int C) {
  if (task == 1) return launch_partition<1, 1>;
  if (C <= 2) return launch_partition<2, 0>;
  if (C <= 4) return launch_partition<4, 0>;
  if (C <= 8) return launch_partition<8, 0>;
  if (C <= 16) return launch_partition<16, 0>;
  return launch_partition<32, 0>;
}

}  // namespace

void scan_u32_ext(const uint32_t* in, int64_t n, uint32_t* out, hipStream_t s,
                  std::vector<DevBuf<uint32_t>>& tmp) {
  scan_u32(in, n, out, s, tmp);
}

void dense_ranks_device(const float* X, int64_t N, int F, DevBuf<uint32_t>& K, DevBuf<float>& SV,
                        DevBuf<int64_t>& SVoff, DevBuf<uint32_t>& S, hipStream_t s,
                        std::vector<DevBuf<uint32_t>>& scan_tmp);

const LnTable& host_ln_table() {
  static LnTable t = [] {
    LnTable x;
    build_ln_table(&x);
    return x;
  }();
  return t;
}

void upload_ln_table(int device) {
  (void)device;
  const LnTable& t = host_ln_table();
  DTE_HIP(hipMemcpyToSymbol(HIP_SYMBOL(g_ln_table), &t, sizeof(LnTable), 0, hipMemcpyHostToDevice));
}

// ------------------------------------------------------------ Builder ---
// Plain launch structure of the reference: at most 100 trees per batch
// (gpudte_algorithm.cc:116-118), at most 512 open nodes per launch chunk
// (gpudte_algorithm.cc:149-151).
constexpr int kPlainTrees = 100;
constexpr int64_t kPlainChunk = 512;

// Largest count with a precomputed x ln x term (512 MB of table).
constexpr int64_t kTMaxCap = (int64_t)1 << 26;

// Largest node a level graph takes (entries): its workgroup-per-item kernel
// needs a scratch slab of this many entries per workgroup.
constexpr int64_t kGraphMaxM = 65536;

__global__ void k_build_T(int64_t* T, int64_t tmax, int F) {
  for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x <= tmax; x += (int64_t)gridDim.x * blockDim.x)
    T[x] = xlnx_fixed((uint64_t)x, F, g_ln_table.inv, g_ln_table.logc);
}

// Device-resident inputs of one fit, shared by every tree batch.
struct FitData {
  DevBuf<float> X;         // canonical copy (-0 -> +0), feature-major
  DevBuf<uint32_t> label;  // classification
  DevBuf<int64_t> yfix;    // regression: fixed-point targets
  int shift = 0;
  DevBuf<uint32_t> K;      // dense ranks (random forest)
  DevBuf<float> SV;
  DevBuf<int64_t> SVoff;
  DevBuf<uint32_t> Sdev;
  DevBuf<int64_t> Ttab;    // fixed-point x ln x terms T[0..Tmax]
  DataArgs d{};
};

void prepare_data(FitData& fd, const float* X_dev_in, const float* y_host, int64_t N, int32_t F,
                  const dte_params& p, hipStream_t s) {
  const int task = p.task;
  const int algorithm = p.algorithm;
  const int C = task == 0 ? p.n_classes : 1;
  if (C > 32) fail(DTE_ERR_INVALID_ARGUMENT, "at most 32 classes are supported");
  if (N > (int64_t)0x7FFFFFFF) fail(DTE_ERR_INVALID_ARGUMENT, "too many rows");
  DevBuf<float>& X = fd.X;
  DevBuf<uint32_t>& label = fd.label;
  DevBuf<int64_t>& yfix = fd.yfix;
  int& shift = fd.shift;
  DevBuf<uint32_t>& K = fd.K;
  DevBuf<float>& SV = fd.SV;
  DevBuf<int64_t>& SVoff = fd.SVoff;
  DevBuf<uint32_t>& Sdev = fd.Sdev;
  DevBuf<int64_t>& Ttab = fd.Ttab;
  DataArgs& d = fd.d;
  // ---- dataset: canonical copy of X, labels / fixed-point targets.
  X.alloc((size_t)N * F);
  DTE_HIP(hipMemcpyAsync(X.p, X_dev_in, sizeof(float) * N * F, hipMemcpyDeviceToDevice, s));
  hipLaunchKernelGGL(k_canon, dim3(grid_for((int64_t)N * F)), dim3(kBlock), 0, s, X.p, (int64_t)N * F);
  if (task == 0) {
    std::vector<uint32_t> lab(N);
    for (int64_t i = 0; i < N; ++i) {
      const int c = (int)y_host[i];
      if (c < 0 || c >= C || (float)c != y_host[i]) fail(DTE_ERR_INVALID_ARGUMENT, "class label out of range");
      lab[i] = (uint32_t)c;
    }
    label.alloc(N);
    DTE_HIP(hipMemcpyAsync(label.p, lab.data(), 4 * N, hipMemcpyHostToDevice, s));
    DTE_HIP(hipStreamSynchronize(s));
  } else {
    double mx = 0;
    for (int64_t i = 0; i < N; ++i) mx = std::max(mx, std::fabs((double)y_host[i]));
    shift = fixed_shift(mx, N);
    std::vector<int64_t> yf(N);
    for (int64_t i = 0; i < N; ++i) yf[i] = std::llrint(std::ldexp((double)y_host[i], shift));
    yfix.alloc(N);
    DTE_HIP(hipMemcpyAsync(yfix.p, yf.data(), 8 * N, hipMemcpyHostToDevice, s));
    DTE_HIP(hipStreamSynchronize(s));
  }
  std::vector<DevBuf<uint32_t>> scan_tmp;
  if (algorithm == DTE_RANDOM_FOREST) {
    {
      Timer tm("dense_ranks", 0.0, s);
      dense_ranks_device(X.p, N, F, K, SV, SVoff, Sdev, s, scan_tmp);
    }
    std::vector<uint32_t> Sh(F);
    DTE_HIP(hipMemcpyAsync(Sh.data(), Sdev.p, 4 * F, hipMemcpyDeviceToHost, s));
    DTE_HIP(hipStreamSynchronize(s));
    for (int f = 0; f < F; ++f) {
      int words = task == 0 ? C : 3;
      int wbits = 0;
      while ((1 << (wbits + 1)) * words <= kHistWords) ++wbits;
      while (((Sh[f] + (1u << wbits) - 1) >> wbits) > (uint32_t)kMaxBuckets) ++wbits;
      if ((1 << wbits) * words > kHistWords)
        fail(DTE_ERR_INVALID_ARGUMENT, "too many distinct values in a feature for the split-search buckets");
    }
  }
  d = DataArgs{};
  d.X = X.p;
  d.K = K.p;
  d.SV = SV.p;
  d.SVoff = SVoff.p;
  d.S = Sdev.p;
  d.N = N;
  d.F = F;
  d.C = task == 0 ? C : 0;
  d.shift = shift;
  d.seed = p.seed;
  d.min_node = p.min_node_size;
  d.max_depth = p.max_depth;
  {
    const LnTable& lt = host_ln_table();
    d.igF = ig_shift((uint64_t)N, task == 0 ? C : 1, lt.inv, lt.logc);
    // Counts of a fit never exceed the root's total weight N.
    d.Tmax = task == 0 ? std::min<int64_t>(N, kTMaxCap) : 0;
    Ttab.alloc(d.Tmax + 1);
    hipLaunchKernelGGL(k_build_T, dim3(grid_for(d.Tmax + 1)), dim3(kBlock), 0, s, Ttab.p, d.Tmax, d.igF);
    d.T = Ttab.p;
  }
}

// A joint build of several extra-trees forests (dte_fit_many): the trees of
// forest f are batch trees [t0, t0 + T) of one build whose level steps serve
// every forest at once.
struct JointForest {
  FitData* fd;
  int64_t N;
  int32_t F;
  int T;
  int t0;
  dte_params p;
};
struct JointSpec {
  std::vector<JointForest> f;
  DataArgs d;  // joint data: forests / forest_of_tree set, C = joint class width
  int C;       // joint class width (the widest forest's)
  int D0, R;   // the widest mandatory and retry draw counts
  int F;       // the widest feature count
};

// Builds trees [t0, t0 + T) of the fit level by level.  plain = 1 is the
// per-node launch structure of the reference (gpudte_algorithm.cc:116-217):
// open nodes in chunks of at most 512, one workgroup per node walking its
// feature draws in order, a host round trip after every chunk; the same
// device functions compute every result, so both modes give the same bits.
Forest* fit_batch(FitData& fd, int t0, int T, int64_t N, int32_t F, const dte_params& p, hipStream_t s,
                  bool plain, const JointSpec* J = nullptr) {
  const auto fit_t0 = std::chrono::steady_clock::now();
  const int task = p.task;
  const int algorithm = p.algorithm;
  const int C = J ? J->C : (task == 0 ? p.n_classes : 1);
  const int k = p.n_features_per_node > 0 ? p.n_features_per_node
                                          : (int)std::round(std::log((double)F)) + 1;
  const int D0 = J ? J->D0 : std::min(k, F);
  const int R = J ? J->R : (k < F ? std::min(10, F - k) : 0);
  const int D = D0 + R;
  const DataArgs d = J ? J->d : fd.d;
  const uint32_t* label = fd.label.p;
  const int64_t* yfix = fd.yfix.p;
  const int shift = fd.shift;
  std::vector<DevBuf<uint32_t>> scan_tmp;
  // per batch tree: rows at its root, forest (joint build), global tree id
  std::vector<int64_t> tN(T, N);
  std::vector<int32_t> tf(T, 0);
  std::vector<int32_t> tg_h(T);
  std::iota(tg_h.begin(), tg_h.end(), t0);
  if (J) {
    for (size_t f = 0; f < J->f.size(); ++f)
      for (int q = 0; q < J->f[f].T; ++q) {
        const int t = J->f[f].t0 + q;
        tN[t] = J->f[f].N;
        tf[t] = (int32_t)f;
        tg_h[t] = J->f[f].p.tree_offset + q;
      }
  }
  auto tree_min_node = [&](int t) { return J ? J->f[tf[t]].p.min_node_size : p.min_node_size; };
  auto tree_max_depth = [&](int t) { return J ? J->f[tf[t]].p.max_depth : p.max_depth; };
  DevBuf<int32_t> tree_global(T);
  DTE_HIP(hipMemcpyAsync(tree_global.p, tg_h.data(), 4 * T, hipMemcpyHostToDevice, s));

  DevBuf<unsigned long long> search_bytes(kMaxTiers), part_bytes(1);  // search: one counter per tier
  DTE_HIP(hipMemsetAsync(search_bytes.p, 0, 8 * kMaxTiers, s));
  DTE_HIP(hipMemsetAsync(part_bytes.p, 0, 8, s));

  // ---- root entries.
  int64_t E = 0;
  std::vector<int64_t> root_start(T), root_m(T);
  DevBuf<uint32_t> rowA, auxA, rowB, auxB;
  DevBuf<int64_t> eyA, eyB;
  DevBuf<uint32_t> root_cls((size_t)T * std::max(C, 1));
  DevBuf<int64_t> root_sum(T);
  DevBuf<uint64_t> root_hash(T);
  DTE_HIP(hipMemsetAsync(root_cls.p, 0, root_cls.bytes(), s));
  DTE_HIP(hipMemsetAsync(root_sum.p, 0, root_sum.bytes(), s));
  DTE_HIP(hipMemsetAsync(root_hash.p, 0, root_hash.bytes(), s));
  if (algorithm == DTE_RANDOM_FOREST) {
    const int64_t TN = (int64_t)T * N;
    DevBuf<uint32_t> mult(TN), flag(TN), pos(TN + 1);
    DTE_HIP(hipMemsetAsync(mult.p, 0, mult.bytes(), s));
    {
      Timer tm("bootstrap", 4.0 * TN, s);
      hipLaunchKernelGGL(k_boot_count, dim3(grid_for(N, kBlock, 4096), T), dim3(kBlock), 0, s, mult.p, N, T,
                         tree_global.p, p.seed);
    }
    hipLaunchKernelGGL(k_nonzero, dim3(grid_for(TN)), dim3(kBlock), 0, s, mult.p, TN, flag.p);
    scan_u32(flag.p, TN, pos.p, s, scan_tmp);
    std::vector<uint32_t> starts(T + 1);
    {
      DevBuf<uint32_t> st(T + 1);
      hipLaunchKernelGGL(k_strided_gather, dim3(grid_for(T + 1)), dim3(kBlock), 0, s, pos.p, N, T + 1, st.p);
      DTE_HIP(hipMemcpyAsync(starts.data(), st.p, 4 * (T + 1), hipMemcpyDeviceToHost, s));
      DTE_HIP(hipStreamSynchronize(s));
    }
    E = starts[T];
    for (int t = 0; t < T; ++t) {
      root_start[t] = starts[t];
      root_m[t] = starts[t + 1] - starts[t];
    }
    rowA.alloc(E);
    auxA.alloc(E);
    rowB.alloc(E);
    auxB.alloc(E);
    if (task == 1) {
      eyA.alloc(E);
      eyB.alloc(E);
    }
    DevBuf<uint32_t> overflow(1);
    DTE_HIP(hipMemsetAsync(overflow.p, 0, 4, s));
    hipLaunchKernelGGL(k_boot_compact, dim3(grid_for(N, kBlock, kBootCompactBlocks), T), dim3(kBlock), 0, s, mult.p,
                       pos.p, N,
                       label, yfix, C, rowA.p, auxA.p, eyA.p, root_cls.p, root_sum.p, root_hash.p,
                       overflow.p);
    if (read_one(overflow.p, s)) fail(DTE_ERR_INTERNAL, "bootstrap multiplicity above 65535");
  } else {
    E = 0;
    for (int t = 0; t < T; ++t) {
      root_start[t] = E;
      root_m[t] = tN[t];
      E += tN[t];
    }
    rowA.alloc(E);
    auxA.alloc(E);
    rowB.alloc(E);
    auxB.alloc(E);
    if (task == 1) {
      eyA.alloc(E);
      eyB.alloc(E);
    }
    if (J) {
      // joint build: each forest's trees hold its own rows, entries of the
      // forests one after another
      std::vector<uint32_t> rc((size_t)T * std::max(C, 1), 0u);
      std::vector<int64_t> rs(T, 0);
      std::vector<uint64_t> rh(T, 0);
      DevBuf<uint32_t> c1(std::max(C, 1));
      DevBuf<int64_t> s1(1);
      DevBuf<uint64_t> h1(1);
      for (const JointForest& jf : J->f) {
        const int64_t off = root_start[jf.t0];
        const uint32_t* lab = jf.fd->label.p;
        const int64_t* yf = jf.fd->yfix.p;
        hipLaunchKernelGGL(k_identity_entries, dim3(grid_for(jf.N, kBlock, 4096), jf.T), dim3(kBlock), 0, s, jf.N,
                           jf.T, lab, yf, rowA.p + off, auxA.p + off, task == 1 ? eyA.p + off : nullptr);
        DTE_HIP(hipMemsetAsync(c1.p, 0, c1.bytes(), s));
        DTE_HIP(hipMemsetAsync(s1.p, 0, 8, s));
        DTE_HIP(hipMemsetAsync(h1.p, 0, 8, s));
        hipLaunchKernelGGL(k_root_stats_identity, dim3(grid_for(jf.N, kBlock, 4096)), dim3(kBlock), 0, s, jf.N, lab,
                           yf, C, c1.p, s1.p, h1.p);
        std::vector<uint32_t> c1h(std::max(C, 1));
        int64_t s1h;
        uint64_t h1h;
        DTE_HIP(hipMemcpyAsync(c1h.data(), c1.p, c1.bytes(), hipMemcpyDeviceToHost, s));
        DTE_HIP(hipMemcpyAsync(&s1h, s1.p, 8, hipMemcpyDeviceToHost, s));
        DTE_HIP(hipMemcpyAsync(&h1h, h1.p, 8, hipMemcpyDeviceToHost, s));
        DTE_HIP(hipStreamSynchronize(s));
        for (int q = 0; q < jf.T; ++q) {
          const int t = jf.t0 + q;
          rs[t] = s1h;
          rh[t] = h1h;
          for (int c = 0; c < std::max(C, 1); ++c) rc[(size_t)t * std::max(C, 1) + c] = c1h[c];
        }
      }
      DTE_HIP(hipMemcpyAsync(root_cls.p, rc.data(), 4 * rc.size(), hipMemcpyHostToDevice, s));
      DTE_HIP(hipMemcpyAsync(root_sum.p, rs.data(), 8 * T, hipMemcpyHostToDevice, s));
      DTE_HIP(hipMemcpyAsync(root_hash.p, rh.data(), 8 * T, hipMemcpyHostToDevice, s));
      DTE_HIP(hipStreamSynchronize(s));  // the host vectors above go out of scope
    } else {
    hipLaunchKernelGGL(k_identity_entries, dim3(grid_for(N, kBlock, 4096), T), dim3(kBlock), 0, s, N, T, label,
                       yfix, rowA.p, auxA.p, eyA.p);
    DevBuf<uint32_t> c1(std::max(C, 1));
    DevBuf<int64_t> s1(1);
    DevBuf<uint64_t> h1(1);
    DTE_HIP(hipMemsetAsync(c1.p, 0, c1.bytes(), s));
    DTE_HIP(hipMemsetAsync(s1.p, 0, 8, s));
    DTE_HIP(hipMemsetAsync(h1.p, 0, 8, s));
    hipLaunchKernelGGL(k_root_stats_identity, dim3(grid_for(N, kBlock, 4096)), dim3(kBlock), 0, s, N, label,
                       yfix, C, c1.p, s1.p, h1.p);
    std::vector<uint32_t> c1h(std::max(C, 1));
    int64_t s1h;
    uint64_t h1h;
    DTE_HIP(hipMemcpyAsync(c1h.data(), c1.p, c1.bytes(), hipMemcpyDeviceToHost, s));
    DTE_HIP(hipMemcpyAsync(&s1h, s1.p, 8, hipMemcpyDeviceToHost, s));
    DTE_HIP(hipMemcpyAsync(&h1h, h1.p, 8, hipMemcpyDeviceToHost, s));
    DTE_HIP(hipStreamSynchronize(s));
    std::vector<uint32_t> rc((size_t)T * std::max(C, 1));
    std::vector<int64_t> rs(T, s1h);
    std::vector<uint64_t> rh(T, h1h);
    for (int t = 0; t < T; ++t)
      for (int c = 0; c < std::max(C, 1); ++c) rc[(size_t)t * std::max(C, 1) + c] = c1h[c];
    DTE_HIP(hipMemcpyAsync(root_cls.p, rc.data(), 4 * rc.size(), hipMemcpyHostToDevice, s));
    DTE_HIP(hipMemcpyAsync(root_sum.p, rs.data(), 8 * T, hipMemcpyHostToDevice, s));
    DTE_HIP(hipMemcpyAsync(root_hash.p, rh.data(), 8 * T, hipMemcpyHostToDevice, s));
    }
  }

  // ---- level-major node arrays (capacity: a tree has at most 2m-1 nodes).
  const int64_t cap = 2 * E + T;
  DevBuf<int32_t> o_tree(cap), o_depth(cap), o_count(cap), o_m(cap), o_feature(cap);
  DevBuf<uint32_t> o_id(cap);
  DevBuf<uint32_t> o_cls(task == 0 ? (size_t)cap * C : 1);
  DevBuf<int64_t> o_sum(task == 1 ? cap : 1);
  DevBuf<uint64_t> o_hash(cap);
  DevBuf<float> o_value(cap), o_thr(cap), o_score(cap);
  DevBuf<int64_t> o_left(cap);

  // roots
  {
    std::vector<int32_t> tr(T), dep(T, 0), cnt(T), mm(T);
    std::vector<uint32_t> idv(T, 0);
    std::vector<int64_t> st(T);
    std::vector<uint8_t> op(T);
    for (int t = 0; t < T; ++t) {
      tr[t] = t;
      cnt[t] = (int32_t)tN[t];
      mm[t] = (int32_t)root_m[t];
      st[t] = root_start[t];
      op[t] = (tN[t] > tree_min_node(t) && 0 < tree_max_depth(t)) ? 1 : 0;
    }
    DTE_HIP(hipMemcpyAsync(o_tree.p, tr.data(), 4 * T, hipMemcpyHostToDevice, s));
    DTE_HIP(hipMemcpyAsync(o_depth.p, dep.data(), 4 * T, hipMemcpyHostToDevice, s));
    DTE_HIP(hipMemcpyAsync(o_count.p, cnt.data(), 4 * T, hipMemcpyHostToDevice, s));
    DTE_HIP(hipMemcpyAsync(o_m.p, mm.data(), 4 * T, hipMemcpyHostToDevice, s));
    DTE_HIP(hipMemcpyAsync(o_id.p, idv.data(), 4 * T, hipMemcpyHostToDevice, s));
    if (task == 0)
      DTE_HIP(hipMemcpyAsync(o_cls.p, root_cls.p, 4 * (size_t)T * C, hipMemcpyDeviceToDevice, s));
    else
      DTE_HIP(hipMemcpyAsync(o_sum.p, root_sum.p, 8 * (size_t)T, hipMemcpyDeviceToDevice, s));
    DTE_HIP(hipMemcpyAsync(o_hash.p, root_hash.p, 8 * (size_t)T, hipMemcpyDeviceToDevice, s));
    // root value (regression mean)
    if (task == 1) {
      std::vector<int64_t> rs(T);
      DTE_HIP(hipMemcpyAsync(rs.data(), root_sum.p, 8 * T, hipMemcpyDeviceToHost, s));
      DTE_HIP(hipStreamSynchronize(s));
      std::vector<float> rv(T);
      for (int t = 0; t < T; ++t)
        rv[t] = (float)std::ldexp((double)rs[t] / (double)tN[t], -(J ? J->f[tf[t]].fd->shift : shift));
      DTE_HIP(hipMemcpyAsync(o_value.p, rv.data(), 4 * T, hipMemcpyHostToDevice, s));
    } else {
      DTE_HIP(hipMemsetAsync(o_value.p, 0, 4 * T, s));
    }
    // working arrays of level 0
  }
  DevBuf<int64_t> w_start(T), w_start_next;
  DevBuf<uint8_t> w_open(T), w_open_next;
  {
    std::vector<int64_t> st(root_start);
    std::vector<uint8_t> op(T);
    for (int t = 0; t < T; ++t) op[t] = (tN[t] > tree_min_node(t) && 0 < tree_max_depth(t)) ? 1 : 0;
    DTE_HIP(hipMemcpyAsync(w_start.p, st.data(), 8 * T, hipMemcpyHostToDevice, s));
    DTE_HIP(hipMemcpyAsync(w_open.p, op.data(), T, hipMemcpyHostToDevice, s));
  }
  DevBuf<uint32_t> lf(T + 1), base(T), lf_next(T + 1), base_next(T);
  {
    std::vector<uint32_t> l(T + 1), b(T, 0);
    for (int t = 0; t <= T; ++t) l[t] = t;
    DTE_HIP(hipMemcpyAsync(lf.p, l.data(), 4 * (T + 1), hipMemcpyHostToDevice, s));
    DTE_HIP(hipMemcpyAsync(base.p, b.data(), 4 * T, hipMemcpyHostToDevice, s));
  }

  // per-level scratch
  DevBuf<uint32_t> flag32, open_pos, open_idx, nchild, cpos, retry_flag, retry_pos, retry_list;
  DevBuf<float> res_score, res_thr, ch_score, ch_thr;
  DevBuf<int32_t> res_feat, ch_feat, l_m, l_count;
  DevBuf<uint32_t> l_cls;
  DevBuf<int64_t> l_sum;
  DevBuf<uint64_t> l_hash;
  DevBuf<uint8_t> split_ok;
  DevBuf<uint32_t> scr_key, scr_aux;
  DevBuf<int64_t> scr_y;
  int64_t scr_stride = 0;

  int num_cu = 256;
  {
    int dev;
    DTE_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    DTE_HIP(hipGetDeviceProperties(&prop, dev));
    num_cu = prop.multiProcessorCount;
  }
  const int search_grid_cap = num_cu * 4;

  SearchFn search = pick_search(task, C);
  BigFn big = pick_big(task, C);
  BigState bigst;
  PartFn partition = pick_part(task, C);
  PartWaveFn partition_wave = pick_part_wave(task, C);
  PartWaveFn partition_seg16 = pick_part_seg<16>(task, C);
  PartWaveFn partition_seg32 = pick_part_seg<32>(task, C);
  PartChunkFn partition_chunks = pick_part_chunks(task, C);
  // host-sized levels: nodes above kPartBig entries partition by chunks
  constexpr Cuts kPartCutsHost{{16, 32, 512, kPartBig, 0x7FFFFFFF, 0}, 5};
  DevBuf<uint32_t> pch_n, pch_off, pch_ml, pch_acc_ml, pch_acc_cls;
  DevBuf<unsigned long long> pch_acc64;

  // Size tiers of the split search: wave per item up to wave_cutoff entries,
  // block per item above (the block kernel itself switches to the bucketed
  // path above its LDS capacity).
  // Nodes above the workgroup's LDS capacity go through the chunk-parallel
  // pipeline (the plain driver keeps one workgroup per node throughout).
  const int cut_team = p.cut_team ? p.cut_team : kDefCutTeam;
  const int cut_block = std::max(cut_team, std::min(kCap, p.cut_block ? p.cut_block : kDefCutBlock));
  bigst.target = p.bucket_target ? p.bucket_target : kDefBucketTarget;
  Cuts cuts{};
  if (algorithm == DTE_RANDOM_FOREST) {
    const int c16 = (task == 1 || C <= 16) ? 16 : 0;
    const int c32 = (task == 1 || C <= 32) ? 32 : 0;
    cuts = Cuts{{c16, std::max(c16, c32), 64, cut_team, cut_block, 0x7FFFFFFF}, 6};
  } else {
    // segment tiers need a lane per class
    const int c16 = (task == 1 || C <= 16) ? 16 : 0;
    const int c32 = (task == 1 || C <= 32) ? 32 : 0;
    const int ce = p.cut_ert ? p.cut_ert : kDefCutErt;
    cuts = Cuts{{c16, std::max(c16, c32), 64, std::min(ce, kErtTeamIpt * kWave), ce, 0x7FFFFFFF}, 6};
  }
  DevBuf<uint32_t> tier_list[kMaxTiers];
  DevBuf<unsigned int> tier_cnt(kMaxTiers);
  auto tier_lists = [&]() {
    TierLists tl{};
    for (int k = 0; k < kMaxTiers; ++k) tl.l[k] = tier_list[k].p;
    return tl;
  };
  const int wave_grid_cap = num_cu * 16;
  LevelArgs lv{};
  int64_t trace_tier[3] = {0, 0, 0};
  // Side stream of the split search's tiers beside the chunk pipeline
  // (created on first use; default priority).
  struct SideStream {
    hipStream_t st = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    hipStream_t get() {
      if (!st) {
        DTE_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        DTE_HIP(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
        DTE_HIP(hipEventCreateWithFlags(&join, hipEventDisableTiming));
      }
      return st;
    }
    ~SideStream() {
      if (st) {
        (void)hipStreamSynchronize(st);
        (void)hipEventDestroy(fork);
        (void)hipEventDestroy(join);
        (void)hipStreamDestroy(st);
      }
    }
  } side;
  auto run_phase = [&](ItemArgs it) {
    for (auto& tl : tier_list) tl.reserve(it.n_groups);
    DTE_HIP(hipMemsetAsync(tier_cnt.p, 0, 4 * kMaxTiers, s));
    hipLaunchKernelGGL(k_tier_split, dim3(std::min(grid_for(it.n_groups), kTierGrid)), dim3(kBlock), 0, s, it.list, it.n_groups,
                       lv.open_idx, lv.m, cuts, tier_lists(), tier_cnt.p, nullptr, nullptr);
    unsigned int cnt[kMaxTiers];
    DTE_HIP(hipMemcpyAsync(cnt, tier_cnt.p, 4 * kMaxTiers, hipMemcpyDeviceToHost, s));
    DTE_HIP(hipStreamSynchronize(s));
    Timer tm("split_search", 0.0, s);
    const bool rf = algorithm == DTE_RANDOM_FOREST;
    // The chunk pipeline (random-forest nodes above cut_block) runs on the
    // library stream with host round trips between its batches; the other
    // tiers touch other nodes and run beside it on a side stream.
    bool others = false;
    for (int tier = 0; tier < 5; ++tier) others = others || cnt[tier] > 0;
    const bool overlap = rf && cnt[5] > 0 && others;
    hipStream_t ts = s;
    if (overlap) {
      ts = side.get();
      DTE_HIP(hipEventRecord(side.fork, s));
      DTE_HIP(hipStreamWaitEvent(ts, side.fork, 0));
    }
    for (int tier = 0; tier < cuts.n; ++tier) {
      if (!cnt[tier]) continue;
      ItemArgs a = it;
      a.list = tier_list[tier].p;
      a.n_groups = cnt[tier];
      a.bytes = it.bytes ? it.bytes + tier : nullptr;
      trace_tier[rf ? (tier <= 2 ? 0 : (tier == 3 ? 1 : 2)) : (tier <= 2 ? 0 : (tier <= 4 ? 1 : 2))] += cnt[tier];
      if (rf && tier == 5) continue;  // below, after the side launches
      int grid;
      if (tier <= 3) {
        // node-major tiers: 4 / 2 / 1 / 1 nodes per wave
        const int npw = tier == 0 ? 4 : (tier == 1 ? 2 : 1);
        grid = (int)std::min<int64_t>((cnt[tier] + npw * kWaves - 1) / (npw * kWaves), wave_grid_cap);
      } else if (rf && tier == 4) {
        grid = (int)std::min<int64_t>(cnt[tier], search_grid_cap);  // one node per workgroup
      } else {
        const int64_t items = (int64_t)cnt[tier] * it.per_group;
        const bool whole = tier == (rf ? 4 : 5);
        const int per_block = whole ? 1 : kWaves;
        const int cap = whole ? search_grid_cap : wave_grid_cap;
        grid = (int)std::min<int64_t>((items + per_block - 1) / per_block, cap);
      }
      Timer tk(search_family(algorithm, tier), 0.0, ts);
      search(algorithm, tier, grid, ts, d, lv, a);
    }
    if (rf && cnt[5]) {
      ItemArgs a = it;
      a.list = tier_list[5].p;
      a.n_groups = cnt[5];
      a.bytes = it.bytes ? it.bytes + 5 : nullptr;
      Timer tk(search_family(algorithm, 5), 0.0, s);
      big(bigst, d, lv, a, tier_list[5].p, cnt[5], s, num_cu);
    }
    if (overlap) {
      DTE_HIP(hipEventRecord(side.join, ts));
      DTE_HIP(hipStreamWaitEvent(s, side.join, 0));
    }
  };
  // Plain launch structure: chunks of at most kPlainChunk open nodes, one
  // workgroup per node, a host round trip after every chunk.
  DevBuf<uint32_t> iota_list;
  auto identity_list = [&](int64_t n) {
    iota_list.reserve(n);
    hipLaunchKernelGGL(k_iota, dim3(grid_for(n)), dim3(kBlock), 0, s, iota_list.p, n);
    return (const uint32_t*)iota_list.p;
  };
  auto plain_phase = [&](ItemArgs it) {
    const uint32_t* lst = it.list ? it.list : identity_list(it.n_groups);
    const int tier = algorithm == DTE_RANDOM_FOREST ? 6 : 5;  // workgroup per node, any size
    Timer tm("split_search", 0.0, s);
    for (int64_t c0 = 0; c0 < it.n_groups; c0 += kPlainChunk) {
      ItemArgs a = it;
      a.list = lst + c0;
      a.n_groups = std::min<int64_t>(kPlainChunk, it.n_groups - c0);
      a.node_major = 1;
      a.bytes = it.bytes ? it.bytes + tier : nullptr;
      Timer tk(search_family(algorithm, tier), 0.0, s);
      search(algorithm, tier, (int)a.n_groups, s, d, lv, a);
      DTE_HIP(hipStreamSynchronize(s));
    }
  };
  const bool trace = std::getenv("DTE_TRACE") != nullptr;

  int64_t off_level = 0;
  int64_t n_level = T;
  int64_t max_m = 0;
  for (int t = 0; t < T; ++t) max_m = std::max(max_m, root_m[t]);
  int64_t big_entries = max_m > cut_block ? E : 0;  // entries in nodes above cut_block

  uint32_t* cur_row = rowA.p;
  uint32_t* cur_aux = auxA.p;
  int64_t* cur_ey = eyA.p;
  uint32_t* nxt_row = rowB.p;
  uint32_t* nxt_aux = auxB.p;
  int64_t* nxt_ey = eyB.p;

  // ---- device-described level steps.  Once every open node of a level
  // fits the device-sized tiers, all further levels run from two captured
  // hipGraphs (even / odd: the ping-pong buffers swap every level) whose
  // kernels read the level's offset and counts from a DevLevel in device
  // memory and run on capped grids; counters are cleared by a kernel.  The
  // buffers are sized once for the largest level possible: a level has at
  // most one node per entry, and at most T * N / (min_node + 1) open nodes.
  DevBuf<unsigned long long> graph_bytes(kMaxTiers + 1);  // search tiers, partition
  int64_t graph_replays = 0;
  auto device_levels = [&](int depth0) {
    int64_t min1 = (int64_t)std::max(p.min_node_size, 0) + 1, W = 0;
    for (int t = 0; t < T; ++t) {
      W += tN[t];
      min1 = std::min<int64_t>(min1, (int64_t)std::max(tree_min_node(t), 0) + 1);
    }
    const int64_t NO = std::max<int64_t>(1, std::min<int64_t>(E, W / min1));
    const int64_t NL = std::max<int64_t>(n_level, std::min<int64_t>(E, 2 * NO));
    auto fit_to = [](auto& b, int64_t n) {
      if ((int64_t)b.n < n) b.alloc((size_t)n);
    };
    {  // the current level's working arrays keep their contents
      DevBuf<int64_t> st((size_t)NL + 1);
      DevBuf<uint8_t> op((size_t)NL + 1);
      DTE_HIP(hipMemcpyAsync(st.p, w_start.p, 8 * n_level, hipMemcpyDeviceToDevice, s));
      DTE_HIP(hipMemcpyAsync(op.p, w_open.p, n_level, hipMemcpyDeviceToDevice, s));
      w_start = std::move(st);
      w_open = std::move(op);
    }
    fit_to(w_start_next, NL + 1);
    fit_to(w_open_next, NL + 1);
    fit_to(open_pos, NL + 1);
    fit_to(open_idx, NO + 1);
    fit_to(res_score, NO * D);
    fit_to(res_thr, NO * D);
    fit_to(res_feat, NO * D);
    fit_to(ch_score, NO);
    fit_to(ch_thr, NO);
    fit_to(ch_feat, NO);
    fit_to(retry_flag, NO);
    fit_to(retry_pos, NO + 1);
    fit_to(retry_list, NO + 1);
    fit_to(l_m, NO);
    fit_to(l_count, NO);
    fit_to(l_cls, task == 0 ? NO * C : 1);
    fit_to(l_sum, NO);
    fit_to(l_hash, NO);
    fit_to(split_ok, NO);
    for (auto& tl : tier_list) fit_to(tl, NO);
    fit_to(nchild, NL);
    fit_to(cpos, NL + 1);
    // scratch slabs of the workgroup-per-item kernel for nodes above
    // cut_block (random forest), one per workgroup of its grid
    const bool big_tier = algorithm == DTE_RANDOM_FOREST && max_m > cut_block;
    if (big_tier && scr_stride < max_m) {
      scr_key.alloc((size_t)search_grid_cap * max_m);
      scr_aux.alloc((size_t)search_grid_cap * max_m);
      if (task == 1) scr_y.alloc((size_t)search_grid_cap * max_m);
      scr_stride = max_m;
    }
    DevBuf<uint32_t> dsums(kDScanBlocks);
    DevBuf<DevLevel> dlv(1);
    DevLevel h{};
    h.off = off_level;
    h.n = n_level;
    h.depth = depth0;
    DTE_HIP(hipMemcpyAsync(dlv.p, &h, sizeof(h), hipMemcpyHostToDevice, s));
    DTE_HIP(hipMemsetAsync(graph_bytes.p, 0, graph_bytes.bytes(), s));
    DevLevel* L = dlv.p;
    const bool rf = algorithm == DTE_RANDOM_FOREST;
    constexpr int kElemGrid = 1024;
    // Extra trees on small data (at most kErtMergeEntries entries over all
    // trees): one launch for all search tiers of a phase.  Its levels are
    // small and latency-bound, and side by side the tiers finish sooner
    // (4,601 x 57, 100 trees: 12.3k -> 15.9k trees/s); on large data the
    // merged kernel's register budget (the largest tier's) costs more
    // occupancy than the overlap gains (64 forests: 8.2k -> 6.9k).
    const bool small_data = E <= kErtMergeEntries;
    const bool ert_merged = !rf && small_data;
    const ErtAllFn ert_all = pick_ert_all(task, C);
    // likewise the partition tiers of a step (both algorithms)
    const PartAllFn partition_all = pick_part_all(task, C);

    auto dscan = [&](const auto* in, const int64_t* n64, const unsigned int* n32, uint32_t* out,
                     unsigned int* total) {
      using In = std::remove_cv_t<std::remove_pointer_t<decltype(in)>>;
      hipLaunchKernelGGL(k_dscan_sums<In>, dim3(kDScanBlocks), dim3(kBlock), 0, s, in, n64, n32, dsums.p);
      hipLaunchKernelGGL(k_dscan_apply<In>, dim3(kDScanBlocks), dim3(kBlock), 0, s, in, n64, n32, dsums.p, out,
                         total);
    };
    // One level step; par = 1 swaps the ping-pong buffers.
    auto enqueue = [&](int par) {
      int64_t* cs = par ? w_start_next.p : w_start.p;
      int64_t* ns = par ? w_start.p : w_start_next.p;
      uint8_t* co = par ? w_open_next.p : w_open.p;
      uint8_t* no = par ? w_open.p : w_open_next.p;
      uint32_t* clf = par ? lf_next.p : lf.p;
      uint32_t* nlf = par ? lf.p : lf_next.p;
      uint32_t* cbase = par ? base_next.p : base.p;
      uint32_t* nbase = par ? base.p : base_next.p;
      uint32_t* crow = par ? nxt_row : cur_row;
      uint32_t* nrw = par ? cur_row : nxt_row;
      uint32_t* caux = par ? nxt_aux : cur_aux;
      uint32_t* nax = par ? cur_aux : nxt_aux;
      int64_t* cey = par ? nxt_ey : cur_ey;
      int64_t* ney = par ? cur_ey : nxt_ey;
      // open list
      dscan((const uint8_t*)co, &L->n, nullptr, open_pos.p, &L->n_open);
      hipLaunchKernelGGL(k_open_list_dev, dim3(kElemGrid), dim3(kBlock), 0, s, co, open_pos.p, (const DevLevel*)L,
                         open_idx.p);
      LevelArgs lvd{};
      lvd.tree = o_tree.p;
      lvd.id = o_id.p;
      lvd.count = o_count.p;
      lvd.m = o_m.p;
      lvd.cls = task == 0 ? o_cls.p : nullptr;
      lvd.sum = task == 1 ? o_sum.p : nullptr;
      lvd.start = cs;
      lvd.open_idx = open_idx.p;
      lvd.tree_global = tree_global.p;
      lvd.row = crow;
      lvd.aux = caux;
      lvd.ey = cey;
      lvd.off_dev = &L->off;
      ItemArgs it{};
      it.D = D;
      it.res_score = res_score.p;
      it.res_thr = res_thr.p;
      it.res_feat = res_feat.p;
      it.scr_key = scr_key.p;
      it.scr_aux = scr_aux.p;
      it.scr_y = scr_y.p;
      it.scr_stride = scr_stride;
      // split search of one phase over the tiers; nodes above cut_block
      // (tier 5) take one workgroup per item through its scratch slab
      auto phase = [&](const uint32_t* list, const unsigned int* n_dev, unsigned int* cnt, int per_group, int j0) {
        hipLaunchKernelGGL(k_tier_split, dim3(kTierGrid), dim3(kBlock), 0, s, list, (int64_t)0,
                           (const uint32_t*)open_idx.p, (const int32_t*)o_m.p, cuts, tier_lists(), cnt, n_dev,
                           (const int64_t*)&L->off);
        if (ert_merged) {
          ErtTiers ts{};
          int64_t acc = 0;
          for (int k = 0; k < 6; ++k) {
            ts.first[k] = acc;
            acc += k <= 4 ? wave_grid_cap : search_grid_cap;
            ts.list[k] = tier_list[k].p;
            ts.n_dev[k] = cnt + k;
            ts.bytes[k] = graph_bytes.p + k;
          }
          ts.first[6] = acc;
          ItemArgs a = it;
          a.per_group = per_group;
          a.j0 = j0;
          ert_all(acc, s, d, lvd, a, ts);
          return;
        }
        for (int tier = 0; tier < cuts.n; ++tier) {
          if (rf && tier == 5 && !big_tier) continue;  // no node above cut_block in these levels
          ItemArgs a = it;
          a.list = tier_list[tier].p;
          a.n_groups = 0;
          a.n_dev = cnt + tier;
          a.per_group = per_group;
          a.j0 = j0;
          a.bytes = graph_bytes.p + tier;
          const int grid = tier <= (rf ? 3 : 4) ? wave_grid_cap : search_grid_cap;
          search(algorithm, rf && tier == 5 ? 6 : tier, grid, s, d, lvd, a);
        }
      };
      phase(nullptr, &L->n_open, L->tier, D0, 0);
      DecideArgs da{};
      da.D0 = D0;
      da.R = R;
      da.D = D;
      da.res_score = res_score.p;
      da.res_thr = res_thr.p;
      da.res_feat = res_feat.p;
      da.ch_score = ch_score.p;
      da.ch_thr = ch_thr.p;
      da.ch_feat = ch_feat.p;
      da.retry_flag = retry_flag.p;
      da.n_open_dev = &L->n_open;
      hipLaunchKernelGGL(k_decide1, dim3(kElemGrid), dim3(kBlock), 0, s, d, lvd, da, algorithm);
      if (R > 0) {
        dscan(retry_flag.p, nullptr, &L->n_open, retry_pos.p, &L->n_retry);
        hipLaunchKernelGGL(k_compact_list_dev, dim3(kElemGrid), dim3(kBlock), 0, s, (const uint32_t*)retry_flag.p,
                           (const uint32_t*)retry_pos.p, (const unsigned int*)&L->n_open, retry_list.p);
        phase(retry_list.p, &L->n_retry, L->rtier, R, D0);
        da.retry_list = retry_list.p;
        da.n_retry_dev = &L->n_retry;
        hipLaunchKernelGGL(k_decide2, dim3(kElemGrid), dim3(kBlock), 0, s, d, lvd, da);
      }
      // partition
      PartArgs pa{};
      pa.ch_feat = ch_feat.p;
      pa.ch_thr = ch_thr.p;
      pa.nrow = nrw;
      pa.naux = nax;
      pa.ney = ney;
      pa.l_m = l_m.p;
      pa.l_count = l_count.p;
      pa.l_cls = l_cls.p;
      pa.l_sum = l_sum.p;
      pa.l_hash = l_hash.p;
      pa.split_ok = split_ok.p;
      pa.bytes = graph_bytes.p + kMaxTiers;
      hipLaunchKernelGGL(k_tier_split, dim3(kTierGrid), dim3(kBlock), 0, s, (const uint32_t*)nullptr, (int64_t)0,
                         (const uint32_t*)open_idx.p, (const int32_t*)o_m.p, kPartCuts, tier_lists(), L->ptier,
                         (const unsigned int*)&L->n_open, (const int64_t*)&L->off);
      PartArgs p0 = pa, p1 = pa, p2 = pa, p3 = pa;
      p0.n_dev = L->ptier + 0;
      p1.n_dev = L->ptier + 1;
      p2.n_dev = L->ptier + 2;
      p3.n_dev = L->ptier + 3;
      p3.list = tier_list[3].p;
      if (small_data) {
        PartTiers pt{};
        const int64_t g[4] = {wave_grid_cap, wave_grid_cap, wave_grid_cap, search_grid_cap};
        int64_t acc = 0;
        for (int k = 0; k < 4; ++k) {
          pt.first[k] = acc;
          acc += g[k];
          pt.list[k] = tier_list[k].p;
          pt.n_dev[k] = L->ptier + k;
        }
        pt.first[4] = acc;
        pt.n_bound = NO;
        partition_all(acc, s, d, lvd, pa, pt);
      } else {
        partition_seg16(wave_grid_cap, s, d, lvd, p0, tier_list[0].p, NO);  // grid from the bound; count from p0.n_dev
        partition_seg32(wave_grid_cap, s, d, lvd, p1, tier_list[1].p, NO);
        partition_wave(wave_grid_cap, s, d, lvd, p2, tier_list[2].p, NO);
        partition(search_grid_cap, s, d, lvd, p3);
      }
      // children
      ChildArgs ca{};
      ca.lvl = L;
      ca.open = co;
      ca.open_pos = open_pos.p;
      ca.split_ok = split_ok.p;
      ca.ch_feat = ch_feat.p;
      ca.ch_thr = ch_thr.p;
      ca.ch_score = ch_score.p;
      ca.l_m = l_m.p;
      ca.l_count = l_count.p;
      ca.l_cls = l_cls.p;
      ca.l_sum = l_sum.p;
      ca.l_hash = l_hash.p;
      ca.nchild = nchild.p;
      ca.cpos = cpos.p;
      ca.lf = clf;
      ca.base = cbase;
      ca.o_tree = o_tree.p;
      ca.o_id = o_id.p;
      ca.o_depth = o_depth.p;
      ca.o_count = o_count.p;
      ca.o_m = o_m.p;
      ca.o_cls = task == 0 ? o_cls.p : nullptr;
      ca.o_sum = o_sum.p;
      ca.o_hash = o_hash.p;
      ca.o_value = o_value.p;
      ca.o_feature = o_feature.p;
      ca.o_thr = o_thr.p;
      ca.o_score = o_score.p;
      ca.o_left = o_left.p;
      ca.start = cs;
      ca.n_start = ns;
      ca.n_open = no;
      hipLaunchKernelGGL(k_children_count, dim3(kElemGrid), dim3(kBlock), 0, s, ca);
      dscan(nchild.p, &L->n, nullptr, cpos.p, &L->n_next);
      hipLaunchKernelGGL(k_children_write, dim3(kElemGrid), dim3(kBlock), 0, s, d, ca);
      hipLaunchKernelGGL(k_tree_update_advance, dim3(grid_for(T + 1)), dim3(kBlock), 0, s, T,
                         (const uint32_t*)cpos.p, (const uint32_t*)clf, (const uint32_t*)cbase, nlf, nbase, L);
      DTE_HIP(hipGetLastError());
    };

    struct Graphs {
      hipGraphExec_t e[2] = {nullptr, nullptr};
      ~Graphs() {
        for (auto x : e)
          if (x) (void)hipGraphExecDestroy(x);
      }
    } gr;
    const auto cap_t0 = std::chrono::steady_clock::now();
    for (int par = 0; par < 2; ++par) {
      hipGraph_t g = nullptr;
      DTE_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      try {
        enqueue(par);
      } catch (...) {
        (void)hipStreamEndCapture(s, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      DTE_HIP(hipStreamEndCapture(s, &g));
      const hipError_t e = hipGraphInstantiate(&gr.e[par], g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      DTE_HIP(e);
    }
    if (profiling_enabled())
      profile_host("graph_capture_host",
                   std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - cap_t0).count());
    int par = 0;
    auto replay = [&](int64_t k) {
      for (int64_t i = 0; i < k; ++i) {
        DTE_HIP(hipGraphLaunch(gr.e[par], s));
        par ^= 1;
        ++graph_replays;
      }
    };
    {
      Timer tm("level_graph", 0.0, s);
      int max_depth = p.max_depth;
      for (int t = 0; t < T; ++t) max_depth = std::max(max_depth, tree_max_depth(t));
      const int64_t steps = (int64_t)max_depth - depth0 + 1;  // levels depth0 .. max_depth
      if (steps <= 64) {
        replay(std::max<int64_t>(steps, 1));
      } else {
        // deep trees: batches of 8 steps; the next batch is queued before the
        // previous one's node count is looked at, so the device never waits
        constexpr int kStepBatch = 8;
        int64_t* pin = nullptr;
        DTE_HIP(hipHostMalloc((void**)&pin, 2 * sizeof(int64_t), hipHostMallocDefault));
        hipEvent_t ev[2] = {nullptr, nullptr};
        try {
          DTE_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
          DTE_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
          for (int64_t b = 0;; ++b) {
            replay(kStepBatch);
            DTE_HIP(hipMemcpyAsync(pin + (b & 1), &L->n, sizeof(int64_t), hipMemcpyDeviceToHost, s));
            DTE_HIP(hipEventRecord(ev[b & 1], s));
            if (b > 0) {
              DTE_HIP(hipEventSynchronize(ev[(b - 1) & 1]));
              if (pin[(b - 1) & 1] == 0) break;
            }
          }
        } catch (...) {
          (void)hipStreamSynchronize(s);
          for (auto e : ev)
            if (e) (void)hipEventDestroy(e);
          (void)hipHostFree(pin);
          throw;
        }
        DTE_HIP(hipStreamSynchronize(s));
        for (auto e : ev) (void)hipEventDestroy(e);
        (void)hipHostFree(pin);
      }
    }
    int64_t fin[2];  // DevLevel.off, DevLevel.n
    DTE_HIP(hipMemcpyAsync(fin, dlv.p, sizeof(fin), hipMemcpyDeviceToHost, s));
    DTE_HIP(hipStreamSynchronize(s));
    if (fin[1] != 0) fail(DTE_ERR_INTERNAL, "level graph: nodes left after the last level step");
    off_level = fin[0];
    n_level = 0;
    if (par) {  // the final per-tree counts are in the other ping-pong buffer
      std::swap(lf, lf_next);
      std::swap(base, base_next);
      std::swap(w_start, w_start_next);
      std::swap(w_open, w_open_next);
      std::swap(cur_row, nxt_row);
      std::swap(cur_aux, nxt_aux);
      std::swap(cur_ey, nxt_ey);
    }
  };
  const bool level_graphs = !plain && p.build_mode == DTE_BUILD_BATCHED;
  if (profiling_enabled()) {
    DTE_HIP(hipStreamSynchronize(s));
    profile_host("fit_setup_host", std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - fit_t0).count());
  }

  for (int depth = 0; n_level > 0; ++depth) {
    // Random forests switch to the level graphs once the nodes above
    // cut_block hold at most 1/64 of the entries and none exceeds
    // kGraphMaxM: inside a graph such nodes take one workgroup per
    // (node, draw) with a scratch slab each (the plain driver's kernel).
    if (level_graphs && (algorithm != DTE_RANDOM_FOREST ||
                         (max_m <= kGraphMaxM && big_entries * 64 <= E))) {
      device_levels(depth);
      break;
    }
    double t_level0 = 0;
    if (trace) {
      DTE_HIP(hipStreamSynchronize(s));
      t_level0 = std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
      trace_tier[0] = trace_tier[1] = trace_tier[2] = 0;
    }
    // ---- open list
    flag32.reserve(n_level);
    open_pos.reserve(n_level + 1);
    hipLaunchKernelGGL(k_open_flags, dim3(grid_for(n_level)), dim3(kBlock), 0, s, w_open.p,
                       n_level, flag32.p);
    scan_u32(flag32.p, n_level, open_pos.p, s, scan_tmp);
    const int64_t n_open = read_one(open_pos.p + n_level, s);
    open_idx.reserve(n_open + 1);
    if (n_open > 0)
      hipLaunchKernelGGL(k_open_list, dim3(grid_for(n_level)), dim3(kBlock), 0, s, w_open.p,
                         open_pos.p, n_level, open_idx.p);

    lv = LevelArgs{};
    lv.tree = o_tree.p + off_level;
    lv.id = o_id.p + off_level;
    lv.count = o_count.p + off_level;
    lv.m = o_m.p + off_level;
    lv.cls = task == 0 ? o_cls.p + off_level * C : nullptr;
    lv.sum = task == 1 ? o_sum.p + off_level : nullptr;
    lv.start = w_start.p;
    lv.open_idx = open_idx.p;
    lv.tree_global = tree_global.p;
    lv.row = cur_row;
    lv.aux = cur_aux;
    lv.ey = cur_ey;

    if (n_open > 0) {
      res_score.reserve(n_open * D);
      res_thr.reserve(n_open * D);
      res_feat.reserve(n_open * D);
      ch_score.reserve(n_open);
      ch_thr.reserve(n_open);
      ch_feat.reserve(n_open);
      retry_flag.reserve(n_open);
      retry_pos.reserve(n_open + 1);
      retry_list.reserve(n_open + 1);
      // scratch slabs of the in-kernel bucketed path (nodes above kCap that
      // do not go through the chunk pipeline)
      if (algorithm == DTE_RANDOM_FOREST && max_m > kCap && plain) {
        const int64_t need = (int64_t)search_grid_cap * max_m;
        if (scr_stride < max_m) {
          scr_key.alloc(need);
          scr_aux.alloc(need);
          if (task == 1) scr_y.alloc(need);
          scr_stride = max_m;
        }
      }
      ItemArgs it{};
      it.list = nullptr;
      it.n_groups = n_open;
      it.per_group = D0;
      it.j0 = 0;
      it.D = D;
      it.res_score = res_score.p;
      it.res_thr = res_thr.p;
      it.res_feat = res_feat.p;
      it.bytes = search_bytes.p;
      it.scr_key = scr_key.p;
      it.scr_aux = scr_aux.p;
      it.scr_y = scr_y.p;
      it.scr_stride = scr_stride;
      if (plain) plain_phase(it); else run_phase(it);
      DecideArgs da{};
      da.n_open = n_open;
      da.D0 = D0;
      da.R = R;
      da.D = D;
      da.res_score = res_score.p;
      da.res_thr = res_thr.p;
      da.res_feat = res_feat.p;
      da.ch_score = ch_score.p;
      da.ch_thr = ch_thr.p;
      da.ch_feat = ch_feat.p;
      da.retry_flag = retry_flag.p;
      hipLaunchKernelGGL(k_decide1, dim3(grid_for(n_open)), dim3(kBlock), 0, s, d, lv, da, algorithm);
      if (R > 0) {
        scan_u32(retry_flag.p, n_open, retry_pos.p, s, scan_tmp);
        const int64_t n_retry = read_one(retry_pos.p + n_open, s);
        if (n_retry > 0) {
          hipLaunchKernelGGL(k_compact_list, dim3(grid_for(n_open)), dim3(kBlock), 0, s,
                             retry_flag.p, retry_pos.p, n_open, retry_list.p);
          it.list = retry_list.p;
          it.n_groups = n_retry;
          it.per_group = R;
          it.j0 = D0;
          if (plain) plain_phase(it); else run_phase(it);
          da.retry_list = retry_list.p;
          da.n_retry = n_retry;
          hipLaunchKernelGGL(k_decide2, dim3(grid_for(n_retry)), dim3(kBlock), 0, s, d, lv, da);
        }
      }
      // ---- partition
      l_m.reserve(n_open);
      l_count.reserve(n_open);
      l_cls.reserve(task == 0 ? n_open * C : 1);
      l_sum.reserve(n_open);
      l_hash.reserve(n_open);
      split_ok.reserve(n_open);
      PartArgs pa{};
      pa.n_open = n_open;
      pa.ch_feat = ch_feat.p;
      pa.ch_thr = ch_thr.p;
      pa.nrow = nxt_row;
      pa.naux = nxt_aux;
      pa.ney = nxt_ey;
      pa.l_m = l_m.p;
      pa.l_count = l_count.p;
      pa.l_cls = l_cls.p;
      pa.l_sum = l_sum.p;
      pa.l_hash = l_hash.p;
      pa.split_ok = split_ok.p;
      pa.bytes = part_bytes.p;
      if (plain) {
        const uint32_t* lst = identity_list(n_open);
        Timer tm("partition", 0.0, s);
        for (int64_t c0 = 0; c0 < n_open; c0 += kPlainChunk) {
          pa.list = lst + c0;
          pa.n_open = std::min<int64_t>(kPlainChunk, n_open - c0);
          partition((int)pa.n_open, s, d, lv, pa);
          DTE_HIP(hipStreamSynchronize(s));
        }
      } else {
        for (auto& tl : tier_list) tl.reserve(n_open);
        DTE_HIP(hipMemsetAsync(tier_cnt.p, 0, 4 * kMaxTiers, s));
        hipLaunchKernelGGL(k_tier_split, dim3(std::min(grid_for(n_open), kTierGrid)), dim3(kBlock), 0, s, (const uint32_t*)nullptr,
                           n_open, lv.open_idx, lv.m, kPartCutsHost, tier_lists(), tier_cnt.p, nullptr, nullptr);
        unsigned int pc[kMaxTiers];
        DTE_HIP(hipMemcpyAsync(pc, tier_cnt.p, 4 * kMaxTiers, hipMemcpyDeviceToHost, s));
        DTE_HIP(hipStreamSynchronize(s));
        Timer tm("partition", 0.0, s);
        if (pc[0]) partition_seg16(wave_grid_cap, s, d, lv, pa, tier_list[0].p, pc[0]);
        if (pc[1]) partition_seg32(wave_grid_cap, s, d, lv, pa, tier_list[1].p, pc[1]);
        if (pc[2])
          partition_wave((int)std::min<int64_t>((pc[2] + kWaves - 1) / kWaves, wave_grid_cap), s, d, lv, pa,
                         tier_list[2].p, pc[2]);
        if (pc[3]) {
          // workgroup partition over the remaining open positions
          pa.list = tier_list[3].p;
          pa.n_open = pc[3];
          partition((int)std::min<int64_t>(pc[3], search_grid_cap), s, d, lv, pa);
        }
        if (pc[4]) {
          // big nodes: chunks of kPartChunk entries, a workgroup each
          const int64_t nb = pc[4];
          pch_n.reserve(nb + 1);
          pch_off.reserve(nb + 1);
          hipLaunchKernelGGL(k_part_nchunks, dim3(std::min(grid_for(nb), 1024)), dim3(kBlock), 0, s,
                             (const uint32_t*)tier_list[4].p, nb, lv.open_idx, lv.m, pch_n.p);
          scan_u32(pch_n.p, nb, pch_off.p, s, scan_tmp);
          const int64_t n_chunks = (int64_t)read_one(pch_off.p + nb, s);
          pch_ml.reserve(n_chunks);
          pch_acc_ml.reserve(nb);
          pch_acc64.reserve(3 * nb);
          pch_acc_cls.reserve(task == 0 ? nb * C : 1);
          DTE_HIP(hipMemsetAsync(pch_acc_ml.p, 0, 4 * nb, s));
          DTE_HIP(hipMemsetAsync(pch_acc64.p, 0, 8 * 3 * nb, s));
          if (task == 0) DTE_HIP(hipMemsetAsync(pch_acc_cls.p, 0, 4 * nb * C, s));
          PartChunks pcs{tier_list[4].p, nb, pch_off.p, pch_ml.p, pch_acc_ml.p,
                         pch_acc64.p, pch_acc64.p + nb, pch_acc64.p + 2 * nb, pch_acc_cls.p};
          pa.list = nullptr;
          pa.n_open = n_open;
          partition_chunks(s, d, lv, pa, pcs, n_chunks);
        }
      }
    }

    // ---- children
    nchild.reserve(n_level);
    cpos.reserve(n_level + 1);
    ChildArgs ca{};
    ca.n_level = n_level;
    ca.off_level = off_level;
    ca.off_next = off_level + n_level;
    ca.depth = depth;
    ca.open = w_open.p;
    ca.open_pos = open_pos.p;
    ca.split_ok = split_ok.p;
    ca.ch_feat = ch_feat.p;
    ca.ch_thr = ch_thr.p;
    ca.ch_score = ch_score.p;
    ca.l_m = l_m.p;
    ca.l_count = l_count.p;
    ca.l_cls = l_cls.p;
    ca.l_sum = l_sum.p;
    ca.l_hash = l_hash.p;
    ca.nchild = nchild.p;
    ca.cpos = cpos.p;
    ca.lf = lf.p;
    ca.base = base.p;
    ca.o_tree = o_tree.p;
    ca.o_id = o_id.p;
    ca.o_depth = o_depth.p;
    ca.o_count = o_count.p;
    ca.o_m = o_m.p;
    ca.o_cls = task == 0 ? o_cls.p : nullptr;
    ca.o_sum = o_sum.p;
    ca.o_hash = o_hash.p;
    ca.o_value = o_value.p;
    ca.o_feature = o_feature.p;
    ca.o_thr = o_thr.p;
    ca.o_score = o_score.p;
    ca.o_left = o_left.p;
    ca.start = w_start.p;
    if (n_open == 0) {
      DTE_HIP(hipMemsetAsync(nchild.p, 0, 4 * n_level, s));
    } else {
      hipLaunchKernelGGL(k_children_count, dim3(grid_for(n_level)), dim3(kBlock), 0, s, ca);
    }
    scan_u32(nchild.p, n_level, cpos.p, s, scan_tmp);
    const int64_t n_next = read_one(cpos.p + n_level, s);
    w_start_next.reserve(n_next + 1);
    w_open_next.reserve(n_next + 1);
    ca.n_start = w_start_next.p;
    ca.n_open = w_open_next.p;
    hipLaunchKernelGGL(k_children_write, dim3(grid_for(n_level)), dim3(kBlock), 0, s, d, ca);
    hipLaunchKernelGGL(k_tree_update, dim3(grid_for(T + 1)), dim3(kBlock), 0, s, T, cpos.p, lf.p, base.p,
                       lf_next.p, base_next.p);
    std::swap(lf, lf_next);
    std::swap(base, base_next);
    std::swap(w_start, w_start_next);
    std::swap(w_open, w_open_next);
    std::swap(cur_row, nxt_row);
    std::swap(cur_aux, nxt_aux);
    std::swap(cur_ey, nxt_ey);
    if (trace) {
      DTE_HIP(hipStreamSynchronize(s));
      fprintf(stderr, "[dte] depth %d nodes %lld open %lld tiers %lld/%lld/%lld : %.3f ms\n", depth,
              (long long)n_level, (long long)n_open, (long long)trace_tier[0], (long long)trace_tier[1],
              (long long)trace_tier[2],
              1000.0 * (std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count() - t_level0));
    }
    off_level += n_level;
    n_level = n_next;
    // largest node of the next level (for scratch sizing) and the entries
    // in nodes above cut_block (whether the level graphs can take over)
    if (n_level > 0 && algorithm == DTE_RANDOM_FOREST) {
      DevBuf<unsigned long long> mx(2);
      DTE_HIP(hipMemsetAsync(mx.p, 0, 16, s));
      hipLaunchKernelGGL(k_level_sizes, dim3(std::min(grid_for(n_level), 1024)), dim3(kBlock), 0, s,
                         o_m.p + off_level, n_level, cut_block, mx.p);
      unsigned long long hm[2];
      DTE_HIP(hipMemcpyAsync(hm, mx.p, 16, hipMemcpyDeviceToHost, s));
      DTE_HIP(hipStreamSynchronize(s));
      max_m = (int64_t)hm[0];
      big_entries = (int64_t)hm[1];
    }
  }
  const int64_t n_nodes = off_level;
  if (profiling_enabled()) {
    unsigned long long sb[kMaxTiers];
    DTE_HIP(hipMemcpyAsync(sb, search_bytes.p, 8 * kMaxTiers, hipMemcpyDeviceToHost, s));
    DTE_HIP(hipStreamSynchronize(s));
    double total = 0;
    for (int tier = 0; tier < kMaxTiers; ++tier) {
      if (!sb[tier]) continue;
      profile_add_bytes(search_family(algorithm, tier), (double)sb[tier]);
      total += (double)sb[tier];
    }
    profile_add_bytes("split_search", total);
    profile_add_bytes("partition", (double)read_one(part_bytes.p, s));
    if (graph_replays) {
      unsigned long long gb[kMaxTiers + 1];
      DTE_HIP(hipMemcpyAsync(gb, graph_bytes.p, sizeof(gb), hipMemcpyDeviceToHost, s));
      DTE_HIP(hipStreamSynchronize(s));
      double g_search = 0;
      for (int tier = 0; tier < kMaxTiers; ++tier) g_search += (double)gb[tier];
      // level steps of the graphs: their split-search bytes (partition bytes
      // beside them) over the whole step time
      profile_add_bytes("level_graph", g_search);
      profile_add_bytes("level_graph_partition", (double)gb[kMaxTiers]);
    }
    profile_count("level_graph_replays", graph_replays);
  }

  // ---- canonical tree-major arrays
  std::vector<uint32_t> per_tree(T);
  DTE_HIP(hipMemcpyAsync(per_tree.data(), base.p, 4 * T, hipMemcpyDeviceToHost, s));
  DTE_HIP(hipStreamSynchronize(s));
  std::vector<int64_t> ts(T + 1, 0);
  for (int t = 0; t < T; ++t) ts[t + 1] = ts[t] + per_tree[t];
  if (ts[T] != n_nodes) fail(DTE_ERR_INTERNAL, "node count mismatch after build");

  auto* fo = new Forest();
  int dev;
  DTE_HIP(hipGetDevice(&dev));
  fo->device = dev;
  fo->task = task;
  fo->algorithm = algorithm;
  fo->n_classes = C;
  fo->n_features = J ? J->F : F;
  fo->n_trees = T;
  fo->fixed_shift = shift;
  fo->seed = p.seed;
  fo->tree_offset = t0;
  fo->fit_rows = algorithm == DTE_RANDOM_FOREST ? N : 0;
  fo->n_nodes = n_nodes;
  fo->n_leaves = (n_nodes + T) / 2;  // per tree: nodes = 2 splits + 1, leaves = splits + 1
  fo->has_stats = true;
  PoolSuspend own;  // the forest outlives the fit
  fo->tree_start.alloc(T + 1);
  DTE_HIP(hipMemcpyAsync(fo->tree_start.p, ts.data(), 8 * (T + 1), hipMemcpyHostToDevice, s));
  fo->feature.alloc(n_nodes);
  fo->left.alloc(n_nodes);
  fo->depth.alloc(n_nodes);
  fo->count.alloc(n_nodes);
  fo->n_entries.alloc(n_nodes);
  fo->threshold.alloc(n_nodes);
  fo->score.alloc(n_nodes);
  fo->value.alloc(n_nodes);
  fo->rowhash.alloc(n_nodes);
  fo->class_counts.alloc(task == 0 ? (size_t)n_nodes * C : 1);
  fo->sum_fixed.alloc(n_nodes);
  FinalArgs fa{};
  fa.n_nodes = n_nodes;
  fa.C = C;
  fa.tree = o_tree.p;
  fa.id = o_id.p;
  fa.tree_start = fo->tree_start.p;
  fa.feature = o_feature.p;
  fa.thr = o_thr.p;
  fa.score = o_score.p;
  fa.left = o_left.p;
  fa.depth = o_depth.p;
  fa.count = o_count.p;
  fa.m = o_m.p;
  fa.hash = o_hash.p;
  fa.cls = task == 0 ? o_cls.p : nullptr;
  fa.sum = task == 1 ? o_sum.p : nullptr;
  fa.value = o_value.p;
  fa.f_feature = fo->feature.p;
  fa.f_thr = fo->threshold.p;
  fa.f_score = fo->score.p;
  fa.f_left = fo->left.p;
  fa.f_depth = fo->depth.p;
  fa.f_count = fo->count.p;
  fa.f_m = fo->n_entries.p;
  fa.f_hash = fo->rowhash.p;
  fa.f_cls = fo->class_counts.p;
  fa.f_sum = fo->sum_fixed.p;
  fa.f_value = fo->value.p;
  hipLaunchKernelGGL(k_finalize, dim3(grid_for(n_nodes)), dim3(kBlock), 0, s, fa);
  build_predict_nodes(*fo, s);
  DTE_HIP(hipStreamSynchronize(s));
  return fo;
}

// Concatenates tree batches into one forest (the reference's
// AggregateModels, gpudte_algorithm.cc:333-397, on canonical arrays).
Forest* merge_forests(std::vector<std::unique_ptr<Forest>>& parts, hipStream_t s) {
  auto* fo = new Forest();
  std::unique_ptr<Forest> guard(fo);
  const Forest& p0 = *parts[0];
  fo->device = p0.device;
  fo->task = p0.task;
  fo->algorithm = p0.algorithm;
  fo->n_classes = p0.n_classes;
  fo->n_features = p0.n_features;
  fo->fixed_shift = p0.fixed_shift;
  fo->seed = p0.seed;
  fo->tree_offset = p0.tree_offset;
  fo->fit_rows = p0.fit_rows;
  fo->has_stats = true;
  int64_t n = 0;
  int T = 0;
  int64_t leaves = 0;
  for (auto& q : parts) {
    n += q->n_nodes;
    T += q->n_trees;
    leaves = (leaves < 0 || q->n_leaves < 0) ? -1 : leaves + q->n_leaves;
  }
  fo->n_nodes = n;
  fo->n_leaves = leaves;
  fo->n_trees = T;
  const int C = fo->n_classes;
  const size_t W = fo->task == 0 ? (size_t)C : 1;
  PoolSuspend own;  // the forest outlives the fit
  fo->tree_start.alloc(T + 1);
  fo->feature.alloc(n);
  fo->left.alloc(n);
  fo->depth.alloc(n);
  fo->count.alloc(n);
  fo->n_entries.alloc(n);
  fo->threshold.alloc(n);
  fo->score.alloc(n);
  fo->value.alloc(n);
  fo->rowhash.alloc(n);
  fo->class_counts.alloc(fo->task == 0 ? (size_t)n * C : 1);
  fo->sum_fixed.alloc(n);
  int64_t off = 0;
  int t_off = 0;
  for (auto& q : parts) {
    const int64_t m = q->n_nodes;
    auto cp = [&](auto* dst, const auto* src, size_t count) {
      DTE_HIP(hipMemcpyAsync(dst, src, sizeof(*src) * count, hipMemcpyDeviceToDevice, s));
    };
    cp(fo->feature.p + off, q->feature.p, m);
    cp(fo->depth.p + off, q->depth.p, m);
    cp(fo->count.p + off, q->count.p, m);
    cp(fo->n_entries.p + off, q->n_entries.p, m);
    cp(fo->threshold.p + off, q->threshold.p, m);
    cp(fo->score.p + off, q->score.p, m);
    cp(fo->value.p + off, q->value.p, m);
    cp(fo->rowhash.p + off, q->rowhash.p, m);
    cp(fo->sum_fixed.p + off, q->sum_fixed.p, m);
    if (fo->task == 0) cp(fo->class_counts.p + off * W, q->class_counts.p, m * W);
    hipLaunchKernelGGL(k_offset_links, dim3(grid_for(m)), dim3(kBlock), 0, s, q->left.p, m, (int32_t)off,
                       fo->left.p + off);
    hipLaunchKernelGGL(k_offset_i64, dim3(grid_for(q->n_trees + 1)), dim3(kBlock), 0, s, q->tree_start.p,
                       (int64_t)q->n_trees + 1, off, fo->tree_start.p + t_off);
    off += m;
    t_off += q->n_trees;
  }
  build_predict_nodes(*fo, s);
  DTE_HIP(hipStreamSynchronize(s));
  return guard.release();
}

Forest* fit_forest(const float* X_dev_in, const float* y_host, int64_t N, int32_t F, const dte_params& p,
                   hipStream_t s) {
  PoolScope pool(s);  // the fit's temporaries come from the stream's scratch pool
  FitData fd;
  prepare_data(fd, X_dev_in, y_host, N, F, p, s);
  if (p.build_mode != DTE_BUILD_PLAIN) return fit_batch(fd, p.tree_offset, p.n_trees, N, F, p, s, false);
  std::vector<std::unique_ptr<Forest>> parts;
  for (int t0 = 0; t0 < p.n_trees; t0 += kPlainTrees)
    parts.emplace_back(
        fit_batch(fd, p.tree_offset + t0, std::min(kPlainTrees, p.n_trees - t0), N, F, p, s, true));
  if (parts.size() == 1) return parts[0].release();
  return merge_forests(parts, s);
}

// Node range [off, off + n) of a joint build's forest, as a forest of its
// own: child links relative to the range, class counts narrowed from the
// joint width Cs to the forest's C.
struct SliceArgs {
  int64_t off, n;
  int Cs, C;
  const int32_t *feature, *left, *depth, *count, *m;
  const float *thr, *score, *value;
  const uint64_t* hash;
  const uint32_t* cls;
  const int64_t* sum;
  int32_t *o_feature, *o_left, *o_depth, *o_count, *o_m;
  float *o_thr, *o_score, *o_value;
  uint64_t* o_hash;
  uint32_t* o_cls;
  int64_t* o_sum;
};

__global__ void k_slice_nodes(SliceArgs a) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < a.n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t g = a.off + i;
    a.o_feature[i] = a.feature[g];
    const int32_t l = a.left[g];
    a.o_left[i] = l < 0 ? -1 : (int32_t)(l - a.off);
    a.o_depth[i] = a.depth[g];
    a.o_count[i] = a.count[g];
    a.o_m[i] = a.m[g];
    a.o_thr[i] = a.thr[g];
    a.o_score[i] = a.score[g];
    a.o_value[i] = a.value[g];
    a.o_hash[i] = a.hash[g];
    a.o_sum[i] = a.sum[g];
    if (a.cls)
      for (int c = 0; c < a.C; ++c) a.o_cls[i * a.C + c] = a.cls[g * a.Cs + c];
  }
}

std::vector<Forest*> fit_joint(const float* const* X, const float* const* y, const int64_t* N, const int32_t* F,
                               const dte_params* ps, int n, hipStream_t s) {
  PoolScope pool(s);
  const bool prof = profiling_enabled();
  auto now = [] { return std::chrono::steady_clock::now(); };
  auto ms = [](std::chrono::steady_clock::time_point a, std::chrono::steady_clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
  };
  const auto h0 = now();
  const int task = ps[0].task;
  std::vector<std::unique_ptr<FitData>> fds(n);
  std::vector<DataArgs> da(n);
  JointSpec J;
  J.C = 1;
  J.D0 = J.R = J.F = 0;
  int T = 0;
  for (int f = 0; f < n; ++f) {
    fds[f] = std::make_unique<FitData>();
    prepare_data(*fds[f], X[f], y[f], N[f], F[f], ps[f], s);
    const int k = ps[f].n_features_per_node > 0 ? ps[f].n_features_per_node
                                                : (int)std::round(std::log((double)F[f])) + 1;
    da[f] = fds[f]->d;
    da[f].D0 = std::min(k, F[f]);
    da[f].R = k < F[f] ? std::min(10, F[f] - k) : 0;
    J.f.push_back(JointForest{fds[f].get(), N[f], F[f], ps[f].n_trees, T, ps[f]});
    T += ps[f].n_trees;
    if (task == 0) J.C = std::max(J.C, ps[f].n_classes);
    J.D0 = std::max(J.D0, da[f].D0);
    J.R = std::max(J.R, da[f].R);
    J.F = std::max(J.F, F[f]);
  }
  for (auto& a : da) a.C = task == 0 ? J.C : 0;  // joint class width in every forest's data
  DevBuf<DataArgs> dforests(n);
  std::vector<int32_t> fot(T);
  for (const JointForest& jf : J.f)
    for (int q = 0; q < jf.T; ++q) fot[jf.t0 + q] = (int32_t)(&jf - J.f.data());
  DevBuf<int32_t> dfot(T);
  DTE_HIP(hipMemcpyAsync(dforests.p, da.data(), sizeof(DataArgs) * n, hipMemcpyHostToDevice, s));
  DTE_HIP(hipMemcpyAsync(dfot.p, fot.data(), 4 * (size_t)T, hipMemcpyHostToDevice, s));
  J.d = da[0];
  J.d.forests = dforests.p;
  J.d.forest_of_tree = dfot.p;
  dte_params pj = ps[0];
  pj.n_trees = T;
  pj.n_classes = task == 0 ? J.C : 0;
  if (prof) {
    DTE_HIP(hipStreamSynchronize(s));
    profile_host("joint_prepare_host", ms(h0, now()));
  }
  const auto h1 = now();
  std::unique_ptr<Forest> all(fit_batch(*fds[0], 0, T, N[0], F[0], pj, s, false, &J));
  const auto h2 = now();
  if (prof) profile_host("joint_build_host", ms(h1, h2));
  // one forest per input, from its trees' node range
  std::vector<int64_t> ts(T + 1);
  DTE_HIP(hipMemcpyAsync(ts.data(), all->tree_start.p, 8 * (size_t)(T + 1), hipMemcpyDeviceToHost, s));
  DTE_HIP(hipStreamSynchronize(s));
  std::vector<std::unique_ptr<Forest>> out;
  PoolSuspend own;  // the forests outlive the fit
  for (const JointForest& jf : J.f) {
    const int64_t off = ts[jf.t0], nn = ts[jf.t0 + jf.T] - off;
    const int Cf = task == 0 ? jf.p.n_classes : 1;
    auto fo = std::make_unique<Forest>();
    fo->device = all->device;
    fo->task = task;
    fo->algorithm = all->algorithm;
    fo->n_classes = Cf;
    fo->n_features = jf.F;
    fo->n_trees = jf.T;
    fo->fixed_shift = jf.fd->shift;
    fo->seed = jf.p.seed;
    fo->tree_offset = jf.p.tree_offset;
    fo->fit_rows = 0;  // extra trees: no bootstrap record
    fo->n_nodes = nn;
    fo->n_leaves = (nn + jf.T) / 2;
    fo->has_stats = true;
    std::vector<int64_t> tsf(jf.T + 1);
    for (int q = 0; q <= jf.T; ++q) tsf[q] = ts[jf.t0 + q] - off;
    fo->tree_start.alloc(jf.T + 1);
    DTE_HIP(hipMemcpyAsync(fo->tree_start.p, tsf.data(), 8 * (size_t)(jf.T + 1), hipMemcpyHostToDevice, s));
    fo->feature.alloc(nn);
    fo->left.alloc(nn);
    fo->depth.alloc(nn);
    fo->count.alloc(nn);
    fo->n_entries.alloc(nn);
    fo->threshold.alloc(nn);
    fo->score.alloc(nn);
    fo->value.alloc(nn);
    fo->rowhash.alloc(nn);
    fo->class_counts.alloc(task == 0 ? (size_t)nn * Cf : 1);
    fo->sum_fixed.alloc(nn);
    SliceArgs sa{off, nn, all->n_classes, task == 0 ? Cf : 0,
                 all->feature.p, all->left.p, all->depth.p, all->count.p, all->n_entries.p,
                 all->threshold.p, all->score.p, all->value.p, all->rowhash.p,
                 task == 0 ? all->class_counts.p : nullptr, all->sum_fixed.p,
                 fo->feature.p, fo->left.p, fo->depth.p, fo->count.p, fo->n_entries.p,
                 fo->threshold.p, fo->score.p, fo->value.p, fo->rowhash.p, fo->class_counts.p, fo->sum_fixed.p};
    hipLaunchKernelGGL(k_slice_nodes, dim3(grid_for(nn)), dim3(kBlock), 0, s, sa);
    build_predict_nodes(*fo, s);
    DTE_HIP(hipStreamSynchronize(s));  // tsf goes out of scope
    out.push_back(std::move(fo));
  }
  all.reset();
  if (prof) profile_host("joint_split_host", ms(h2, now()));
  std::vector<Forest*> res;
  for (auto& f : out) res.push_back(f.release());
  return res;
}

void build_predict_nodes(Forest& f, hipStream_t s) {
  PoolSuspend own;
  f.pnodes.alloc(f.n_nodes);
  hipLaunchKernelGGL(k_predict_nodes, dim3(grid_for(f.n_nodes)), dim3(kBlock), 0, s, f.n_nodes, f.n_classes,
                     f.task, f.feature.p, f.threshold.p, f.left.p, f.class_counts.p, f.value.p, f.pnodes.p);
}

}  // namespace dte
