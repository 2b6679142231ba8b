"""Known-answer and closed-form checks that pin the CPU oracle's building
blocks (no GPU).

Each check is chosen so that a plausible slip in the oracle -- a wrong
Philox round constant, a transposed counter word, a dropped entropy term, a
mis-signed variance term, an off-by-one boundary -- fails it.
"""
import itertools
import math
from fractions import Fraction

import numpy as np
import pytest

import pyoracle as O

MASK64 = (1 << 64) - 1


# --------------------------------------------------------------- Philox ----
# Random123 kat_vectors for philox4x32 with 10 rounds (Salmon et al. 2011).
PHILOX_KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,want", PHILOX_KAT)
def test_philox_known_answers(ctr, key, want):
    assert tuple(int(v) for v in O.philox(ctr, key)) == want


def _philox_py(ctr, key, rounds=10):
    """Straight transcription of the SC'11 round function, for the counter
    layout check below (independent of the oracle's C code)."""
    M0, M1, W0, W1 = 0xD2511F53, 0xCD9E8D57, 0x9E3779B9, 0xBB67AE85
    c = list(ctr)
    k0, k1 = key
    for r in range(rounds):
        if r:
            k0, k1 = (k0 + W0) & 0xFFFFFFFF, (k1 + W1) & 0xFFFFFFFF
        p0, p1 = M0 * c[0], M1 * c[2]
        c = [((p1 >> 32) ^ c[1] ^ k0) & 0xFFFFFFFF, p1 & 0xFFFFFFFF,
             ((p0 >> 32) ^ c[3] ^ k1) & 0xFFFFFFFF, p0 & 0xFFFFFFFF]
    return c


def test_philox_python_transcription_matches_kat():
    for ctr, key, want in PHILOX_KAT:
        assert tuple(_philox_py(ctr, key)) == want


@pytest.mark.parametrize("seed", [0, 1, 324123, (1 << 63) + 12345])
def test_draw_counter_layout(seed):
    """A draw is word idx%4 of Philox(counter=(idx/4, node, tree, stream),
    key=(seed lo, seed hi))."""
    L = O.lib()
    rng = np.random.default_rng(seed & 0xFFFF)
    for _ in range(40):
        stream = int(rng.choice([0x746F6F62, 0x74616566, 0x00726874]))
        tree, node = int(rng.integers(0, 1 << 20)), int(rng.integers(0, 1 << 30))
        idx = int(rng.integers(0, 1 << 40))
        got = L.oracle_rand_u32(seed, stream, tree, node, idx)
        want = _philox_py(((idx >> 2) & 0xFFFFFFFF, node, tree, stream), (seed & 0xFFFFFFFF, seed >> 32))[idx & 3]
        assert got == want


# --------------------------------------------------------------- x ln x ----
def test_xlnx_matches_libm():
    xs = [0, 1, 2, 3, 4, 5, 127, 128, 129, 255, 256, 257, 1000, 65535, 65536, 10 ** 6, 2 ** 31 - 1, 2 ** 31,
          2 ** 40 + 3, 2 ** 52 - 1]
    rng = np.random.default_rng(1)
    xs += [int(v) for v in np.exp(rng.uniform(0, math.log(2 ** 48), 4000)).astype(np.int64)]
    for x in xs:
        got = O.xlnx(x)
        if x <= 1:
            assert got == 0.0
            continue
        want = x * math.log(x)
        assert abs(got - want) <= 7e-16 * want, (x, got, want)  # both sides carry ~1.5 ulp


def test_xlnx_table_boundaries_are_monotone():
    # mantissa-table boundaries (j changes) must not break monotonicity of x ln x.
    xs = sorted({(2 ** e) * (128 + j) // 128 + d for e in range(8, 40, 3) for j in range(128) for d in (-1, 0, 1)})
    vals = [O.xlnx(x) for x in xs]
    assert all(b > a for a, b in zip(vals, vals[1:]))


# ------------------------------------------------------- information gain ----
def _ig_libm(node, left):
    """Entropy of the node minus the weighted entropy of the two sides, in
    bits, in fp64 with libm log (gpudte.cu:76-117 written out)."""
    node, left = np.asarray(node, np.int64), np.asarray(left, np.int64)
    right = node - left

    def H(c):
        n = c.sum()
        if n == 0:
            return 0.0
        return -sum((v / n) * math.log2(v / n) for v in c if v > 0)

    n, nl, nr = node.sum(), left.sum(), right.sum()
    return H(node) - (nl / n) * H(left) - (nr / n) * H(right)


@pytest.mark.parametrize("node,left,want", [
    ([2, 2], [2, 0], 1.0),          # [[2,0],[0,2]]: a perfect split of two classes
    ([2, 2], [1, 1], 0.0),          # [[1,1],[1,1]]: no information
    ([1, 1, 1, 1], [1, 1, 0, 0], 1.0),
    ([4, 4, 4, 4], [4, 4, 4, 0], 2.0 - 0.75 * math.log2(3)),
    ([8, 0], [3, 0], 0.0),          # a pure node
    ([3, 5, 0], [3, 0, 0], _ig_libm([3, 5, 0], [3, 0, 0])),
])
def test_info_gain_closed_forms(node, left, want):
    assert O.info_gain(node, left) == pytest.approx(want, abs=1e-12)


def test_info_gain_random_counts_vs_libm():
    rng = np.random.default_rng(5)
    for trial in range(3000):
        C = int(rng.integers(2, 21))
        scale = int(rng.choice([3, 30, 3000, 300000, 3000000]))
        node = rng.integers(0, scale, C).astype(np.uint32)
        if node.sum() == 0:
            node[0] = 1
        left = (node * rng.random(C)).astype(np.uint32)
        got = O.info_gain(node, left)
        want = _ig_libm(node, left)
        # fixed-point terms carry 2^-F absolute error each; F >= 20 bits
        # below the largest term, so the bits value is good to ~1e-12.
        assert abs(got - want) <= 1e-11 + 1e-12 * abs(want), (node, left, got, want)


def test_info_gain_uses_fit_wide_shift():
    """A smaller fixed-point shift (a larger fit's) changes the value by no
    more than the term quantisation."""
    node, left = [700, 300, 11], [600, 10, 11]
    base = O.info_gain(node, left)
    for F in (O.ig_shift(10 ** 6, 3), O.ig_shift(10 ** 9, 3)):
        assert O.info_gain(node, left, F) == pytest.approx(base, abs=1e-9)


@pytest.mark.parametrize("C", [2, 3, 10, 20, 64])
def test_ig_shift_cannot_overflow(C):
    for n in [2, 3, 100, 4601, 10 ** 6, 4 * 10 ** 6, 2 ** 32 - 1]:
        F = O.ig_shift(n, C)
        # the largest sum of 2C+2 terms, each <= T(n), stays below 2^62
        assert (2 * C + 2) * (O.xlnx(n) + 1.0) * 2.0 ** F < 2.0 ** 62
        assert F >= 8


# ------------------------------------------------------ variance criterion ----
def test_variance_gain_vs_exact_sse_reduction():
    """n_l n_r/(n_l+n_r) (mean_l - mean_r)^2 equals SSE(parent) - SSE(l) -
    SSE(r); checked in exact rationals on fixed-point targets."""
    rng = np.random.default_rng(9)
    for trial in range(400):
        m = int(rng.integers(2, 40))
        y = rng.normal(0, float(rng.choice([1e-3, 1.0, 1e3])), m)
        w = rng.integers(1, 4, m)
        shift = O.lib().oracle_fixed_shift(float(np.abs(y).max()), int(w.sum()))
        yfix = [int(np.rint(np.ldexp(v, shift))) for v in y]
        cut = int(rng.integers(1, m))

        def sse(idx):
            n = sum(int(w[i]) for i in idx)
            s = sum(Fraction(int(w[i]) * yfix[i]) for i in idx)
            q = sum(Fraction(int(w[i]) * yfix[i] * yfix[i]) for i in idx)
            return q - s * s / n

        L, R = range(cut), range(cut, m)
        exact = (sse(range(m)) - sse(L) - sse(R)) / Fraction(2) ** (2 * shift)
        nl, nr = sum(int(w[i]) for i in L), sum(int(w[i]) for i in R)
        sl, sr = sum(int(w[i]) * yfix[i] for i in L), sum(int(w[i]) * yfix[i] for i in R)
        got = O.variance_gain(nl, sl, nr, sr, shift)
        assert got == pytest.approx(float(exact), rel=1e-12, abs=1e-300)


@pytest.mark.parametrize("max_abs,total", [(1.0, 10), (0.7, 4_000_000), (1e6, 1000), (1e-30, 5), (0.0, 100)])
def test_fixed_shift_bounds(max_abs, total):
    s = O.lib().oracle_fixed_shift(max_abs, total)
    assert s <= 60
    # |y| * 2^s summed over the total weight stays inside int64
    assert max_abs * 2.0 ** s * total < 2.0 ** 63


# ------------------------------------------------------------- midpoint ----
def test_midpoint_separates_consecutive_values():
    rng = np.random.default_rng(3)
    a = rng.standard_normal(20000).astype(np.float32) * np.float32(10) ** rng.integers(-30, 30, 20000).astype(np.float32)
    gaps = rng.integers(1, 4, a.size)
    b = a.copy()
    for i in range(a.size):  # the next 1-3 representable floats above a
        for _ in range(int(gaps[i])):
            b[i] = np.nextafter(b[i], np.float32(np.inf))
    extra = np.array([[-np.finfo(np.float32).max, 0.0], [3.0e38, 3.4e38], [-1e-45, 0.0], [0.0, 1e-45],
                      [1.0, 2.0], [-2.0, -1.0]], np.float32)
    pairs = list(zip(a, b)) + [tuple(r) for r in extra]
    for x, z in pairs:
        m = np.float32(O.lib().oracle_midpoint(float(x), float(z)))
        assert x < m <= z, (x, z, m)


def test_midpoint_is_the_fp32_average_when_exact():
    for x, z in [(1.0, 2.0), (-3.0, 5.0), (0.25, 0.5)]:
        assert O.lib().oracle_midpoint(x, z) == np.float32((np.float32(x) + np.float32(z)) * np.float32(0.5))


# -------------------------------------------------- best split, brute force ----
def _boundary_gains(values, labels, weights, C):
    """Every boundary between distinct values with its threshold, scored on
    the realised partition {x < thr} with libm fp64."""
    node = np.zeros(C, np.int64)
    for lab, w in zip(labels, weights):
        node[lab] += w
    uniq = sorted(set(float(v) for v in values))
    out = []
    for lo, hi in zip(uniq, uniq[1:]):
        thr = float(np.float32(O.lib().oracle_midpoint(lo, hi)))
        left = np.zeros(C, np.int64)
        for v, lab, w in zip(values, labels, weights):
            if v < thr:
                left[lab] += w
        out.append((_ig_libm(node, left), thr))
    return out


TIE = 1e-9  # gains closer than this are equal up to the fixed-point rounding


def test_rf_best_split_brute_force():
    """The chosen (score, threshold) is the best boundary: the score is the
    maximum gain (1e-5 relative), the threshold is the lowest one attaining
    it, and a node whose every boundary has zero gain (up to rounding) gets
    threshold 0 and score 0 or a rounding-level gain."""
    rng = np.random.default_rng(11)
    pinned = 0
    for trial in range(1500):
        m = int(rng.integers(1, 17))
        C = int(rng.integers(2, 5))
        pool = rng.standard_normal(int(rng.integers(1, 6))).astype(np.float32)
        values = rng.choice(pool, m) if rng.random() < 0.5 else rng.standard_normal(m).astype(np.float32)
        labels = rng.integers(0, C, m).astype(np.int32)
        weights = rng.integers(1, 4, m).astype(np.uint32)
        score, thr = O.rf_best_split(values, labels, weights, C)
        gains = _boundary_gains(values, labels, weights, C)
        best = max([g for g, _ in gains], default=0.0)
        ctx = (values, labels, weights)
        if best <= TIE:
            assert abs(float(score)) <= TIE, ctx
            assert thr == 0.0 or any(t == thr for _, t in gains), ctx
            continue
        assert float(score) == pytest.approx(best, rel=1e-5), ctx
        near = [t for g, t in gains if g >= best - TIE]
        if len(near) == 1:
            assert thr == np.float32(near[0]), ctx
            pinned += 1
        else:
            assert thr in [np.float32(t) for t in near], ctx
    assert pinned > 500


def test_rf_best_split_enumerates_all_boundaries_tiny():
    """n <= 6 with all label assignments: the chosen threshold's partition
    attains the maximum over every subset that a threshold can cut."""
    vals = np.array([-1.5, -0.0, 0.0, 0.5, 0.5, 2.0], np.float32)
    for labs in itertools.product(range(2), repeat=vals.size):
        labels = np.array(labs, np.int32)
        w = np.ones(vals.size, np.uint32)
        score, thr = O.rf_best_split(vals, labels, w, 2)
        gains = _boundary_gains(np.where(vals == 0, np.float32(0), vals), labels, w, 2)
        best = max(g for g, _ in gains)
        if best <= TIE:
            assert abs(float(score)) <= TIE
        else:
            assert float(score) == pytest.approx(best, rel=1e-5)
            assert min(t for g, t in gains if g >= best - TIE) == thr
