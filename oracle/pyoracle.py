"""ctypes front end of the CPU oracle (test infrastructure only).

Imported by tests/, by __graft_entry__.smoke() and by bench.py's
cpu_baseline leg, always as the checker.  The product never imports it.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None


class OracleParams(ctypes.Structure):
    _fields_ = [
        ("algorithm", ctypes.c_int32),
        ("task", ctypes.c_int32),
        ("n_trees", ctypes.c_int32),
        ("k", ctypes.c_int32),
        ("max_depth", ctypes.c_int32),
        ("min_node_size", ctypes.c_int32),
        ("n_classes", ctypes.c_int32),
        ("n_threads", ctypes.c_int32),
        ("seed", ctypes.c_uint64),
    ]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError("oracle library missing: run `make -C oracle`")
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.c_void_p
        L.oracle_fit.argtypes = [P, P, ctypes.c_int64, ctypes.c_int32, ctypes.POINTER(OracleParams),
                                 ctypes.POINTER(ctypes.c_void_p)]
        L.oracle_fit.restype = ctypes.c_int
        L.oracle_forest_free.argtypes = [P]
        L.oracle_forest_n_nodes.argtypes = [P]
        L.oracle_forest_n_nodes.restype = ctypes.c_int64
        L.oracle_forest_n_trees.argtypes = [P]
        L.oracle_forest_n_trees.restype = ctypes.c_int32
        L.oracle_forest_stat_width.argtypes = [P]
        L.oracle_forest_stat_width.restype = ctypes.c_int32
        L.oracle_forest_fixed_shift.argtypes = [P]
        L.oracle_forest_fixed_shift.restype = ctypes.c_int32
        L.oracle_forest_get.argtypes = [P] * 13
        L.oracle_forest_recompute_scores.argtypes = [P, P, P, ctypes.c_int64, ctypes.c_int32, P]
        L.oracle_predict.argtypes = [P, P, ctypes.c_int64, P, P]
        L.oracle_philox4x32_10.argtypes = [P, P, P]
        L.oracle_rand_u32.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                      ctypes.c_uint64]
        L.oracle_rand_u32.restype = ctypes.c_uint32
        L.oracle_xlnx.argtypes = [ctypes.c_uint64]
        L.oracle_xlnx.restype = ctypes.c_double
        L.oracle_info_gain.argtypes = [P, P, ctypes.c_int32, ctypes.c_int32]
        L.oracle_ig_shift.argtypes = [ctypes.c_uint64, ctypes.c_int32]
        L.oracle_ig_shift.restype = ctypes.c_int32
        L.oracle_info_gain.restype = ctypes.c_double
        L.oracle_variance_gain.argtypes = [ctypes.c_int64] * 4 + [ctypes.c_int32]
        L.oracle_variance_gain.restype = ctypes.c_double
        L.oracle_fixed_shift.argtypes = [ctypes.c_double, ctypes.c_int64]
        L.oracle_fixed_shift.restype = ctypes.c_int32
        L.oracle_midpoint.argtypes = [ctypes.c_float, ctypes.c_float]
        L.oracle_midpoint.restype = ctypes.c_float
        L.oracle_rf_best_split.argtypes = [P, P, P, ctypes.c_int64, ctypes.c_int32, P]
        L.oracle_rf_best_split.restype = ctypes.c_float
        L.oracle_last_error.restype = ctypes.c_char_p
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


class Forest:
    """Canonical node arrays (tree-major, level order inside each tree)."""

    FIELDS = ("tree_start", "feature", "threshold", "left", "depth", "count", "n_entries", "score",
              "rowhash", "class_counts", "sum_fixed", "value")

    def __init__(self, handle, task):
        self._h = handle
        L = lib()
        self.task = task
        self.n_nodes = L.oracle_forest_n_nodes(handle)
        self.n_trees = L.oracle_forest_n_trees(handle)
        self.stat_width = L.oracle_forest_stat_width(handle)
        self.fixed_shift = L.oracle_forest_fixed_shift(handle)
        n = self.n_nodes
        self.tree_start = np.zeros(self.n_trees + 1, np.int64)
        self.feature = np.zeros(n, np.int32)
        self.threshold = np.zeros(n, np.float32)
        self.left = np.zeros(n, np.int32)
        self.depth = np.zeros(n, np.int32)
        self.count = np.zeros(n, np.int32)
        self.n_entries = np.zeros(n, np.int32)
        self.score = np.zeros(n, np.float32)
        self.rowhash = np.zeros(n, np.uint64)
        self.class_counts = np.zeros(n * self.stat_width if task == 0 else 0, np.uint32)
        self.sum_fixed = np.zeros(n, np.int64)
        self.value = np.zeros(n, np.float32)
        L.oracle_forest_get(handle, *[_ptr(getattr(self, k)) if getattr(self, k).size else None
                                      for k in self.FIELDS])
        if task == 0:
            self.class_counts = self.class_counts.reshape(n, self.stat_width)

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.oracle_forest_free(self._h)
            self._h = None

    def predict(self, X):
        X = np.ascontiguousarray(X, np.float32)
        n_rows = X.shape[1]
        if self.task == 0:
            votes = np.zeros((n_rows, self.stat_width), np.int32)
            lib().oracle_predict(self._h, _ptr(X), n_rows, _ptr(votes), None)
            return votes
        sums = np.zeros(n_rows, np.float32)
        lib().oracle_predict(self._h, _ptr(X), n_rows, None, _ptr(sums))
        return sums

    def recompute_scores(self, X, y):
        out = np.zeros(self.n_nodes, np.float64)
        lib().oracle_forest_recompute_scores(self._h, _ptr(X), _ptr(y), X.shape[1], X.shape[0], _ptr(out))
        return out


def fit(X, y, *, algorithm, task=0, n_trees, k=0, max_depth=1000, min_node_size=5, n_classes=0,
        seed=0, n_threads=0):
    """X: (F, N) float32 feature-major; y: (N,) float32."""
    X = np.ascontiguousarray(X, np.float32)
    y = np.ascontiguousarray(y, np.float32)
    F, N = X.shape
    p = OracleParams(algorithm, task, n_trees, k, max_depth, min_node_size, n_classes, n_threads, seed)
    h = ctypes.c_void_p()
    rc = lib().oracle_fit(_ptr(X), _ptr(y), N, F, ctypes.byref(p), ctypes.byref(h))
    if rc != 0:
        raise RuntimeError(lib().oracle_last_error().decode())
    return Forest(h, task)


def philox(ctr, key):
    c = np.asarray(ctr, np.uint32)
    k = np.asarray(key, np.uint32)
    out = np.zeros(4, np.uint32)
    lib().oracle_philox4x32_10(_ptr(c), _ptr(k), _ptr(out))
    return out


def xlnx(x):
    return lib().oracle_xlnx(int(x))


def info_gain(node, left, F=0):
    node = np.asarray(node, np.uint32)
    left = np.asarray(left, np.uint32)
    return lib().oracle_info_gain(_ptr(node), _ptr(left), node.size, F)


def ig_shift(total_weight, n_classes):
    return lib().oracle_ig_shift(int(total_weight), n_classes)


def variance_gain(nl, sl, nr, sr, shift):
    return lib().oracle_variance_gain(nl, sl, nr, sr, shift)


def rf_best_split(values, labels, weights, n_classes):
    v = np.ascontiguousarray(values, np.float32)
    lab = np.ascontiguousarray(labels, np.int32)
    w = np.ascontiguousarray(weights, np.uint32)
    thr = np.zeros(1, np.float32)
    s = lib().oracle_rf_best_split(_ptr(v), _ptr(lab), _ptr(w), v.size, n_classes, _ptr(thr))
    return np.float32(s), thr[0]


STREAM_BOOT = 0x746F6F62  # bootstrap draws (same counter layout as the build)
STREAM_IMP = 0x00706D69   # per-visit row draws of the permutation pass
_FLT_MAX = np.float32(3.4028234663852886e38)


def rand_u32(seed, stream, tree, node, idx):
    """Word idx % 4 of Philox4x32-10 at counter (idx / 4, node, tree, stream), key = seed."""
    o = philox((idx >> 2, node, tree, stream), (seed & 0xFFFFFFFF, seed >> 32))
    return int(o[idx & 3])


def oob_evaluate(forest, X, y, *, seed, tree_offset=0, permuted_feature=-1):
    """Out-of-bag pairs of a random forest, restating lib_ensembles/source/
    gpurf.cu:290-438 (gpurf_oob_estimate / gpurf_feature_importance) with the
    build's counter-based draws: every (tree, row) pair whose bootstrap left
    the row out is routed through the tree; nodes splitting on
    permuted_feature read row rand_u32(seed, STREAM_IMP, tree, row, depth) % N
    (the reference draws curand % N per visit).  Returns (pairs, correct
    pairs) for classification, (pairs, sum of squared errors) for
    regression (the reference sums leaf values there).  Pure Python: for
    test-sized inputs."""
    X = np.asarray(X, np.float32)
    F, N = X.shape
    pairs, ok, sse = 0, 0, 0.0
    for t in range(forest.n_trees):
        tg = tree_offset + t
        inbag = np.zeros(N, bool)
        for i in range(N):
            inbag[rand_u32(seed, STREAM_BOOT, tg, 0, i) % N] = True
        for r in range(N):
            if inbag[r]:
                continue
            g = int(forest.tree_start[t])
            depth = 0
            while forest.feature[g] >= 0:
                f = int(forest.feature[g])
                src = r
                if f == permuted_feature:
                    src = rand_u32(seed, STREAM_IMP, tg, r, depth) % N
                x = X[f, src]
                right = 0 if (x == -_FLT_MAX or x < forest.threshold[g]) else 1
                g = int(forest.left[g]) + right
                depth += 1
            pairs += 1
            if forest.task == 0:
                ok += int(np.argmax(forest.class_counts[g]) == int(y[r]))
            else:
                e = float(forest.value[g]) - float(y[r])
                sse += e * e
    return pairs, (ok if forest.task == 0 else sse)
