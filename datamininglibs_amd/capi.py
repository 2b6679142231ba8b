"""ctypes binding of libdte (include/dte.h).

Python is a test / benchmark driver here; the product is the C ABI and the
C++ API above it.  Loading fails loudly when the HIP library is missing:
there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdte.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "dte.h")

RANDOM_FOREST = 0
EXTRA_TREES = 1
CLASSIFICATION = 0
REGRESSION = 1
BUILD_BATCHED = 0
BUILD_PLAIN = 1
BUILD_BATCHED_HOST = 2

OK = 0
ERR_INVALID_ARGUMENT = 1
ERR_HIP = 2
ERR_OUT_OF_MEMORY = 3
ERR_NO_DEVICE = 4
ERR_ADMISSION = 5
ERR_INTERNAL = 6
ERR_NOT_FOUND = 7
ERR_IO = 8


class DteError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"libdte error {code}: {msg}")
        self.code = code


class Params(ctypes.Structure):
    _fields_ = [
        ("algorithm", ctypes.c_int32),
        ("task", ctypes.c_int32),
        ("n_trees", ctypes.c_int32),
        ("n_features_per_node", ctypes.c_int32),
        ("max_depth", ctypes.c_int32),
        ("min_node_size", ctypes.c_int32),
        ("n_classes", ctypes.c_int32),
        ("device_id", ctypes.c_int32),
        ("seed", ctypes.c_uint64),
        ("build_mode", ctypes.c_int32),
        ("tree_offset", ctypes.c_int32),
        ("cut_team", ctypes.c_int32),
        ("cut_block", ctypes.c_int32),
        ("bucket_target", ctypes.c_int32),
        ("cut_ert", ctypes.c_int32),
        ("reserved", ctypes.c_int32 * 2),
    ]


class Node(ctypes.Structure):
    _fields_ = [
        ("child_start", ctypes.c_int32),
        ("child_count", ctypes.c_int32),
        ("probability_start", ctypes.c_int32),
        ("attribute", ctypes.c_int32),
        ("split_point", ctypes.c_float),
    ]


NODE_DTYPE = np.dtype([("child_start", np.int32), ("child_count", np.int32),
                       ("probability_start", np.int32), ("attribute", np.int32),
                       ("split_point", np.float32)])


class ForestInfo(ctypes.Structure):
    _fields_ = [
        ("n_trees", ctypes.c_int32),
        ("n_classes", ctypes.c_int32),
        ("n_features", ctypes.c_int32),
        ("task", ctypes.c_int32),
        ("algorithm", ctypes.c_int32),
        ("fixed_shift", ctypes.c_int32),
        ("n_nodes", ctypes.c_int64),
        ("n_leaves", ctypes.c_int64),
        ("n_probabilities", ctypes.c_int64),
        ("device_bytes", ctypes.c_int64),
    ]


class ServerStats(ctypes.Structure):
    _fields_ = [
        ("requests_done", ctypes.c_int64),
        ("rows_done", ctypes.c_int64),
        ("batches_launched", ctypes.c_int64),
        ("requests_rejected", ctypes.c_int64),
        ("busy_ms", ctypes.c_double),
        ("hbm_in_use", ctypes.c_int64),
        ("stage_ms", ctypes.c_double),
        ("h2d_ms", ctypes.c_double),
        ("kernel_ms", ctypes.c_double),
        ("d2h_ms", ctypes.c_double),
        ("scatter_ms", ctypes.c_double),
        ("groups_launched", ctypes.c_int64),
    ]


_lib = None
P = ctypes.c_void_p


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libdte.so not built ({LIB_PATH}); run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        L.dte_params_default.argtypes = [ctypes.c_int32, ctypes.POINTER(Params)]
        L.dte_fit.argtypes = [P, P, ctypes.c_int64, ctypes.c_int32, ctypes.POINTER(Params), ctypes.POINTER(P)]
        L.dte_fit_device.argtypes = L.dte_fit.argtypes
        L.dte_forest_info_get.argtypes = [P, ctypes.POINTER(ForestInfo)]
        L.dte_forest_export.argtypes = [P, P, ctypes.c_int64, P, ctypes.c_int64]
        L.dte_forest_import.argtypes = [P, ctypes.c_int64, P, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32,
                                        ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(P)]
        L.dte_forest_canonical.argtypes = [P] * 13
        L.dte_forest_free.argtypes = [P]
        L.dte_predict.argtypes = [P, P, ctypes.c_int64, P]
        L.dte_predict_noalloc.argtypes = [P, P, ctypes.c_int64, P]
        L.dte_predict_device.argtypes = [P, P, ctypes.c_int64, ctypes.c_int64, ctypes.c_int32, P, P, P]
        L.dte_server_create.argtypes = [ctypes.c_int32, ctypes.c_int64, ctypes.c_int32, ctypes.POINTER(P)]
        L.dte_server_load_forest.argtypes = [P, P, ctypes.POINTER(ctypes.c_int32)]
        L.dte_server_submit.argtypes = [P, ctypes.c_int32, P, ctypes.c_int64, ctypes.POINTER(ctypes.c_uint64)]
        L.dte_server_poll.argtypes = [P, ctypes.c_uint64, P, ctypes.POINTER(ctypes.c_int32)]
        L.dte_server_wait.argtypes = [P, ctypes.c_uint64, P]
        L.dte_server_get_stats.argtypes = [P, ctypes.POINTER(ServerStats)]
        L.dte_server_destroy.argtypes = [P]
        L.dte_profiling.argtypes = [ctypes.c_int32]
        L.dte_profile_query.argtypes = [ctypes.c_int32, ctypes.c_char_p, ctypes.c_int32,
                                        ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_double),
                                        ctypes.POINTER(ctypes.c_double)]
        L.dte_device_count.argtypes = [ctypes.POINTER(ctypes.c_int32)]
        L.dte_device_alloc.argtypes = [ctypes.c_int32, ctypes.c_int64, ctypes.POINTER(P)]
        L.dte_device_free.argtypes = [ctypes.c_int32, P]
        L.dte_memcpy_to_device.argtypes = [ctypes.c_int32, P, P, ctypes.c_int64]
        L.dte_memcpy_to_host.argtypes = [ctypes.c_int32, P, P, ctypes.c_int64]
        L.dte_last_error.argtypes = [ctypes.c_char_p, ctypes.c_int64]
        _lib = L
    return _lib


def _check(rc):
    if rc != OK:
        buf = ctypes.create_string_buffer(1024)
        lib().dte_last_error(buf, 1024)
        raise DteError(rc, buf.value.decode(errors="replace"))


def _ptr(a):
    return a.ctypes.data_as(P) if a is not None else None


def default_params(algorithm=RANDOM_FOREST, **kw):
    p = Params()
    _check(lib().dte_params_default(algorithm, ctypes.byref(p)))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def device_count():
    n = ctypes.c_int32(0)
    _check(lib().dte_device_count(ctypes.byref(n)))
    return n.value


class Forest:
    def __init__(self, handle):
        self._h = handle
        info = ForestInfo()
        _check(lib().dte_forest_info_get(handle, ctypes.byref(info)))
        self.info = info
        self.n_trees = info.n_trees
        self.n_classes = info.n_classes
        self.n_features = info.n_features
        self.task = info.task
        self.n_nodes = info.n_nodes

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            lib().dte_forest_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def canonical(self):
        """Canonical arrays (same names as the oracle's Forest)."""
        n = self.n_nodes
        C = self.n_classes
        out = {
            "tree_start": np.zeros(self.n_trees + 1, np.int64),
            "feature": np.zeros(n, np.int32),
            "threshold": np.zeros(n, np.float32),
            "left": np.zeros(n, np.int32),
            "depth": np.zeros(n, np.int32),
            "count": np.zeros(n, np.int32),
            "n_entries": np.zeros(n, np.int32),
            "score": np.zeros(n, np.float32),
            "rowhash": np.zeros(n, np.uint64),
            "class_counts": np.zeros((n, C) if self.task == CLASSIFICATION else (0,), np.uint32),
            "sum_fixed": np.zeros(n, np.int64),
            "value": np.zeros(n, np.float32),
        }
        order = ["tree_start", "feature", "threshold", "left", "depth", "count", "n_entries", "score",
                 "rowhash", "class_counts", "sum_fixed", "value"]
        _check(lib().dte_forest_canonical(self._h, *[_ptr(out[k]) if out[k].size else None for k in order]))
        return out

    def save(self, path):
        _check(lib().dte_forest_save(self._h, os.fsencode(path)))

    def export(self):
        """Reference model layout: (nodes structured array, probability array)."""
        nodes = np.zeros(self.n_nodes, NODE_DTYPE)
        probs = np.zeros(self.info.n_probabilities, np.float32)
        _check(lib().dte_forest_export(self._h, _ptr(nodes), nodes.size, _ptr(probs), probs.size))
        return nodes, probs

    def predict(self, X):
        """X: (F, n) feature-major float32.  Votes (n, C) or sums (n,)."""
        X = np.ascontiguousarray(X, np.float32)
        if X.ndim != 2 or X.shape[0] != self.n_features:
            raise ValueError(f"X must be (n_features={self.n_features}, n_rows), got {X.shape}")
        n = X.shape[1]
        if self.task == CLASSIFICATION:
            out = np.zeros((n, self.n_classes), np.float32)
        else:
            out = np.zeros(n, np.float32)
        _check(lib().dte_predict(self._h, _ptr(X), n, _ptr(out)))
        return out

    def predict_noalloc(self, X):
        """predict() through dte_predict_noalloc: no device allocation, the
        library's fixed 64 MiB scratch holds the samples and votes."""
        X = np.ascontiguousarray(X, np.float32)
        if X.ndim != 2 or X.shape[0] != self.n_features:
            raise ValueError(f"X must be (n_features={self.n_features}, n_rows), got {X.shape}")
        n = X.shape[1]
        out = np.zeros((n, self.n_classes) if self.task == CLASSIFICATION else n, np.float32)
        _check(lib().dte_predict_noalloc(self._h, _ptr(X), n, _ptr(out)))
        return out


def fit(X, y, params):
    """X: (F, N) float32 feature-major host array; y: (N,) float32."""
    X = np.ascontiguousarray(X, np.float32)
    y = np.ascontiguousarray(y, np.float32)
    F, N = X.shape
    h = P()
    _check(lib().dte_fit(_ptr(X), _ptr(y), N, F, ctypes.byref(params), ctypes.byref(h)))
    return Forest(h)


def fit_device(X_dev_ptr, n_rows, n_features, y, params):
    y = np.ascontiguousarray(y, np.float32)
    h = P()
    _check(lib().dte_fit_device(P(X_dev_ptr), _ptr(y), n_rows, n_features, ctypes.byref(params),
                                ctypes.byref(h)))
    return Forest(h)


def autotune(X_dev_ptr, n_rows, n_features, y, params, trial_trees=0):
    """Split-search tuning for this data set (dte_autotune): (tuned params, seconds spent)."""
    y = np.ascontiguousarray(y, np.float32)
    tuned = Params()
    secs = ctypes.c_double()
    _check(lib().dte_autotune(P(X_dev_ptr), _ptr(y), n_rows, n_features, ctypes.byref(params), int(trial_trees),
                              ctypes.byref(tuned), ctypes.byref(secs)))
    return tuned, secs.value


def fit_many(X_dev_ptrs, n_rows, n_features, ys, params, max_concurrent=0):
    """Independent fits (dte_fit_many): extra-trees forests of one task are
    built jointly, one level step serving all of them; other fits run
    concurrently on separate streams."""
    n = len(X_dev_ptrs)
    ys = [np.ascontiguousarray(y, np.float32) for y in ys]
    xs = (P * n)(*[P(x) for x in X_dev_ptrs])
    ts = (P * n)(*[P(y.ctypes.data) for y in ys])
    rows = (ctypes.c_int64 * n)(*[int(r) for r in n_rows])
    feats = (ctypes.c_int32 * n)(*[int(f) for f in n_features])
    ps = (Params * n)(*params)
    out = (P * n)()
    _check(lib().dte_fit_many(n, xs, ts, rows, feats, ps, int(max_concurrent), out))
    return [Forest(P(out[i])) for i in range(n)]


def oob_evaluate(forest, X_dev_ptr, y, n_rows, permuted_feature=-1):
    """Out-of-bag pass (dte_oob_evaluate): (pairs, correct pairs) or (pairs, squared error sum)."""
    y = np.ascontiguousarray(y, np.float32)
    out = (ctypes.c_double * 2)()
    _check(lib().dte_oob_evaluate(forest.handle, P(X_dev_ptr), _ptr(y), int(n_rows), int(permuted_feature), out))
    return out[0], out[1]


def permutation_importance(forest, X_dev_ptr, y, n_rows):
    """Out-of-bag score and, per feature, its drop when the feature is permuted
    (accuracy drop for classification, MSE increase for regression)."""
    pairs, v = oob_evaluate(forest, X_dev_ptr, y, n_rows)
    base = v / pairs if pairs else 0.0
    imp = np.zeros(forest.n_features)
    for f in range(forest.n_features):
        pf, vf = oob_evaluate(forest, X_dev_ptr, y, n_rows, f)
        s = vf / pf if pf else 0.0
        imp[f] = base - s if forest.task == 0 else s - base
    return base, imp


def load_forest(path, device_id=0):
    h = P()
    _check(lib().dte_forest_load(os.fsencode(path), device_id, ctypes.byref(h)))
    return Forest(h)


def import_forest(nodes, probs, n_trees, n_classes, n_features, task, device_id=0):
    nodes = np.ascontiguousarray(nodes, NODE_DTYPE)
    probs = np.ascontiguousarray(probs, np.float32)
    h = P()
    _check(lib().dte_forest_import(_ptr(nodes), nodes.size, _ptr(probs), probs.size, n_trees, n_classes,
                                   n_features, task, device_id, ctypes.byref(h)))
    return Forest(h)


def profiling(on):
    _check(lib().dte_profiling(1 if on else 0))


def profile_reset():
    _check(lib().dte_profile_reset())


def release_cached_memory():
    """Frees the device scratch that fits keep cached; returns the bytes freed."""
    b = ctypes.c_int64()
    _check(lib().dte_release_cached_memory(ctypes.byref(b)))
    return b.value


def profile_stats():
    L = lib()
    n = L.dte_profile_query(-1, None, 0, None, None, None)
    out = {}
    for i in range(n):
        name = ctypes.create_string_buffer(64)
        la = ctypes.c_int64()
        ms = ctypes.c_double()
        by = ctypes.c_double()
        L.dte_profile_query(i, name, 64, ctypes.byref(la), ctypes.byref(ms), ctypes.byref(by))
        out[name.value.decode()] = {"launches": la.value, "total_ms": ms.value, "bytes": by.value}
    return out


def declared_symbols():
    """Function names declared in include/dte.h."""
    import re
    txt = open(HEADER_PATH).read()
    return sorted(set(re.findall(r"^int (dte_\w+)\(", txt, re.M)))


class DeviceArray:
    """A block of HBM owned through libdte (no torch needed)."""

    def __init__(self, nbytes, device=0):
        self.device = device
        self.nbytes = int(nbytes)
        p = P()
        _check(lib().dte_device_alloc(device, self.nbytes, ctypes.byref(p)))
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, a, device=0):
        a = np.ascontiguousarray(a)
        d = cls(a.nbytes, device)
        _check(lib().dte_memcpy_to_device(device, P(d.ptr), _ptr(a), a.nbytes))
        return d

    def to_numpy(self, dtype, shape):
        out = np.empty(shape, dtype)
        _check(lib().dte_memcpy_to_host(self.device, _ptr(out), P(self.ptr), out.nbytes))
        return out

    def free(self):
        if self.ptr:
            lib().dte_device_free(self.device, P(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Server:
    """Prediction server: resident forests, continuous batching (dte_server_*)."""

    def __init__(self, device=0, hbm_budget_bytes=0, max_batch_rows=65536):
        h = P()
        _check(lib().dte_server_create(device, hbm_budget_bytes, max_batch_rows, ctypes.byref(h)))
        self._h = h
        self._meta = {}

    def load(self, forest):
        slot = ctypes.c_int32()
        _check(lib().dte_server_load_forest(self._h, forest.handle, ctypes.byref(slot)))
        self._meta[slot.value] = (forest.n_features, forest.n_classes if forest.task == CLASSIFICATION else 1)
        return slot.value

    def submit(self, slot, rows):
        """rows: (n, F) row-major float32.  Returns a ticket."""
        rows = np.ascontiguousarray(rows, np.float32)
        # an unknown slot is the library's to report (DTE_ERR_NOT_FOUND)
        if slot in self._meta and (rows.ndim != 2 or rows.shape[1] != self._meta[slot][0]):
            raise ValueError(f"rows must be (n_rows, n_features={self._meta[slot][0]}), got {rows.shape}")
        t = ctypes.c_uint64()
        _check(lib().dte_server_submit(self._h, slot, _ptr(rows), rows.shape[0], ctypes.byref(t)))
        return t.value, rows.shape[0], slot

    def _out(self, req):
        _, n, slot = req
        width = self._meta[slot][1]
        return np.zeros((n, width) if width > 1 else n, np.float32)

    def wait(self, req):
        out = self._out(req)
        _check(lib().dte_server_wait(self._h, req[0], _ptr(out)))
        return out

    def poll(self, req):
        out = self._out(req)
        done = ctypes.c_int32()
        _check(lib().dte_server_poll(self._h, req[0], _ptr(out), ctypes.byref(done)))
        return out if done.value else None

    def stats(self):
        s = ServerStats()
        _check(lib().dte_server_get_stats(self._h, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in ServerStats._fields_}

    def close(self):
        if self._h:
            _check(lib().dte_server_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
