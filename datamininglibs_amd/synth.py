"""Synthetic workloads of the tree-ensemble build / predict benchmarks.

Every generator is deterministic from its seed (numpy PCG64) and returns
feature-major float32 samples X with shape (F, N) -- the layout of
MlDataFrame<float>::GetSamples() -- and float32 targets y of shape (N,).
"""
import numpy as np


def spam_like(seed=0, n=4601, f=57, n_informative=8, pos_rate=0.39):
    """4,601 x 57 nonnegative heavy-tailed features (40% zeros, the rest
    lognormal(0, 1.5)); binary labels from a sparse logistic function of 8
    features with the intercept solved for a 39% positive rate."""
    rng = np.random.default_rng(seed)
    X = rng.lognormal(0.0, 1.5, size=(f, n)).astype(np.float32)
    X[rng.random((f, n)) < 0.4] = 0.0
    S = rng.choice(f, n_informative, replace=False)
    w = rng.standard_normal(n_informative)
    z = (w[:, None] * np.log1p(X[S].astype(np.float64))).sum(0)
    u = rng.random(n)
    lo, hi = -50.0, 50.0
    for _ in range(200):
        b = 0.5 * (lo + hi)
        rate = np.mean(u < 1.0 / (1.0 + np.exp(-(z + b))))
        if rate < pos_rate:
            lo = b
        else:
            hi = b
    y = (u < 1.0 / (1.0 + np.exp(-(z + 0.5 * (lo + hi))))).astype(np.float32)
    return X, y


def gaussian_classes(seed=0, n=1_000_000, f=128, n_informative=16, n_classes=10):
    """N(0,1) features; label = argmax_c (W x_S + g)_c with W ~ N(0,1) on 16
    informative features and i.i.d. Gumbel noise g."""
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((f, n), dtype=np.float32)
    S = rng.choice(f, min(n_informative, f), replace=False)
    W = rng.standard_normal((n_classes, S.size)).astype(np.float32)
    logits = W @ X[S]
    logits += rng.gumbel(size=logits.shape).astype(np.float32)
    y = np.argmax(logits, axis=0).astype(np.float32)
    return X, y


def sine_regression(seed=0, n=4_000_000, f=64, noise=0.1):
    """U[0,1) features; y = sin(4 x0) + x1 x2 + N(0, 0.1) in fp32."""
    rng = np.random.default_rng(seed)
    X = rng.random((f, n), dtype=np.float32)
    y = (np.sin(4.0 * X[0]) + X[1] * X[2] + noise * rng.standard_normal(n)).astype(np.float32)
    return X, y


def uneven_forests(seed=0, n_forests=64, n_min=2000, n_max=200_000):
    """Specs of 64 independent forests: rows log-uniform in [2k, 200k],
    features from {16,32,64,128,256}, 2..8 classes; data generated lazily."""
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(n_forests):
        n = int(round(np.exp(rng.uniform(np.log(n_min), np.log(n_max)))))
        f = int(rng.choice([16, 32, 64, 128, 256]))
        c = int(rng.integers(2, 9))
        specs.append({"index": i, "n": n, "f": f, "classes": c, "seed": int(rng.integers(1 << 31))})
    return specs


def uneven_forest_data(spec):
    return gaussian_classes(seed=spec["seed"], n=spec["n"], f=spec["f"], n_informative=spec["f"],
                            n_classes=spec["classes"])


def request_sizes(seed=0, n_requests=1000, max_rows=65536):
    """Rows per request, log-uniform in [1, 65536]."""
    rng = np.random.default_rng(seed)
    return np.maximum(1, np.round(np.exp(rng.uniform(0.0, np.log(max_rows), n_requests)))).astype(np.int64)
