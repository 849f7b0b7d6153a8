"""Seeded synthetic inputs of the five benchmark workloads (no datasets are
available offline).  Base seed 20261202 + workload index.

  mixture10_50d      N=10,000 x 50-D, 10 isotropic unit-variance Gaussians,
                     centres ~ N(0, 16 I), equal weights
  sparse784          N=60,000 x 784-D, 10 clusters, each with a support of
                     ~19% of the dimensions, values U[0, 255] on the support
  powerlaw40_50d     N=1,300,000 x 50-D, 40 Gaussians with sizes ~ c^-1.5,
                     random rotations, covariance eigenvalues 1/i, centres ~ N(0, 16 I)
  batch_jobs         256 jobs, N log-uniform in [500, 50,000], 50-D, 5..20-cluster mixtures
  serving_stream     Poisson arrivals, N log-uniform in [1,000, 200,000], d in {2, 3}
"""
import numpy as np

BASE_SEED = 20261202


def mixture_isotropic(n=10_000, dim=50, clusters=10, seed=BASE_SEED + 0):
    rng = np.random.default_rng(seed)
    centres = rng.normal(0.0, 4.0, size=(clusters, dim))
    labels = rng.integers(0, clusters, size=n)
    X = centres[labels] + rng.standard_normal((n, dim))
    return X.astype(np.float32), labels


def sparse_clusters(n=60_000, dim=784, clusters=10, density=0.19, seed=BASE_SEED + 1):
    rng = np.random.default_rng(seed)
    support = [rng.choice(dim, size=int(round(density * dim)), replace=False) for _ in range(clusters)]
    labels = rng.integers(0, clusters, size=n)
    X = np.zeros((n, dim), np.float32)
    for c in range(clusters):
        idx = np.nonzero(labels == c)[0]
        X[np.ix_(idx, support[c])] = rng.uniform(0, 255, size=(idx.size, support[c].size)).astype(np.float32)
    return X, labels


def powerlaw_mixture(n=1_300_000, dim=50, clusters=40, exponent=1.5, seed=BASE_SEED + 2):
    rng = np.random.default_rng(seed)
    w = np.arange(1, clusters + 1, dtype=np.float64) ** (-exponent)
    w /= w.sum()
    sizes = np.floor(w * n).astype(np.int64)
    sizes[0] += n - sizes.sum()
    sd = 1.0 / np.sqrt(np.arange(1, dim + 1, dtype=np.float64))
    X = np.empty((n, dim), np.float32)
    labels = np.empty(n, np.int32)
    o = 0
    for c in range(clusters):
        Q, _ = np.linalg.qr(rng.standard_normal((dim, dim)))
        centre = rng.normal(0.0, 4.0, size=dim)
        z = rng.standard_normal((sizes[c], dim)) * sd
        X[o:o + sizes[c]] = (z @ Q.T + centre).astype(np.float32)
        labels[o:o + sizes[c]] = c
        o += sizes[c]
    perm = rng.permutation(n)
    return X[perm], labels[perm]


def gaussian_job(n, dim=50, seed=0):
    rng = np.random.default_rng(seed)
    clusters = int(rng.integers(5, 21))
    centres = rng.normal(0.0, 4.0, size=(clusters, dim))
    labels = rng.integers(0, clusters, size=n)
    X = centres[labels] + rng.standard_normal((n, dim))
    return X.astype(np.float32), labels


def batch_sizes(jobs=256, lo=500, hi=50_000, seed=BASE_SEED + 3):
    rng = np.random.default_rng(seed)
    return np.exp(rng.uniform(np.log(lo), np.log(hi), size=jobs)).astype(np.int64)


def serving_requests(count, rate_per_s, lo=1_000, hi=200_000, seed=BASE_SEED + 4):
    """List of (arrival_time_s, n, d, job_seed)."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate_per_s, size=count))
    ns = np.exp(rng.uniform(np.log(lo), np.log(hi), size=count)).astype(np.int64)
    ds = rng.choice([2, 3], size=count)
    return [(float(t[i]), int(ns[i]), int(ds[i]), int(seed * 1000 + i)) for i in range(count)]


def random_init(n, d, seed=0):
    """N(0, 1e-4^2) start, the reference's initialisation (CU/utils.cu:185-192)."""
    rng = np.random.default_rng(seed)
    return (1e-4 * rng.standard_normal((n, d))).astype(np.float32)
