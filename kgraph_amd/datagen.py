"""Seeded synthetic datasets for the benchmark workloads.

numpy.random.RandomState streams are frozen (NEP 19), so the same seed gives
the same bytes on any numpy version and on any machine.
"""
import hashlib

import numpy as np


def uniform(n, d, seed):
    """n x d float32 rows i.i.d. uniform[0, 1)."""
    return np.random.RandomState(seed).random_sample((n, d)).astype(np.float32)


def integer_mixture(n, d, centers, seed_centers, seed_rows, sigma=30.0, center_hi=120.0, chunk=1 << 17):
    """Integer-valued Gaussian mixture: centres mu ~ U[0, center_hi)^d; each row
    picks a centre c ~ U{0..centers-1} and draws x = clip(rint(|N(mu_c,
    sigma^2)|), 0, 255).  Values are integers <= 255, so every squared
    distance (<= d * 255^2 < 2^24 for d <= 256) is exact in FP32."""
    mu = np.random.RandomState(seed_centers).uniform(0.0, center_hi, size=(centers, d))
    rs = np.random.RandomState(seed_rows)
    lab = rs.randint(0, centers, size=n)
    out = np.empty((n, d), dtype=np.float32)
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        x = rs.normal(0.0, sigma, size=(e - s, d)) + mu[lab[s:e]]
        out[s:e] = np.clip(np.rint(np.abs(x)), 0, 255)
    return out


def angular_mixture(n, d, centers, seed_centers, seed_rows, sigma=0.3, chunk=1 << 17):
    """L2-normalised Gaussian mixture: centres g/||g||, g ~ N(0, I); rows
    x = mu_c + sigma * N(0, I), then x / ||x||."""
    g = np.random.RandomState(seed_centers).normal(size=(centers, d))
    mu = g / np.linalg.norm(g, axis=1, keepdims=True)
    rs = np.random.RandomState(seed_rows)
    lab = rs.randint(0, centers, size=n)
    out = np.empty((n, d), dtype=np.float32)
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        x = mu[lab[s:e]] + sigma * rs.normal(size=(e - s, d))
        out[s:e] = x / np.linalg.norm(x, axis=1, keepdims=True)
    return out


def anisotropic_mixture(n, d, centers, seed_centers, seed_sigma, seed_rows, chunk=1 << 17):
    """Gaussian mixture with per-cluster diagonal sigma ~ U[0.05, 0.5]^d."""
    mu = np.random.RandomState(seed_centers).normal(size=(centers, d))
    sig = np.random.RandomState(seed_sigma).uniform(0.05, 0.5, size=(centers, d))
    rs = np.random.RandomState(seed_rows)
    lab = rs.randint(0, centers, size=n)
    out = np.empty((n, d), dtype=np.float32)
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        out[s:e] = mu[lab[s:e]] + sig[lab[s:e]] * rs.normal(size=(e - s, d))
    return out


def sha256(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


# The benchmark workloads (sizes, seeds, parameters).
WORKLOADS = {
    "uniform-100k-d16": dict(data=lambda: uniform(100_000, 16, 1), queries=lambda: uniform(1_000, 16, 2),
                             build=dict(K=25, L=100, S=10, R=100, iterations=30, reverse=-1),
                             search=dict(K=10, P=[100])),
    "intmix-1m-d128": dict(data=lambda: integer_mixture(1_000_000, 128, 1000, 3, 4),
                           queries=lambda: integer_mixture(10_000, 128, 1000, 3, 5),
                           build=dict(K=100, L=150, S=10, R=100, iterations=15, reverse=-1),
                           search=dict(K=10, P=[50, 100, 200])),
}
