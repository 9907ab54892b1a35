"""Seeded synthetic point sets for the CORE-SG workloads.

Gaussian mixtures with per-component spread plus uniform noise over the
bounding box of the clustered points, row-major float32. The same generator
feeds the benchmark and the parity tests, so a workload is fully named by its
parameters and seed.
"""
from __future__ import annotations

import numpy as np


def gaussian_mixture(n: int, d: int, components: int, std_lo: float, std_hi: float,
                     noise: float, seed: int, centre_box: float = 10.0) -> np.ndarray:
    """n points in d dims: `components` isotropic Gaussians with centres uniform
    in [-centre_box, centre_box]^d and std uniform in [std_lo, std_hi], plus a
    fraction `noise` of points uniform over the clustered points' bounding box.
    Rows are shuffled so noise is not clustered at the end."""
    rng = np.random.default_rng(seed)
    centres = rng.uniform(-centre_box, centre_box, size=(components, d))
    stds = rng.uniform(std_lo, std_hi, size=components)
    n_noise = int(round(noise * n))
    n_clu = n - n_noise
    assign = rng.integers(0, components, size=n_clu)
    pts = np.empty((n, d), dtype=np.float32)
    pts[:n_clu] = (centres[assign] + stds[assign, None] * rng.standard_normal((n_clu, d))).astype(
        np.float32)
    if n_noise:
        lo = pts[:n_clu].min(axis=0) if n_clu else np.full(d, -centre_box)
        hi = pts[:n_clu].max(axis=0) if n_clu else np.full(d, centre_box)
        pts[n_clu:] = rng.uniform(lo, hi, size=(n_noise, d)).astype(np.float32)
    return pts[rng.permutation(n)].copy()


def anisotropic_mixture(n: int, d: int, components: int, eig_lo: float, eig_hi: float,
                        seed: int, centre_box: float = 10.0) -> np.ndarray:
    """Mixture with anisotropic covariances: eigenvalues log-uniform in
    [eig_lo, eig_hi] under a random rotation per component."""
    rng = np.random.default_rng(seed)
    centres = rng.uniform(-centre_box, centre_box, size=(components, d))
    assign = rng.integers(0, components, size=n)
    pts = np.empty((n, d), dtype=np.float32)
    for c in range(components):
        idx = np.nonzero(assign == c)[0]
        if idx.size == 0:
            continue
        q, _ = np.linalg.qr(rng.standard_normal((d, d)))
        ev = np.exp(rng.uniform(np.log(eig_lo), np.log(eig_hi), size=d))
        z = rng.standard_normal((idx.size, d)) * np.sqrt(ev)
        pts[idx] = (centres[c] + z @ q.T).astype(np.float32)
    return pts


# Named workloads (sizes and seeds of the benchmark configurations).
WORKLOADS = {
    # 20k points, d=16, 10 components, std U[0.3,2.0], 5% noise, k_max=16
    "mixture-20k-d16": dict(n=20_000, d=16, components=10, std_lo=0.3, std_hi=2.0,
                            noise=0.05, seed=0, k_max=16),
    # 1M points, d=24, 32 components, std U[0.2,3.0], 2% noise, k_max=32
    "mixture-1m-d24": dict(n=1_000_000, d=24, components=32, std_lo=0.2, std_hi=3.0,
                           noise=0.02, seed=1, k_max=32),
}


def workload(name: str, n: int | None = None) -> tuple[np.ndarray, int]:
    """Points and k_max of a named workload (optionally truncated to n rows by
    generating with the same parameters at a smaller n)."""
    w = dict(WORKLOADS[name])
    k_max = w.pop("k_max")
    if n is not None:
        w["n"] = n
    return gaussian_mixture(**w), k_max


def pad_rows(x: np.ndarray) -> np.ndarray:
    """Zero-pad rows to a multiple of 8 floats (the stored row layout)."""
    n, d = x.shape
    stride = (d + 7) & ~7
    out = np.zeros((n, stride), dtype=np.float32)
    out[:, :d] = x
    return out
