"""Seeded synthetic trajectories of the shapes the knn_rms workloads use.

Frames are float32 nm, [n_frames, natoms, 3].  Base structure: a 3-D random
walk with step 0.38 nm and directions uniform on the sphere.  Each frame gets
a uniformly random rotation (Shoemake quaternions) and a translation
U(-5, 5)^3 nm, so every pair needs a real superposition.

  trajectory(...)          per-frame noise sigma ~ U(0.05, 0.3) nm
  metastable_trajectory()  frames drawn from `n_centres` centres (base walk
                           plus N(0, sigma_centre^2) per centre) with
                           N(0, 0.1^2) noise; a fraction of frames are
                           near-duplicates: a random earlier frame plus
                           N(0, (1e-4)^2) and a fresh rigid transform.
"""
from __future__ import annotations

import numpy as np


def random_walk(natoms: int, rng: np.random.Generator, step: float = 0.38) -> np.ndarray:
    v = rng.normal(size=(natoms, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    v[0] = 0.0
    return np.cumsum(step * v, axis=0)


def random_rotations(n: int, rng: np.random.Generator) -> np.ndarray:
    """n uniformly distributed proper rotations (Shoemake 1992)."""
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    a, b = np.sqrt(1 - u1), np.sqrt(u1)
    w = a * np.sin(2 * np.pi * u2)
    x = a * np.cos(2 * np.pi * u2)
    y = b * np.sin(2 * np.pi * u3)
    z = b * np.cos(2 * np.pi * u3)
    R = np.empty((n, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z)
    R[:, 0, 1] = 2 * (x * y - z * w)
    R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w)
    R[:, 1, 1] = 1 - 2 * (x * x + z * z)
    R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w)
    R[:, 2, 1] = 2 * (y * z + x * w)
    R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def _rigid(pre: np.ndarray, rng: np.random.Generator, box: float = 5.0) -> np.ndarray:
    n = pre.shape[0]
    R = random_rotations(n, rng)
    t = rng.uniform(-box, box, size=(n, 1, 3))
    return (np.einsum("nij,naj->nai", R, pre) + t).astype(np.float32)


def trajectory(n_frames: int, natoms: int, seed: int, base_seed: int | None = None,
               sigma_range=(0.05, 0.3), chunk: int = 4096) -> np.ndarray:
    """Frames = R (base + sigma_f N(0,1)) + t with per-frame sigma_f ~ U(sigma_range)."""
    base = random_walk(natoms, np.random.default_rng(seed if base_seed is None else base_seed))
    rng = np.random.default_rng([seed, 1])
    out = np.empty((n_frames, natoms, 3), np.float32)
    for f0 in range(0, n_frames, chunk):
        n = min(chunk, n_frames - f0)
        sig = rng.uniform(*sigma_range, size=(n, 1, 1))
        pre = base[None] + sig * rng.normal(size=(n, natoms, 3))
        out[f0:f0 + n] = _rigid(pre, rng)
    return out


def metastable_trajectory(n_frames: int, natoms: int, seed: int, n_centres: int = 50,
                          sigma_centre: float = 0.3, sigma_frame: float = 0.1,
                          dup_fraction: float = 0.05, dup_sigma: float = 1e-4,
                          chunk: int = 4096) -> np.ndarray:
    rng = np.random.default_rng([seed, 2])
    base = random_walk(natoms, np.random.default_rng(seed))
    centres = base[None] + sigma_centre * rng.normal(size=(n_centres, natoms, 3))
    label = rng.integers(0, n_centres, size=n_frames)
    pre = np.empty((n_frames, natoms, 3), np.float32)
    for f0 in range(0, n_frames, chunk):
        n = min(chunk, n_frames - f0)
        pre[f0:f0 + n] = centres[label[f0:f0 + n]] + sigma_frame * rng.normal(size=(n, natoms, 3))
    dup = np.nonzero(rng.random(n_frames) < dup_fraction)[0]
    dup = dup[dup > 0]
    src = (rng.random(len(dup)) * dup).astype(np.int64)  # a random earlier frame
    for f, s in zip(dup, src):
        pre[f] = pre[s].astype(np.float64) + dup_sigma * rng.normal(size=(natoms, 3))
    out = np.empty_like(pre)
    for f0 in range(0, n_frames, chunk):
        n = min(chunk, n_frames - f0)
        out[f0:f0 + n] = _rigid(pre[f0:f0 + n].astype(np.float64), rng)
    return out


def masses(natoms: int, kind: str = "uniform", seed: int = 0) -> np.ndarray:
    """Per-atom weights: 'uniform' (all 1, a single atom type) or 'backbone'
    (the N, CA, C pattern of a protein backbone with standard masses)."""
    if kind == "uniform":
        return np.ones(natoms, np.float32)
    if kind == "backbone":
        pat = np.array([14.0067, 12.011, 12.011], np.float32)
        return np.resize(pat, natoms).astype(np.float32)
    if kind == "random":
        return np.random.default_rng(seed).uniform(1.0, 32.0, natoms).astype(np.float32)
    if kind == "wide":
        # hydrogen to heavy-metal range
        return np.random.default_rng(seed).uniform(1.0, 200.0, natoms).astype(np.float32)
    raise ValueError(kind)


def extended_chain(n_frames: int, natoms: int, seed: int, sigma: float = 0.05) -> np.ndarray:
    """Nearly straight chains: atoms 0.38 nm apart along a line with
    N(0, sigma^2) transverse and axial noise per frame, then a random rigid
    motion.  Close to collinear, so two of the key matrix's eigenvalues
    nearly coincide (the hard case of the QCP Newton iteration)."""
    rng = np.random.default_rng([seed, 3])
    line = np.zeros((natoms, 3))
    line[:, 0] = 0.38 * np.arange(natoms)
    line -= line.mean(axis=0)
    pre = line[None] + sigma * rng.normal(size=(n_frames, natoms, 3))
    return _rigid(pre, rng)


def data_vectors(n: int, vdim: int, seed: int, kind: str = "clusters",
                 n_centres: int = 20) -> np.ndarray:
    """Seeded FP64 data vectors [n, vdim] of the kinds knn_data reads from
    .pts files.

      clusters   n_centres Gaussian centres N(0, 1) plus N(0, 0.3^2) noise,
                 with 5 % near-duplicates (an earlier vector plus N(0, 1e-6^2))
      dihedral   sin/cos pairs of angles around n_centres conformations
                 (the angles_to_sincos form), noise 0.3 rad
      offset     clusters shifted by 50 in every element (a large common
                 component the correlation metric removes and the Euclidean
                 one must not lose precision on)
    """
    rng = np.random.default_rng(seed)
    if kind == "dihedral":
        half = (vdim + 1) // 2
        centres = rng.uniform(-np.pi, np.pi, (n_centres, half))
        ang = centres[rng.integers(0, n_centres, n)] + rng.normal(0.0, 0.3, (n, half))
        out = np.empty((n, 2 * half))
        out[:, 0::2] = np.sin(ang)
        out[:, 1::2] = np.cos(ang)
        return np.ascontiguousarray(out[:, :vdim])
    centres = rng.normal(0.0, 1.0, (n_centres, vdim))
    x = centres[rng.integers(0, n_centres, n)] + rng.normal(0.0, 0.3, (n, vdim))
    dup = np.nonzero(rng.random(n) < 0.05)[0]
    dup = dup[dup > 0]
    if len(dup):
        src = (rng.random(len(dup)) * dup).astype(np.int64)
        x[dup] = x[src] + rng.normal(0.0, 1e-6, (len(dup), vdim))
    if kind == "offset":
        x += 50.0
    return x
