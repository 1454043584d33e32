"""Shared input builders for the pair-kernel tests (host side, numpy)."""
import numpy as np

from oracle import nep_oracle as O


def random_unit_quats(rng, n):
    q = rng.standard_normal((n, 4))
    return q / np.linalg.norm(q, axis=1, keepdims=True)


def random_pairs(rng, n_bodies, n_pairs, rmin=0.9, rmax=2.4):
    """Orientations of n_bodies, pairs (i != j) and random separations t."""
    q = random_unit_quats(rng, n_bodies).astype(np.float32)
    i = rng.integers(0, n_bodies, n_pairs)
    j = (i + 1 + rng.integers(0, n_bodies - 1, n_pairs)) % n_bodies
    d = rng.standard_normal((n_pairs, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    t = (d * rng.uniform(rmin, rmax, (n_pairs, 1))).astype(np.float32)
    return q, i.astype(np.int32), j.astype(np.int32), t


def oracle_clouds(q, i, j, t, sites):
    q64 = q.astype(np.float64)
    t64 = t.astype(np.float64)
    b = np.asarray(sites, dtype=np.float64)
    pi = O.quat_rotate(q64[i][:, None, :], b[None])
    pj = O.quat_rotate(q64[j][:, None, :], b[None]) + t64[:, None, :]
    return np.concatenate([pi, pj], axis=1)


def rowwise_rel(a, b):
    """max_k |a-b| / max_k |b| per row."""
    return np.max(np.abs(a - b), axis=1) / np.maximum(np.max(np.abs(b), axis=1), 1e-300)


def normwise_rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))
