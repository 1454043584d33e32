"""Seeded synthetic systems for the benchmark workloads.

Positions: a jittered simple-cubic lattice filling the periodic box (random
non-overlapping placement of anisotropic bodies at these packing fractions
needs compression; the lattice gives the stated density and pair counts).
Cubes at packing fraction <= 0.30 are additionally made non-overlapping with
an exact separating-axis test and orientation re-draws.
Quaternions: uniform on S^3 (Shoemake; 01/GeometryUtils.py:4-10), scalar first.
Velocities: N(0, kT/m).  Angular momentum: body L_k ~ N(0, kT I_k), stored as
p = 2 q (x) (0, L) so that s = q* (x) p / 2 = L (IntegrateStepOneKernels.py:244-248).
"""
import math

import numpy as np

from .shapes import Shape, get_shape
from .system import SystemSpec


def box_length(n, shape: Shape, phi):
    return (n * shape.volume / phi) ** (1.0 / 3.0)


def random_quaternions(rng, n):
    u = rng.random((n, 3))
    a = np.sqrt(1.0 - u[:, 0])
    b = np.sqrt(u[:, 0])
    return np.stack([a * np.sin(2 * np.pi * u[:, 1]), a * np.cos(2 * np.pi * u[:, 1]),
                     b * np.sin(2 * np.pi * u[:, 2]), b * np.cos(2 * np.pi * u[:, 2])], axis=1)


def _qmul(a, b):
    s = a[:, 0] * b[:, 0] - np.sum(a[:, 1:] * b[:, 1:], axis=1)
    v = a[:, :1] * b[:, 1:] + b[:, :1] * a[:, 1:] + np.cross(a[:, 1:], b[:, 1:])
    return np.concatenate([s[:, None], v], axis=1)


def _rotmat(q):
    w, x, y, z = q
    a = w * w - (x * x + y * y + z * z)
    return np.array([[a + 2 * x * x, 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), a + 2 * y * y, 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), a + 2 * z * z]])


def cubes_overlap(t, Ra, Rb, half=0.5):
    """Separating-axis test for two cubes of half-edge `half`, centres 0 and t."""
    A, B = Ra, Rb
    axes = [A[:, k] for k in range(3)] + [B[:, k] for k in range(3)]
    for i in range(3):
        for j in range(3):
            c = np.cross(A[:, i], B[:, j])
            n = np.linalg.norm(c)
            if n > 1e-12:
                axes.append(c / n)
    for ax in axes:
        ra = half * np.sum(np.abs(A.T @ ax))
        rb = half * np.sum(np.abs(B.T @ ax))
        if abs(np.dot(t, ax)) > ra + rb:
            return False
    return True


def lattice_positions(rng, n, L, jitter=0.15):
    m = int(math.ceil(n ** (1.0 / 3.0)))
    a = L / m
    idx = np.array([(i, j, k) for i in range(m) for j in range(m) for k in range(m)])[:n]
    if len(idx) < n:
        raise ValueError("lattice too small")
    sel = rng.permutation(m ** 3)[:n] if m ** 3 > n else np.arange(n)
    idx = np.array([(i, j, k) for i in range(m) for j in range(m) for k in range(m)])[sel]
    x = (idx + 0.5) * a - 0.5 * L + rng.uniform(-jitter, jitter, (n, 3)) * a
    x = np.where(x > 0.5 * L, x - L, x)
    x = np.where(x < -0.5 * L, x + L, x)
    return x


def _resolve_cube_overlaps(rng, x, q, L, attempts=2000):
    """Greedy sequential orientations: redraw q_i until cube i overlaps no earlier cube."""
    n = len(x)
    R = [None] * n
    for i in range(n):
        t = x[:i] - x[i]
        t = np.where(t > 0.5 * L, t - L, t)
        t = np.where(t < -0.5 * L, t + L, t)
        d = np.linalg.norm(t, axis=1)
        near = np.nonzero(d < math.sqrt(3.0))[0]
        if np.any(d[near] < 1.0):
            raise RuntimeError("cube centres closer than one edge")
        for _ in range(attempts):
            Ri = _rotmat(q[i])
            if not any(cubes_overlap(t[j], Ri, R[j]) for j in near):
                R[i] = Ri
                break
            q[i] = random_quaternions(rng, 1)[0]
        else:
            raise RuntimeError("could not place non-overlapping cubes")
    return q


def make_system(shape_name, n, phi, kT=1.0, dt=0.001, seed=0, rc=None, skin=0.5, non_overlap=None,
                thermo=None):
    shape = get_shape(shape_name)
    rng = np.random.default_rng(seed)
    L = box_length(n, shape, phi)
    if non_overlap is None:
        non_overlap = shape_name == "cube8" and n <= 4096
    x = lattice_positions(rng, n, L, jitter=0.02 if non_overlap else 0.15)
    q = random_quaternions(rng, n)
    if non_overlap and shape_name == "cube8":
        q = _resolve_cube_overlaps(rng, x, q, L)
    mass = np.ones(n)
    v = rng.standard_normal((n, 3)) * np.sqrt(kT / mass)[:, None]
    moi = np.asarray(shape.moi)
    Lb = rng.standard_normal((n, 3)) * np.sqrt(kT * moi)[None, :]
    p = 2.0 * _qmul(q, np.concatenate([np.zeros((n, 1)), Lb], axis=1))
    return SystemSpec(x, v, q, p, mass, L, kT, dt, shape.moi, shape.cutoff if rc is None else rc, skin=skin,
                      thermo=thermo)


def replica_specs(n_systems, seed, shape_name="cube8", n_range=(64, 4096), phi_range=(0.2, 0.5),
                  kT_range=(0.5, 2.0)):
    """Independent systems for batched runs: N log-uniform, packing and kT uniform (seeded).
    Positions are a jittered lattice with random orientations (bodies may overlap); the packing
    fraction is capped so that every box is at least 2.05 cutoffs wide (minimum image)."""
    shape = get_shape(shape_name)
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_systems):
        n = int(round(math.exp(rng.uniform(math.log(n_range[0]), math.log(n_range[1])))))
        phi = min(float(rng.uniform(*phi_range)), n * shape.volume / (2.05 * shape.cutoff) ** 3)
        kT = float(rng.uniform(*kT_range))
        out.append(make_system(shape_name, n, phi, kT=kT, seed=int(rng.integers(1 << 30)), non_overlap=False))
    return out


# Named workloads (sizes as in the benchmark definitions).
WORKLOADS = {
    "cube512": dict(shape="cube8", n=512, phi=0.30, kT=1.0, dt=0.001),
    "tet32k": dict(shape="tet10", n=32768, phi=0.45, kT=1.0, dt=0.001),
    "cyl262k": dict(shape="cyl12", n=262144, phi=0.35, kT=1.0, dt=0.001),
}
