"""Rigid-body shapes: body-frame interaction sites and per-shape constants.

The reference registry is nep_simulator/DescriptorGenerator.py:58-125
(cube_v2, th_v1, pbp_v3: sites, en_min/en_max, COM cutoff, nep cutoff).
The benchmark shapes cube8 / tet10 / cyl12 add unit mass and the solid-body
inertia of the shape; their COM cutoff equals the descriptor cutoff.
"""
import math
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Shape:
    name: str
    sites: np.ndarray            # (Ns, 3) body frame
    en_min: float
    en_max: float
    cutoff: float                # COM cutoff of the pair filter
    nep_cutoff: float            # descriptor cutoff r_nep
    volume: float = 1.0          # particle volume (packing fraction)
    moi: tuple = (1.0, 1.0, 1.0)  # diagonal inertia for unit mass
    hypers: tuple = (11, 6, 4)   # (nrad, nang, lmax) used with this shape by default
    meta: dict = field(default_factory=dict)

    @property
    def n_sites(self):
        return len(self.sites)


def _tet10(edge=1.0):
    a = edge / (2.0 * math.sqrt(2.0))
    v = np.array([[a, a, a], [a, -a, -a], [-a, a, -a], [-a, -a, a]])
    mids = [0.5 * (v[i] + v[j]) for i in range(4) for j in range(i + 1, 4)]
    return np.vstack([v, np.array(mids)])


def _cyl12(length=3.0, radius=0.5, n=12, twist_deg=90.0):
    # helix of n sites at the surface radius, azimuth spread over the twist,
    # evenly spaced along the axis; after the template of
    # 05/PtsPlacement.py:35-62 (pts_along_axis), centroid removed
    ang = np.linspace(0.0, np.deg2rad(twist_deg), n)
    s = np.zeros((n, 3))
    s[:, 0] = radius * np.cos(ang)
    s[:, 1] = radius * np.sin(ang)
    s[:, 2] = np.linspace(-0.5 * length, 0.5 * length, n)
    return s - s.mean(axis=0)


def _registry():
    cube8 = 0.5 * np.array([[sx, sy, sz] for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)], dtype=float)
    cyl_r, cyl_l = 0.5, 3.0
    return {
        "cube_v2": Shape("cube_v2", 1.7 * np.array([[1.0, 0, 0], [-1.0, 0, 0], [0, 1.0, 0], [0, -1.0, 0],
                                                   [0, 0, 1.0], [0, 0, -1.0]]),
                         -5.1, 15.0, 5.80, 4.5, hypers=(10, 8, 4)),
        "th_v1": Shape("th_v1", np.array([[-0.45, -0.45, -0.45], [-0.45, 0.45, 0.45], [0.45, -0.45, 0.45],
                                          [0.45, 0.45, -0.45]]),
                       -6.7, 15.0, 5.75, 4.5, hypers=(10, 8, 4)),
        "pbp_v3": Shape("pbp_v3", np.array([[-1.46946, 2.02254, 0.0], [-2.37764, -0.77254, 0.0],
                                            [-0.0, -2.5, 0.0], [2.37764, -0.77254, 0.0],
                                            [1.46946, 2.02254, 0.0], [0.0, 0.0, 3.5], [0.0, 0.0, -3.5]]),
                        -2.5, 15.0, 8.1, 7.5, hypers=(10, 8, 4)),
        "cube8": Shape("cube8", cube8, 0.0, 1.0, 2.5, 2.5, volume=1.0, moi=(1 / 6, 1 / 6, 1 / 6)),
        "tet10": Shape("tet10", _tet10(), 0.0, 1.0, 2.5, 2.5, volume=1.0 / (6.0 * math.sqrt(2.0)),
                       moi=(0.05, 0.05, 0.05)),
        "cyl12": Shape("cyl12", _cyl12(cyl_l, cyl_r), 0.0, 1.0, 4.0, 4.0, volume=math.pi * cyl_r ** 2 * cyl_l,
                       moi=((3 * cyl_r ** 2 + cyl_l ** 2) / 12.0, (3 * cyl_r ** 2 + cyl_l ** 2) / 12.0,
                            cyl_r ** 2 / 2.0),
                       hypers=(17, 6, 6)),
    }


SHAPES = _registry()


def get_shape(name):
    try:
        return SHAPES[name]
    except KeyError:
        raise ValueError(f"shape not found: {name!r} (known: {sorted(SHAPES)})") from None
