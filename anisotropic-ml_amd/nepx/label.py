"""Label potential of the reference's data generator: soft bead-bead repulsion between rigid bodies
(01/SamplerUtils.py:555-631).

  E_ij = sum_{a in i, b in j} eps ((r0 - d_ab) / sigma)^2  for d_ab < r0,  eps 5, sigma 1.1, r0 2^(1/6)

BeadPotential runs it as a pair model of the MD step (nepx_bead_pair_forces: analytic forces and
torques in the per-pair layout of the NEP kernels, so the gather and the integrator are shared), e.g.
to validate learned forces against the potential the labels came from.  calc_en_unred labels rows
of the unreduced pair table (ZXZ Euler angles of body 2, then its spherical position) on the GPU.
"""
import numpy as np
import torch

from . import _lib

EPS, SIGMA, R0 = 5.0, 1.1, 2.0 ** (1.0 / 6.0)


def cube_structure_beads():
    """The 152-bead cube of 01/cube_structure.gsd: surface points of a 6x6x6 grid of spacing 0.7
    centred at the origin, in the file's order (z fastest, then x, then y)."""
    g = -1.75 + 0.7 * np.arange(6)
    return np.array([(x, y, z) for y in g for x in g for z in g if max(abs(x), abs(y), abs(z)) > 1.75 - 1e-9])


def matrix_to_quat(R):
    """Unit quaternions (w, x, y, z) of rotation matrices (P, 3, 3) (Shepperd's method: the largest of
    4w^2, 4x^2, 4y^2, 4z^2 is read from the diagonal, the other components from off-diagonal sums)."""
    R = np.asarray(R, dtype=np.float64)
    m00, m11, m22 = R[:, 0, 0], R[:, 1, 1], R[:, 2, 2]
    tr = m00 + m11 + m22
    d10, d01 = R[:, 1, 0], R[:, 0, 1]
    d20, d02 = R[:, 2, 0], R[:, 0, 2]
    d21, d12 = R[:, 2, 1], R[:, 1, 2]
    k = np.argmax(np.stack([tr, m00, m11, m22], axis=1), axis=1)
    q = np.empty((len(R), 4))
    with np.errstate(divide="ignore", invalid="ignore"):
        s0 = 2.0 * np.sqrt(np.maximum(1.0 + tr, 0.0))
        s1 = 2.0 * np.sqrt(np.maximum(1.0 + m00 - m11 - m22, 0.0))
        s2 = 2.0 * np.sqrt(np.maximum(1.0 + m11 - m00 - m22, 0.0))
        s3 = 2.0 * np.sqrt(np.maximum(1.0 + m22 - m00 - m11, 0.0))
        cand = [np.stack([0.25 * s0, (d21 - d12) / s0, (d02 - d20) / s0, (d10 - d01) / s0], 1),
                np.stack([(d21 - d12) / s1, 0.25 * s1, (d01 + d10) / s1, (d02 + d20) / s1], 1),
                np.stack([(d02 - d20) / s2, (d01 + d10) / s2, 0.25 * s2, (d12 + d21) / s2], 1),
                np.stack([(d10 - d01) / s3, (d02 + d20) / s3, (d12 + d21) / s3, 0.25 * s3], 1)]
    for c in range(4):
        q[k == c] = cand[c][k == c]
    return q / np.linalg.norm(q, axis=1, keepdims=True)


def euler_zxz(angles):
    """R = Rz(a0) Rx(a1) Rz(a2) for rows of angles (P, 3): the rotation of the unreduced table
    (01/SamplerUtils.py:589-605)."""
    a = np.asarray(angles, dtype=np.float64)
    s1, c1, s2, c2, s3, c3 = np.sin(a[:, 0]), np.cos(a[:, 0]), np.sin(a[:, 1]), np.cos(a[:, 1]), \
        np.sin(a[:, 2]), np.cos(a[:, 2])
    return np.stack([np.stack([c1 * c3 - c2 * s1 * s3, -c1 * s3 - c2 * c3 * s1, s1 * s2], 1),
                     np.stack([c3 * s1 + c1 * c2 * s3, -s1 * s3 + c1 * c2 * c3, -c1 * s2], 1),
                     np.stack([s2 * s3, c3 * s2, c2], 1)], axis=1)


class BeadPotential:
    """Pair model: body-frame beads (n, 3) with the soft repulsion above."""

    def __init__(self, beads=None, eps=EPS, sigma=SIGMA, r0=R0, device="cuda"):
        b = cube_structure_beads() if beads is None else np.asarray(beads, dtype=np.float64)
        if not 0 < len(b) <= 256:
            raise ValueError("1..256 beads per body")
        self.beads = b
        self.eps, self.sigma, self.r0 = float(eps), float(sigma), float(r0)
        self.device = torch.device(device)
        self.b_dev = torch.from_numpy(b.astype(np.float32).reshape(-1)).to(self.device)

    @property
    def reach(self):
        """Largest centre distance at which two bodies interact."""
        return 2.0 * float(np.linalg.norm(self.beads, axis=1).max()) + self.r0

    def pair_forces(self, qf_ptr, pi_ptr, pj_ptr, pt_ptr, n_dev_ptr, max_pairs, out_ptr, clip=0.0, stream=None):
        _lib.check(_lib.lib.nepx_bead_pair_forces(_lib.ptr(self.b_dev), len(self.beads), self.eps, self.sigma,
                                                  self.r0, qf_ptr, pi_ptr, pj_ptr, pt_ptr, n_dev_ptr,
                                                  int(max_pairs), float(clip), out_ptr,
                                                  _lib.stream_handle(stream)), "nepx_bead_pair_forces")


def unred_pairs(unred):
    """(q_j, t) of each unreduced row: body 1 at the origin unrotated, body 2 rotated by the ZXZ
    angles unred[:, 0:3] and placed at spherical (r, azimuth, polar) = unred[:, 3:6]."""
    u = np.asarray(unred, dtype=np.float64)
    st = np.sin(u[:, 5])
    t = np.stack([u[:, 3] * st * np.cos(u[:, 4]), u[:, 3] * st * np.sin(u[:, 4]), u[:, 3] * np.cos(u[:, 5])], 1)
    return matrix_to_quat(euler_zxz(u[:, :3])), t


def calc_en_unred(unred, potential=None, device="cuda"):
    """Label energy of each unreduced table row on the GPU (float64 array, FP32 arithmetic)."""
    pot = potential or BeadPotential(device=device)
    dev = pot.device
    qj, t = unred_pairs(unred)
    P = len(qj)
    if P == 0:
        return np.zeros(0)
    qf = np.zeros((2 * P, 4), np.float32)
    qf[:P, 0] = 1.0
    qf[P:] = qj
    pt = np.zeros((P, 4), np.float32)
    pt[:, :3] = t
    qf_d = torch.from_numpy(qf).to(dev)
    pi = torch.arange(P, dtype=torch.int32, device=dev)
    pj = pi + P
    pt_d = torch.from_numpy(pt).to(dev)
    out = torch.empty(P * 12, dtype=torch.float32, device=dev)
    n_dev = torch.tensor([P], dtype=torch.int32, device=dev)
    pot.pair_forces(_lib.ptr(qf_d), _lib.ptr(pi), _lib.ptr(pj), _lib.ptr(pt_d), _lib.ptr(n_dev), P, _lib.ptr(out))
    return out.view(P, 12)[:, 3].double().cpu().numpy()
