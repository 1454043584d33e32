"""The hand-written adjoint behind torch.autograd, checked with FP64 central finite differences.

* small system: autograd gradients of the GPU energy (dU/dX = -F, dU/drot = -tau) against central
  differences (h = 1e-4) of the FP64 oracle energy and against the oracle's analytic gradient;
* the 32,768-tetrahedron workload: GPU forces/torques of 256 sampled particles against central
  differences (h = 1e-4, FP64) of each particle's pair energies.
Tolerance: 1e-4 relative (max-norm over the checked components).
"""
import math

import numpy as np
import pytest
import torch

from oracle import nep_oracle as O

pytestmark = pytest.mark.gpu
HYP = (11, 6, 4)


def _axis_quat(ax, ang):
    q = np.zeros(4)
    q[0] = math.cos(0.5 * ang)
    q[1 + ax] = math.sin(0.5 * ang)
    return q


def test_autograd_function_matches_fd_and_oracle(gpu):
    import nepx
    from nepx.autograd import nep_energy
    from nepx.configs import make_system
    spec = make_system("tet10", 120, 0.08, seed=31)
    net = nepx.EnergyNet.random(40, [64, 64], seed=3, device=gpu)
    ar = nepx.Arena([spec], nepx.get_shape("tet10"), HYP, net, device=gpu)
    x = torch.tensor(spec.x, device=gpu, requires_grad=True)
    rot = torch.zeros((spec.n, 3), dtype=torch.float64, device=gpu, requires_grad=True)
    q32 = torch.tensor(spec.q, device=gpu).to(torch.float32).to(torch.float64)
    U = nep_energy(x, rot, q32, ar)
    U.backward()
    gx, gr = x.grad.cpu().numpy(), rot.grad.cpu().numpy()
    sh = O.shape_table()["tet10"]
    Q = q32.cpu().numpy()
    layers = net.layers_f64()
    F, T, pairs, _ = O.system_forces(spec.x, Q, sh, spec.L, HYP, layers)
    assert len(pairs) > 50
    assert np.abs(gx + F).max() < 1e-4 * np.abs(F).max()
    assert np.abs(gr + T).max() < 1e-4 * np.abs(T).max()
    U64 = O.total_energy(spec.x, Q, sh, spec.L, HYP, layers)
    assert abs(U.item() - U64) < 1e-5 * abs(U64)
    # FP64 central differences on a few particles
    h = 1e-4
    d = np.sqrt(O.pair_d2(spec.x, *np.triu_indices(spec.n, 1), spec.L))
    for i in (0, 17, 63):
        for k in range(3):
            xp, xm = spec.x.copy(), spec.x.copy()
            xp[i, k] += h
            xm[i, k] -= h
            fd = (O.total_energy(xp, Q, sh, spec.L, HYP, layers) - O.total_energy(xm, Q, sh, spec.L, HYP, layers)) / (2 * h)
            if np.min(np.abs(d - 2.5)) > 3 * h:   # no pair crosses the cutoff within the stencil
                assert abs(fd - gx[i, k]) < 1e-4 * np.abs(F).max()
            Qp, Qm = Q.copy(), Q.copy()
            Qp[i] = O.quat_mul(_axis_quat(k, h), Q[i])
            Qm[i] = O.quat_mul(_axis_quat(k, -h), Q[i])
            fd = (O.total_energy(spec.x, Qp, sh, spec.L, HYP, layers) -
                  O.total_energy(spec.x, Qm, sh, spec.L, HYP, layers)) / (2 * h)
            assert abs(fd - gr[i, k]) < 1e-4 * np.abs(T).max()


def _local_fd(X, Q, i, nbrs, sh, L, layers, h):
    """Central-difference force and torque on particle i from its pair energies (one batched oracle call)."""
    pairs = np.stack([np.full(len(nbrs), i), nbrs], 1)
    clouds = []
    for k in range(3):
        for s in (1.0, -1.0):
            Xs = X.copy()
            Xs[i, k] += s * h
            clouds.append(O.pair_clouds(Xs, Q, pairs, sh["sites"], L)[0])
    for k in range(3):
        for s in (1.0, -1.0):
            Qs = Q.copy()
            Qs[i] = O.quat_mul(_axis_quat(k, s * h), Q[i])
            clouds.append(O.pair_clouds(X, Qs, pairs, sh["sites"], L)[0])
    G = O.descriptor(np.concatenate(clouds), sh["nep_cutoff"], *HYP)
    e = O.ffnn_forward(G, layers, "tanh").reshape(12, len(nbrs)).sum(1)
    f = -(e[0:6:2] - e[1:6:2]) / (2 * h)
    t = -(e[6:12:2] - e[7:12:2]) / (2 * h)
    return f, t


def test_tet32k_forces_central_fd_256_particles(gpu):
    import nepx
    from nepx.configs import make_system
    spec = make_system("tet10", 32768, 0.45, seed=2024)
    net = nepx.EnergyNet.random(40, [64, 64], seed=7, device=gpu)
    ar = nepx.Arena([spec], nepx.get_shape("tet10"), HYP, net, device=gpu)
    ar.init()
    ar.filter_pairs()
    ar.pair_fused()
    ar.gather()
    torch.cuda.synchronize()
    got = ar.system_arrays(0)
    sh = O.shape_table()["tet10"]
    layers = net.layers_f64()
    X = spec.x
    Q = spec.q.astype(np.float32).astype(np.float64)
    rng = np.random.default_rng(5)
    h = 1e-4
    r_cut = 2.5
    chosen, F_fd, T_fd, F_gpu, T_gpu = [], [], [], [], []
    for i in rng.permutation(spec.n):
        d = np.sqrt(O.pair_d2(X, np.full(spec.n, i), np.arange(spec.n), spec.L))
        d[i] = np.inf
        if np.min(np.abs(d - r_cut)) < 3 * h:
            continue   # a pair would cross the cutoff inside the stencil
        nb = np.nonzero(d < r_cut)[0]
        f, t = _local_fd(X, Q, i, nb, sh, spec.L, layers, h)
        F_fd.append(f)
        T_fd.append(t)
        F_gpu.append(got["F"][i])
        T_gpu.append(got["T"][i])
        chosen.append(i)
        if len(chosen) == 256:
            break
    F_fd, T_fd, F_gpu, T_gpu = map(np.array, (F_fd, T_fd, F_gpu, T_gpu))
    eF = np.abs(F_gpu - F_fd).max() / np.abs(F_fd).max()
    eT = np.abs(T_gpu - T_fd).max() / np.abs(T_fd).max()
    print(f"tet32k FD check on {len(chosen)} particles: F {eF:.2e} T {eT:.2e}")
    assert len(chosen) == 256
    assert eF < 1e-4 and eT < 1e-4
