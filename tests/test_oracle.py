"""Pins the FP64 oracle against library routines, closed forms, invariants and brute force.

Nothing here re-types an oracle formula and compares it with itself: every
check uses an independent route (scipy special functions / Rotation /
cKDTree, torch nn.Linear + autograd, finite differences, the addition
theorem, symmetry, or a closed-form solution).
"""
import math
import os

import numpy as np
import pytest
import torch
from scipy.spatial import cKDTree
from scipy.spatial.transform import Rotation
from scipy.special import eval_chebyt, eval_legendre

from oracle import nep_oracle as O

HYP = (11, 6, 4)


def _rng(s=0):
    return np.random.default_rng(s)


def _unit_q(rng, n):
    q = rng.standard_normal((n, 4))
    return q / np.linalg.norm(q, axis=1, keepdims=True)


# ---------------------------------------------------------------- placement
def test_quat_rotate_matches_scipy_rotation():
    rng = _rng(1)
    q = _unit_q(rng, 50)
    v = rng.standard_normal((50, 3))
    want = Rotation.from_quat(q[:, [1, 2, 3, 0]]).apply(v)   # scipy is scalar-last
    assert np.allclose(O.quat_rotate(q, v), want, atol=1e-13)
    # R(q) R(q*) = I and |R b| = |b|
    assert np.allclose(O.quat_rotate(O.quat_conj(q), O.quat_rotate(q, v)), v, atol=1e-13)
    assert np.allclose(np.linalg.norm(O.quat_rotate(q, v), axis=1), np.linalg.norm(v, axis=1))


def test_quat_mul_matches_scipy_composition():
    rng = _rng(2)
    a, b = _unit_q(rng, 20), _unit_q(rng, 20)
    ab = O.quat_mul(a, b)
    want = (Rotation.from_quat(a[:, [1, 2, 3, 0]]) * Rotation.from_quat(b[:, [1, 2, 3, 0]])).as_quat()
    want = want[:, [3, 0, 1, 2]]
    sign = np.sign(np.sum(ab * want, axis=1, keepdims=True))
    assert np.allclose(ab, sign * want, atol=1e-13)


def test_shape_tables():
    T = O.shape_table()
    tet = T["tet10"]["sites"]
    d = np.linalg.norm(tet[:4, None] - tet[None, :4], axis=-1)
    assert np.allclose(d[np.triu_indices(4, 1)], 1.0)           # regular tetrahedron, edge 1
    assert np.allclose(tet.mean(0), 0.0)
    cube = T["cube8"]["sites"]
    assert np.allclose(np.abs(cube), 0.5) and len(cube) == 8
    cyl = T["cyl12"]["sites"]
    assert len(cyl) == 12 and np.allclose(cyl.mean(0), 0.0)
    assert np.isclose(np.ptp(cyl[:, 2]), 3.0) and np.allclose(np.diff(cyl[:, 2]), 3.0 / 11)
    # circle through three sites in the xy plane: every site at radius 0.5, azimuth spread 90 degrees
    (ax, ay), (bx, by), (cx, cy) = cyl[0, :2], cyl[6, :2], cyl[11, :2]
    d = 2 * (ax * (by - cy) + bx * (cy - ay) + cx * (ay - by))
    ux = ((ax * ax + ay * ay) * (by - cy) + (bx * bx + by * by) * (cy - ay) + (cx * cx + cy * cy) * (ay - by)) / d
    uy = ((ax * ax + ay * ay) * (cx - bx) + (bx * bx + by * by) * (ax - cx) + (cx * cx + cy * cy) * (bx - ax)) / d
    rel = cyl[:, :2] - np.array([ux, uy])
    assert np.allclose(np.linalg.norm(rel, axis=1), 0.5)
    ang = np.unwrap(np.arctan2(rel[:, 1], rel[:, 0]))
    assert np.isclose(abs(ang[-1] - ang[0]), np.pi / 2)


# ---------------------------------------------------------------- bases
def test_chebyshev_basis_against_scipy_and_closed_forms():
    rho = np.linspace(0.0, 1.3, 41)
    f, df = O._cheb_basis(rho, 11)
    x = 2 * (rho - 1) ** 2 - 1
    fc = 0.5 * (1 + np.cos(np.pi * rho))
    for n in range(12):
        assert np.allclose(f[n], 0.5 * (eval_chebyt(n, x) + 1) * fc, atol=1e-13)
    th = np.linspace(0, np.pi, 17)
    for n in range(12):                                      # T_n(cos t) = cos(n t)
        assert np.allclose(eval_chebyt(n, np.cos(th)), np.cos(n * th), atol=1e-12)
    assert np.allclose(f[:, 0], 1.0)                         # rho = 0: x = 1, fc = 1
    assert np.allclose(O._cheb_basis(np.array([1.0]), 11)[0], 0.0, atol=1e-15)  # rho = 1: fc = 0
    h = 1e-6
    num = (O._cheb_basis(rho + h, 11)[0] - O._cheb_basis(rho - h, 11)[0]) / (2 * h)
    assert np.allclose(df, num, atol=1e-6)


def test_legendre_against_scipy():
    c = np.linspace(-1, 1, 33)
    P, dP = O._legendre(c, 6)
    for l in range(7):
        assert np.allclose(P[l], eval_legendre(l, c), atol=1e-13)
    h = 1e-6
    cc = np.linspace(-0.9, 0.9, 11)
    num = (O._legendre(cc + h, 6)[0] - O._legendre(cc - h, 6)[0]) / (2 * h)
    assert np.allclose(O._legendre(cc, 6)[1], num, atol=1e-6)


# ---------------------------------------------------------------- descriptor
def _clouds(seed, P=6, nd=20, scale=1.2):
    return _rng(seed).normal(scale=scale, size=(P, nd, 3))


def test_descriptor_brute_force_loop():
    """Pure-Python double loop over all (a, b), diagonal included (SingleKernelDebug.py:23-73)."""
    cl = _clouds(3, P=2, nd=8)
    r_nep, (nrad, nang, lmax) = 2.5, (5, 3, 3)
    G = O.descriptor(cl, r_nep, nrad, nang, lmax)
    for p in range(2):
        c = cl[p] - cl[p].mean(0)
        r = np.linalg.norm(c, axis=1)
        rho = r / r_nep
        x = 2 * (rho - 1) ** 2 - 1
        fc = 0.5 * (1 + np.cos(np.pi * rho))
        f = [[0.5 * (eval_chebyt(n, x[a]) + 1) * fc[a] for a in range(8)] for n in range(nrad + 1)]
        want = [sum(f[n]) for n in range(nrad + 1)]
        for n in range(nang + 1):
            for l in range(1, lmax + 1):
                s = 0.0
                for a in range(8):
                    for b in range(8):
                        s += f[n][a] * f[n][b] * eval_legendre(l, np.dot(c[a], c[b]) / (r[a] * r[b]))
                want.append(s)
        assert np.allclose(G[p], want, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("nd,hyp", [(20, (11, 6, 4)), (24, (17, 6, 6)), (12, (10, 8, 4)), (16, (11, 6, 4))])
def test_addition_theorem_matches_direct_form(nd, hyp):
    cl = _clouds(4, P=8, nd=nd)
    G1 = O.descriptor(cl, 2.5, *hyp)
    G2 = O.descriptor_harmonic(cl, 2.5, *hyp)
    assert np.max(np.abs(G1 - G2)) < 1e-12 * np.max(np.abs(G1))


def test_descriptor_invariances():
    cl = _clouds(5, P=4)
    G = O.descriptor(cl, 2.5, *HYP)
    R = Rotation.random(random_state=7).as_matrix()
    assert np.allclose(O.descriptor(cl @ R.T + np.array([3.0, -1.0, 0.5]), 2.5, *HYP), G, rtol=1e-12, atol=1e-12)
    perm = _rng(6).permutation(20)
    assert np.allclose(O.descriptor(cl[:, perm], 2.5, *HYP), G, rtol=1e-12, atol=1e-12)


def test_pair_cloud_symmetric_under_swap():
    rng = _rng(8)
    sites = O.shape_table()["tet10"]["sites"]
    X = rng.uniform(-3, 3, (2, 3))
    Q = _unit_q(rng, 2)
    a, _ = O.pair_clouds(X, Q, np.array([[0, 1]]), sites, 20.0)
    b, _ = O.pair_clouds(X, Q, np.array([[1, 0]]), sites, 20.0)
    assert np.allclose(O.descriptor(a, 2.5, *HYP), O.descriptor(b, 2.5, *HYP), rtol=1e-12)


def _torch_descriptor(cl, r_nep, nrad, nang, lmax):
    """Autograd-able torch form (library differentiation of the same function)."""
    c = cl - cl.mean(dim=1, keepdim=True)
    r = torch.linalg.norm(c, dim=-1)
    u = c / r[..., None]
    rho = r / r_nep
    x = 2 * (rho - 1) ** 2 - 1
    fc = 0.5 * (1 + torch.cos(math.pi * rho))
    T = [torch.ones_like(x), x]
    for n in range(1, nrad):
        T.append(2 * x * T[n] - T[n - 1])
    f = torch.stack([(0.5 * (t + 1)) * fc for t in T[: nrad + 1]])
    cos = torch.einsum("pak,pbk->pab", u, u)
    P = [torch.ones_like(cos), cos]
    for l in range(1, lmax):
        P.append(((2 * l + 1) * cos * P[l] - l * P[l - 1]) / (l + 1))
    out = [f.sum(-1).T]
    for n in range(nang + 1):
        for l in range(1, lmax + 1):
            out.append(torch.einsum("pa,pab,pb->p", f[n], P[l], f[n])[:, None])
    return torch.cat(out, dim=1)


def test_descriptor_grad_against_autograd_and_fd():
    cl = _clouds(9, P=5)
    dG = _rng(10).standard_normal((5, 40))
    g = O.descriptor_grad(cl, dG, 2.5, *HYP)
    t = torch.tensor(cl, requires_grad=True)
    Gt = _torch_descriptor(t, 2.5, *HYP)
    assert np.allclose(Gt.detach().numpy(), O.descriptor(cl, 2.5, *HYP), rtol=1e-12, atol=1e-12)
    (Gt * torch.tensor(dG)).sum().backward()
    assert np.allclose(t.grad.numpy(), g, rtol=1e-10, atol=1e-10)
    h = 1e-6
    for (p, a, k) in [(0, 0, 0), (1, 7, 2), (4, 19, 1), (3, 10, 0)]:
        cp, cm = cl.copy(), cl.copy()
        cp[p, a, k] += h
        cm[p, a, k] -= h
        num = np.sum((O.descriptor(cp, 2.5, *HYP) - O.descriptor(cm, 2.5, *HYP))[p] * dG[p]) / (2 * h)
        assert abs(num - g[p, a, k]) < 1e-6 * max(1.0, abs(num))


def test_reference_cloud_fixture():
    """Cube pair clouds shipped with the reference (01/datadir/x_12pts_500_s42.npy, 64 sampled)."""
    import os
    path = os.path.join(os.path.dirname(__file__), "golden", "x12pts_cube_v2_sample.npz")
    z = np.load(path)
    cl = z["clouds"].astype(np.float64)
    G = O.descriptor(cl, 4.5, 10, 8, 4)
    assert np.allclose(G, z["G"], rtol=1e-12, atol=1e-12)
    assert np.allclose(O.descriptor_harmonic(cl, 4.5, 10, 8, 4), G, rtol=1e-11, atol=1e-11)
    # the clouds have the cube_v2 site geometry: 3 antipodal pairs per body at radius 1.7
    for body in (cl[:, :6], cl[:, 6:]):
        ctr = body.mean(1, keepdims=True)
        assert np.allclose(np.linalg.norm(body - ctr, axis=-1), 1.7, atol=1e-5)


# ---------------------------------------------------------------- pairs
@pytest.mark.parametrize("n,L,r", [(300, 10.0, 2.5), (150, 6.0, 3.0), (40, 5.1, 2.5)])
def test_brute_pairs_match_ckdtree(n, L, r):
    rng = _rng(n)
    X = rng.uniform(-L / 2, L / 2, (n, 3))
    got = {tuple(p) for p in O.brute_pairs(X, L, r)}
    tree = cKDTree(X + L / 2, boxsize=L)
    want = {tuple(sorted(p)) for p in tree.query_pairs(r)}
    assert got == want


# ---------------------------------------------------------------- network
def test_ffnn_matches_torch_linear():
    rng = _rng(11)
    layers = O.random_layers(40, [64, 64], seed=3)
    G = rng.uniform(0, 3, (30, 40))
    for act, mod in (("tanh", torch.nn.Tanh), ("relu", torch.nn.ReLU)):
        net = torch.nn.Sequential(torch.nn.Linear(40, 64), mod(), torch.nn.Linear(64, 64), mod(),
                                  torch.nn.Linear(64, 1)).double()
        with torch.no_grad():
            for lin, (W, b) in zip([net[0], net[2], net[4]], layers):
                lin.weight.copy_(torch.tensor(W))
                lin.bias.copy_(torch.tensor(b + 0.1))
        lay = [(W, b + 0.1) for W, b in layers]
        x = torch.tensor(G, requires_grad=True)
        y = net(x)[:, 0]
        y.sum().backward()
        yo, go = O.ffnn_forward_backward(G, lay, act)
        assert np.allclose(yo, y.detach().numpy(), rtol=1e-12, atol=1e-12)
        assert np.allclose(go, x.grad.numpy(), rtol=1e-12, atol=1e-12)


# ---------------------------------------------------------------- forces
def _small_system(seed=12, n=24, L=7.0):
    rng = _rng(seed)
    # jittered lattice, no pair within 1e-3 of the cutoff so central FD stays on one branch
    while True:
        X = rng.uniform(-L / 2, L / 2, (n, 3))
        d = np.sqrt(O.pair_d2(X, *np.triu_indices(n, 1), L))
        if np.all(np.abs(d - 2.5) > 1e-3) and d.min() > 0.8:
            return X, _unit_q(rng, n)


def test_pair_forces_match_central_fd():
    X, Q = _small_system()
    sh = O.shape_table()["tet10"]
    L = 7.0
    layers = O.random_layers(40, [64, 64], seed=5, std_scale=0.05)
    F, T, pairs, _ = O.system_forces(X, Q, sh, L, HYP, layers)
    assert len(pairs) > 10
    h = 1e-4
    E = lambda X_, Q_: O.total_energy(X_, Q_, sh, L, HYP, layers)
    for i in (0, 5, 17):
        for k in range(3):
            Xp, Xm = X.copy(), X.copy()
            Xp[i, k] += h
            Xm[i, k] -= h
            num = -(E(Xp, Q) - E(Xm, Q)) / (2 * h)
            assert abs(num - F[i, k]) < 1e-6 * max(1e-3, np.abs(F).max())
            e = np.zeros(3)
            e[k] = 1.0
            dq = lambda s: np.concatenate([[math.cos(s * h / 2)], math.sin(s * h / 2) * e])
            Qp, Qm = Q.copy(), Q.copy()
            Qp[i] = O.quat_mul(dq(1), Q[i])           # lab-frame rotation about X_i
            Qm[i] = O.quat_mul(dq(-1), Q[i])
            num = -(E(X, Qp) - E(X, Qm)) / (2 * h)
            assert abs(num - T[i, k]) < 1e-6 * max(1e-3, np.abs(T).max())


def test_pair_balance_laws():
    X, Q = _small_system(seed=13)
    sh = O.shape_table()["tet10"]
    pairs = O.brute_pairs(X, 7.0, 2.5)
    r = O.pair_forces(X, Q, pairs, sh, 7.0, HYP, O.random_layers(40, [64, 64], seed=6))
    # angular momentum about the origin: tau_i + tau_j + t x F_j = 0 with F_j = -F_i
    assert np.allclose(r["Ti"] + r["Tj"] - np.cross(r["t"], r["F"]), 0.0, atol=1e-10 * np.abs(r["F"]).max())


def test_reference_fd_stencil_converges_to_analytic():
    X, Q = _small_system(seed=14)
    sh = O.shape_table()["tet10"]
    L = 7.0
    layers = O.random_layers(40, [64, 64], seed=7, std_scale=0.05)
    pairs = O.brute_pairs(X, L, 2.5)
    r = O.pair_forces(X, Q, pairs, sh, L, HYP, layers)
    clouds, t = O.pair_clouds(X, Q, pairs, sh["sites"], L)
    f, ti, tj = O.fd_pair_forces(clouds, 10, t, 2.5, HYP, layers, "tanh", 0.0, 1.0, 1e-6, 1e-6)
    sc = np.abs(r["F"]).max()
    assert np.abs(f - r["F"]).max() < 1e-4 * sc
    assert np.abs(ti - r["Ti"]).max() < 1e-4 * np.abs(r["Ti"]).max()
    assert np.abs(tj - r["Tj"]).max() < 1e-4 * np.abs(r["Tj"]).max()


# ---------------------------------------------------------------- integrator
def _state(n, seed):
    rng = _rng(seed)
    q = _unit_q(rng, n)
    Lb = rng.standard_normal((n, 3))
    p = 2.0 * O.quat_mul(q, np.concatenate([np.zeros((n, 1)), Lb], 1))
    return dict(x=rng.uniform(-2, 2, (n, 3)), v=rng.standard_normal((n, 3)), q=q, p=p, m=np.ones(n)), Lb


def test_free_particle_translation_and_rot_ke_convention():
    st, Lb = _state(10, 20)
    moi = (0.3, 0.5, 0.7)
    th = dict(xi=0, eta=0, xi_r=0, eta_r=0, theta=1.0, e_r=1.0)
    new, Kt, Kr = O.nvt_step(st, np.zeros((10, 3)), np.zeros((10, 3)), 0.01, 1.0, 1.0, 100.0, moi, th)
    assert np.allclose(new["x"], st["x"] + 0.01 * st["v"], rtol=0, atol=1e-15)
    assert np.allclose(new["v"], st["v"])
    # s = q* p / 2 is the body angular momentum: K_r = sum L^2 / 2I at the initial state
    assert np.allclose(O.rot_ke(st["q"], st["p"], moi), np.sum(Lb ** 2 / (2 * np.array(moi)), 1))


def test_spherical_top_closed_form():
    """I1 = I2 = I3: the body rotates at constant omega = L/I about a fixed lab axis."""
    st, Lb = _state(5, 21)
    I = 0.4
    dt = 0.003
    th = dict(xi=0, eta=0, xi_r=0, eta_r=0, theta=1.0, e_r=1.0)
    s = st
    for _ in range(50):
        s, _, _ = O.nvt_step(s, np.zeros((5, 3)), np.zeros((5, 3)), dt, 1.0, 1.0, 100.0, (I, I, I), th)
    Llab = O.quat_rotate(st["q"], Lb)                    # lab-frame angular momentum
    w = Llab / I
    ang = np.linalg.norm(w, axis=1) * dt * 50
    axis = w / np.linalg.norm(w, axis=1, keepdims=True)
    rot = np.concatenate([np.cos(ang / 2)[:, None], np.sin(ang / 2)[:, None] * axis], 1)
    want = O.quat_mul(rot, st["q"])
    sign = np.sign(np.sum(want * s["q"], 1, keepdims=True))
    assert np.allclose(s["q"], sign * want, atol=1e-10)
    assert np.allclose(O.quat_rotate(s["q"], 0.5 * O.quat_mul(O.quat_conj(s["q"]), s["p"])[:, 1:]), Llab, atol=1e-10)


def test_no_squish_time_reversible():
    st, _ = _state(6, 22)
    moi = (0.2, 0.6, 0.9)
    p1, q1 = O.no_squish(st["p"], st["q"], moi, 0.01)
    p0, q0 = O.no_squish(p1, q1, moi, -0.01)
    assert np.allclose(p0, st["p"], atol=1e-13) and np.allclose(q0, st["q"], atol=1e-13)


def test_thermostat_fixed_point_and_direction():
    th = dict(xi=0.2, eta=0.1, xi_r=-0.1, eta_r=0.0)
    N_dof, kT = 30.0, 1.5
    K_eq = 0.5 * N_dof * kT                              # T = kT exactly
    u = O.thermo_update(th, K_eq, K_eq, N_dof, N_dof, 0.01, 1.0, kT)
    assert u["xi"] == th["xi"] and u["xi_r"] == th["xi_r"]
    assert np.isclose(u["eta"], th["eta"] + 0.2 * 0.01)
    assert np.isclose(u["theta"], math.exp(-0.5 * 0.2 * 0.01))
    hot = O.thermo_update(th, 2 * K_eq, 0.5 * K_eq, N_dof, N_dof, 0.01, 1.0, kT)
    assert hot["xi"] > th["xi"] and hot["xi_r"] < th["xi_r"]


def test_descriptor_weights_equal_site_multiplicity():
    """Integer per-site weights k_a give the descriptor of the cloud with site a repeated k_a times
    (the radial sum and both angular sums count it k_a times).  The clouds are point-symmetric with
    symmetric weights, so the repeated cloud keeps the centroid the weighted one is taken about."""
    rng = np.random.default_rng(11)
    X2 = rng.normal(0, 1, (5, 3, 3))
    sym = np.concatenate([X2, -X2], axis=1)
    k = np.array([2, 1, 3, 2, 1, 3])
    G_w = O.descriptor(sym, 3.0, 5, 4, 3, weights=k.astype(float))
    G_r = O.descriptor(np.repeat(sym, k, axis=1), 3.0, 5, 4, 3)
    assert np.allclose(G_w, G_r, rtol=1e-12, atol=1e-12)
    assert not np.allclose(G_w, O.descriptor(sym, 3.0, 5, 4, 3))


# ---- label potential (01/SamplerUtils.py:555-631) ----------------------------

def test_cube_beads_equal_reference_structure_file():
    from nepx import gsd
    x = gsd.read_frame(os.path.join(os.path.dirname(__file__), "golden", "cube_structure.gsd"))["particles/position"]
    assert np.abs(O.cube_structure_beads() - x).max() < 1e-6


def test_euler_matrices_are_intrinsic_zxz():
    from scipy.spatial.transform import Rotation
    a = np.random.default_rng(0).uniform(-3.2, 3.2, (20, 3))
    assert np.abs(Rotation.from_euler("ZXZ", a).as_matrix() - O.euler_zxz(a)).max() < 1e-12


def test_bead_energy_closed_form_and_symmetry():
    one = np.zeros((1, 3))
    X = np.array([[0.0, 0.0, 0.0], [0.8, 0.3, -0.2]])
    Q = np.array([[1.0, 0, 0, 0], [1.0, 0, 0, 0]])
    r = O.bead_pair_forces(X, Q, np.array([[0, 1]]), one, 50.0)
    d = np.linalg.norm(X[1])
    assert np.isclose(r["E"][0], 5.0 * ((2 ** (1 / 6) - d) / 1.1) ** 2, rtol=1e-14)
    far = O.bead_pair_forces(np.array([[0.0, 0, 0], [1.2, 0, 0]]), Q, np.array([[0, 1]]), one, 50.0)
    assert far["E"][0] == 0.0 and np.all(far["F"] == 0.0)
    beads = O.cube_structure_beads()
    rng = np.random.default_rng(2)
    X = np.array([[0.0, 0, 0], [5.0, 1.0, -0.5]])
    Q = O.quat_mul(np.eye(4)[[0, 0]], rng.normal(size=(2, 4)))
    Q /= np.linalg.norm(Q, axis=1, keepdims=True)
    a = O.bead_pair_forces(X, Q, np.array([[0, 1]]), beads, 50.0)
    b = O.bead_pair_forces(X[::-1].copy(), Q[::-1].copy(), np.array([[0, 1]]), beads, 50.0)
    assert a["E"][0] > 0 and np.isclose(a["E"][0], b["E"][0], rtol=1e-12)
    assert np.allclose(a["F"], -b["F"], rtol=1e-10, atol=1e-12)
    assert np.allclose(a["Ti"], b["Tj"], rtol=1e-10, atol=1e-12)


def test_bead_forces_central_fd():
    beads = O.cube_structure_beads()
    rng = np.random.default_rng(3)
    X = np.array([[0.0, 0, 0], [4.6, 1.3, -0.9]])
    Q = rng.normal(size=(2, 4))
    Q /= np.linalg.norm(Q, axis=1, keepdims=True)
    pr = np.array([[0, 1]])
    r = O.bead_pair_forces(X, Q, pr, beads, 50.0)
    assert r["E"][0] > 0
    h = 1e-6
    for k in range(3):
        Xp, Xm = X.copy(), X.copy()
        Xp[0, k] += h
        Xm[0, k] -= h
        fd = -(O.bead_pair_forces(Xp, Q, pr, beads, 50.0)["E"][0] - O.bead_pair_forces(Xm, Q, pr, beads, 50.0)["E"][0]) / (2 * h)
        assert abs(fd - r["F"][0, k]) < 1e-5 * np.abs(r["F"]).max()
        for body, key in ((0, "Ti"), (1, "Tj")):
            dq = np.zeros(4)
            dq[0], dq[1 + k] = np.cos(h / 2), np.sin(h / 2)
            Qp, Qm = Q.copy(), Q.copy()
            Qp[body] = O.quat_mul(dq, Q[body])
            Qm[body] = O.quat_mul(dq * np.array([1, -1, -1, -1]), Q[body])
            fd = -(O.bead_pair_forces(X, Qp, pr, beads, 50.0)["E"][0] - O.bead_pair_forces(X, Qm, pr, beads, 50.0)["E"][0]) / (2 * h)
            assert abs(fd - r[key][0, k]) < 1e-5 * np.abs(r[key]).max()


def test_calc_en_unred_equals_pair_form():
    """The unreduced-table energy equals the pair form with body 2's quaternion from the ZXZ matrix."""
    from scipy.spatial.transform import Rotation
    beads = O.cube_structure_beads()
    rng = np.random.default_rng(4)
    u = np.column_stack([rng.uniform(-3.14, 3.14, 6), rng.uniform(0, 3.14, 6), rng.uniform(-3.14, 3.14, 6),
                         rng.uniform(4.3, 6.0, 6), rng.uniform(-3.14, 3.14, 6), rng.uniform(0, 3.14, 6)])
    E = O.calc_en_unred(u, beads)
    assert (E > 0).sum() >= 4
    for k in range(6):
        qx = Rotation.from_matrix(O.euler_zxz(u[k:k + 1, :3])).as_quat()[0]      # scalar last
        q = np.array([[1.0, 0, 0, 0], [qx[3], qx[0], qx[1], qx[2]]])
        st = np.sin(u[k, 5])
        t = u[k, 3] * np.array([st * np.cos(u[k, 4]), st * np.sin(u[k, 4]), np.cos(u[k, 5])])
        r = O.bead_pair_forces(np.array([[0.0, 0, 0], t]), q, np.array([[0, 1]]), beads, 100.0)
        assert np.isclose(r["E"][0], E[k], rtol=1e-12)
