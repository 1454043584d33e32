"""Edge cases of the step: a system with no pair inside the cutoff (empty compacted pair set) batched
with a two-body system, stepped against the FP64 oracle integrator (1e-12), and a single-pair
system's forces against the oracle (1e-4)."""
import numpy as np
import pytest
import torch

from oracle import nep_oracle as O

pytestmark = pytest.mark.gpu
HYP = (11, 6, 4)


def _lattice_spec(n_side, spacing, seed):
    from nepx.configs import random_quaternions
    from nepx.system import SystemSpec
    rng = np.random.default_rng(seed)
    g = (np.arange(n_side) - (n_side - 1) / 2) * spacing
    x = np.array([(a, b, c) for a in g for b in g for c in g], dtype=np.float64)
    n = len(x)
    q = random_quaternions(rng, n)
    v = rng.normal(0, 1, (n, 3))
    L_body = rng.normal(0, 0.3, (n, 3))
    p = 2.0 * O.quat_mul(q, np.concatenate([np.zeros((n, 1)), L_body], axis=1))
    moi = (1 / 12, 1 / 12, 1 / 12)
    return SystemSpec(x, v, q, p, 1.0, n_side * spacing, 1.0, 0.001, moi, 2.5)


def test_empty_pair_set_steps_like_free_flight(gpu):
    import nepx
    dilute = _lattice_spec(4, 3.0, 1)                    # 64 bodies 3.0 apart: no pair within rc = 2.5
    two = _lattice_spec(3, 1.8, 2)                       # 27 bodies 1.8 apart: many pairs
    net = nepx.EnergyNet.random(40, [64, 64], seed=5, device=gpu)
    ar = nepx.Arena([dilute, two], nepx.get_shape("tet10"), HYP, net, device=gpu)
    ar.init()
    th0 = ar.thermo()[0]
    s0 = ar.system_arrays(0)
    ar.step()
    torch.cuda.synchronize()
    ar.check()
    s1 = ar.system_arrays(0)
    assert np.all(s1["F"] == 0.0) and np.all(s1["T"] == 0.0) and np.all(s1["upot"] == 0.0)
    assert np.any(ar.system_arrays(1)["F"] != 0.0)
    th = {k: th0[k] for k in ("xi", "eta", "xi_r", "eta_r", "theta", "e_r")}
    new, Kt, Kr = O.nvt_step(dict(x=s0["x"], v=s0["v"], q=s0["q"], p=s0["p"], m=s0["mass"]), np.zeros((64, 3)),
                             np.zeros((64, 3)), dilute.dt, dilute.kT, dilute.tau, dilute.L, dilute.moi, th)
    for k in "xvqp":
        e = np.max(np.abs(s1[k] - new[k])) / np.max(np.abs(new[k]))
        assert e < 1e-12, (k, e)


def test_single_pair_forces(gpu):
    import nepx
    from nepx.system import SystemSpec
    x = np.array([[0.0, 0.0, 0.0], [1.3, 0.4, -0.5]])
    q = np.array([[1.0, 0, 0, 0], [0.5, 0.5, 0.5, 0.5]])
    spec = SystemSpec(x, np.zeros((2, 3)), q, np.zeros((2, 4)), 1.0, 6.0, 1.0, 0.001, (0.1, 0.1, 0.1), 2.5)
    net = nepx.EnergyNet.random(40, [64, 64], seed=8, device=gpu)
    ar = nepx.Arena([spec], nepx.get_shape("tet10"), HYP, net, device=gpu)
    ar.init()
    ar.filter_pairs()
    ar.pair_fused()
    ar.gather()
    torch.cuda.synchronize()
    got = ar.system_arrays(0)
    assert ar.n_pairs() == 1
    F, T, pairs, r = O.system_forces(x, q.astype(np.float32).astype(np.float64), O.shape_table()["tet10"], 6.0,
                                     HYP, net.layers_f64())
    assert len(pairs) == 1
    assert np.abs(got["F"] - F).max() < 1e-4 * np.abs(F).max()
    assert np.abs(got["T"] - T).max() < 1e-4 * np.abs(T).max()
    assert np.allclose(got["F"][0], -got["F"][1], rtol=0, atol=1e-12)
