"""The reference call surface (DescriptorGenerator / ForceEvaluator / Sim / NVT) on the HIP path.

Inputs: the reference's own initial configuration (200 tetrahedra, box 33,
tests/golden/th_init_N200.gsd) with the th_v1 shape and hypers (10, 8, 4),
and a seeded TheEnergyNet(150, 3, 47) with ReLU (the reference network
layout; the shipped checkpoints are not used).  Tolerances as in
test_gpu_pair.py; the forward-difference path is compared with the FP64
oracle running the same stencil.
"""
import os

import numpy as np
import pytest
import torch
from scipy.spatial import cKDTree

from oracle import nep_oracle as O

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")
HYP = (10, 8, 4)


def _init():
    from nepx.GsdHandler import GsdHandler
    return GsdHandler(os.path.join(GOLD, "th_init_N200.gsd"))


def _net(gpu, act="relu", D=47):
    import nepx
    return nepx.EnergyNet.random(D, [150, 150, 150, 150], seed=21, act=act, en_min=-6.7, en_max=15.0, device=gpu)


def _kdtree_pairs(x, L, r):
    tree = cKDTree(x + 0.5 * L, leafsize=12, boxsize=L + 0.0001)   # as Sim.setNeighborList (Sim.py:85-92)
    return tree.query_pairs(r=r, output_type="ndarray")


def _dg_fe(gpu, h, net, fd=False, dx=1e-5, hyp=HYP):
    from nepx.DescriptorGenerator import DescriptorGenerator
    from nepx.ForceEvaluator import ForceEvaluator
    dg = DescriptorGenerator()
    dg.setDevice(0)
    dg.setHyperParameters([*hyp, 4.5])
    dg.setBoxSize(h.getLx())
    dg.set_dx_dteta(dx, dx)
    dg.setShape("th_v1")
    dg.move_data_to_device()
    dg.fd_compat = fd
    fe = ForceEvaluator()
    fe.setDevice(0)
    fe.setNN_dimensions(150, 3)
    fe.setNd(dg.getN_descriptor())
    fe.set_model(net)
    fe.set_energy_range(*dg.get_en_min_max())
    fe.setNparticles(len(h.get_central_pos()))
    fe.attach(dg)
    return dg, fe


@pytest.mark.parametrize("hyp", [HYP, (10, 10, 10)])
def test_descriptor_generator_and_force_evaluator(gpu, hyp):
    """(10, 10, 10) is the reference's documented run (sim_input.py:10: th_v1, --hypers 10 10 10 4.5,
    a TheEnergyNet(150, 3, 121) checkpoint); D = 121 takes eight l-range passes per pair."""
    import nepx
    h = _init()
    x = h.get_central_pos().astype(np.float64)
    q = h.getOrientations().astype(np.float64)
    L = h.getLx()
    net = _net(gpu, D=nepx.n_descriptor(*hyp))
    dg, fe = _dg_fe(gpu, h, net, hyp=hyp)
    nlist = torch.from_numpy(_kdtree_pairs(x, L, 5.75 + 2.0)).to(gpu)
    G, pp, P = dg.generate_nep_descriptors(torch.from_numpy(x).to(gpu), torch.from_numpy(q).to(gpu), nlist)
    want = O.brute_pairs(x, L, 5.75)
    assert P == len(want) and np.array_equal(pp.cpu().numpy(), want)
    sh = O.shape_table()["th_v1"]
    q32 = q.astype(np.float32).astype(np.float64)
    clouds, _ = O.pair_clouds(x, q32, want, sh["sites"], L)
    G64 = O.descriptor(clouds, 4.5, *hyp)
    eG = np.max(np.abs(G.cpu().numpy() - G64), axis=1) / np.max(np.abs(G64), axis=1)
    F, T = fe.evaluate_interactions(G, pp, P)
    F64, T64, _, r = O.system_forces(x, q32, sh, L, hyp, net.layers_f64(), "relu", -6.7, 15.0)
    eF = np.abs(F.cpu().numpy() - F64).max() / np.abs(F64).max()
    eT = np.abs(T.cpu().numpy() - T64).max() / np.abs(T64).max()
    eE = np.abs(fe.e0.cpu().numpy()[:, 0] - r["E"]).max() / np.abs(r["E"]).max()
    print(f"th_v1 N=200 hypers {hyp}: {P} pairs, G {eG.max():.2e} E {eE:.2e} F {eF:.2e} T {eT:.2e}")
    assert eG.max() < 1e-5 and eE < 1e-5 and eF < 1e-4 and eT < 1e-4


def test_fd_compat_stencil_matches_oracle(gpu):
    h = _init()
    x = h.get_central_pos().astype(np.float64)
    q = h.getOrientations().astype(np.float64)
    L = h.getLx()
    net = _net(gpu)
    dx = 1e-3
    dg, fe = _dg_fe(gpu, h, net, fd=True, dx=dx)
    nlist = torch.from_numpy(_kdtree_pairs(x, L, 7.75)).to(gpu)
    G, pp, P = dg.generate_nep_descriptors(torch.from_numpy(x).to(gpu), torch.from_numpy(q).to(gpu), nlist)
    assert G.shape == (10 * P, 47)
    sh = O.shape_table()["th_v1"]
    q32 = q.astype(np.float32).astype(np.float64)
    pairs = pp.cpu().numpy()
    clouds, t = O.pair_clouds(x, q32, pairs, sh["sites"], L)
    st = O.stencil_clouds(clouds, 4, t, dx, dx)
    G64 = O.descriptor(st.reshape(-1, 8, 3), 4.5, *HYP)
    eG = np.max(np.abs(G.cpu().numpy() - G64), axis=1) / np.max(np.abs(G64), axis=1)
    assert eG.max() < 1e-5, eG.max()
    F, T = fe.evaluate_interactions(G, pp, P)
    f, ti, tj = O.fd_pair_forces(clouds, 4, t, 4.5, HYP, net.layers_f64(), "relu", -6.7, 15.0, dx, dx)
    F64, T64 = O.scatter_forces(len(x), pairs, f, ti, tj)
    eF = np.abs(F.cpu().numpy() - F64).max() / np.abs(F64).max()
    eT = np.abs(T.cpu().numpy() - T64).max() / np.abs(T64).max()
    # forward differences of FP32 energies carry a rounding error of about eps32 |E| / dx per pair
    # (the reference computes them the same way); bound it and add the 1e-4 relative budget
    E = fe.e0.cpu().numpy()[:, 0]
    deg = np.bincount(pairs.ravel(), minlength=len(x)).max()
    noise = 8 * deg * np.finfo(np.float32).eps * np.abs(E).max() / dx
    dF = np.abs(F.cpu().numpy() - F64).max()
    dT = np.abs(T.cpu().numpy() - T64).max()
    print(f"fd_compat: G {eG.max():.2e} F {eF:.2e} T {eT:.2e} (FP32 FD noise bound {noise:.2e})")
    assert dF < noise + 1e-4 * np.abs(F64).max() and dT < noise + 1e-4 * np.abs(T64).max()


def _nvt(gpu, net, component=False, fd=False, hyp=HYP):
    from nepx.NVT import NVT
    sim = NVT()
    sim.component_path = component
    sim.fd_compat = fd
    sim.setDevice(0)
    sim.initializeSystem(os.path.join(GOLD, "th_init_N200.gsd"))
    sim.setNepDescriptors([*hyp, 4.5])
    sim.setEvaluator(net=net)
    sim.setShape("th_v1")
    sim.setNeighborList()
    sim.setkT(0.5)
    sim.setdt(0.005)
    return sim


@pytest.mark.parametrize("hyp", [HYP, (10, 10, 10)])
def test_sim_fast_path_equals_method_by_method(gpu, hyp):
    """(10, 10, 10): the reference's documented run (sim_input.py:10) with a D = 121 network."""
    import nepx
    net = _net(gpu, D=nepx.n_descriptor(*hyp))
    a = _nvt(gpu, net, hyp=hyp)
    a.run(20)
    b = _nvt(gpu, net, component=True, hyp=hyp)
    b.run(20)
    for prop in ("central_pos", "orientations", "velocities", "angmom"):
        u, v = getattr(a, prop).cpu().numpy(), getattr(b, prop).cpu().numpy()
        err = float(np.abs(u - v).max())
        assert err < 1e-10 * max(1.0, np.abs(v).max()), (prop, err)
    ta, tb = a.thermostat(), b.thermostat()
    assert abs(ta["xi"] - tb["xi"]) < 1e-12 and abs(ta["Kt"] - tb["Kt"]) < 1e-9 * ta["Kt"]


def test_sim_gsd_dump_and_exact_resume(gpu, tmp_path):
    from nepx import gsd
    net = _net(gpu)
    a = _nvt(gpu, net)
    a.setGsdDump(10, str(tmp_path / "run.gsd"))
    a.run(20)
    assert gsd.n_frames(str(tmp_path / "run.gsd")) == 2
    # resume from the last dumped frame (step 10) and compare with an uninterrupted run at step 20
    b = _nvt(gpu, net)
    b.setGsdDump(0, None)
    b.run(10)
    fr = gsd.read_frame(str(tmp_path / "run.gsd"))
    assert int(fr["configuration/step"][0]) == 10
    assert np.array_equal(fr["log/nepx/position_f64"], b.central_pos.cpu().numpy()), "dumped state differs"
    th = fr["log/nepx/thermostat"]
    tb = b.thermostat()
    assert th[0] == tb["xi"] and th[2] == tb["xi_r"]
    from nepx.NVT import NVT
    c = NVT()
    c.setDevice(0)
    c.initializeSystem(str(tmp_path / "run.gsd"))
    c.setNepDescriptors([10, 8, 4, 4.5])
    c.setEvaluator(net=net)
    c.setShape("th_v1")
    c.setNeighborList()
    c.setkT(0.5)
    c.setdt(0.005)
    c.run(10)
    err = float(np.abs(c.central_pos.cpu().numpy() - a.central_pos.cpu().numpy()).max())
    assert err < 1e-9, err


def test_sim_dump_labels_follow_device_steps_after_overflow(gpu, tmp_path):
    """A run that starts with an overflowed list (its steps are skipped until the host grows the
    list): every frame is labelled with the device's step count, labels are strictly increasing, and
    each frame holds the state an unconstrained run has at that step."""
    from nepx import gsd
    net = _net(gpu)
    ref = _nvt(gpu, net)
    ref.setGsdDump(5, str(tmp_path / "ref.gsd"))
    ref.run(30)
    a = _nvt(gpu, net)
    a.setGsdDump(5, str(tmp_path / "ovf.gsd"))
    a.move_data_to_device()
    a.arena._alloc_list(64)                # far below the ~570 pairs in range: the forced build overflows
    a.arena.build_list(True)
    a.runner._capture()
    assert a.arena.flags()[1] == 1
    a.run(30)
    n = gsd.n_frames(str(tmp_path / "ovf.gsd"))
    assert n == gsd.n_frames(str(tmp_path / "ref.gsd")) == 6
    steps = []
    for k in range(n):
        fa = gsd.read_frame(str(tmp_path / "ovf.gsd"), k)
        fr = gsd.read_frame(str(tmp_path / "ref.gsd"), k)
        steps.append(int(fa["configuration/step"][0]))
        assert steps[-1] == int(fr["configuration/step"][0])
        assert np.array_equal(fa["log/nepx/position_f64"], fr["log/nepx/position_f64"]), steps[-1]
    assert steps == [0, 5, 10, 15, 20, 25]
    assert a.arena.device_steps() == 30
