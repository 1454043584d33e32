"""Neighbour list, per-step filter, gather and the FP64 integrator against the oracle.

Neighbour sets are compared bit-exactly (sorted int pairs).  Forces/torques
within 1e-4 relative (max-norm over the system) of FP64; one integrator step
within 1e-12 relative of the FP64 oracle given the same forces.
"""
import numpy as np
import pytest
import torch

from oracle import nep_oracle as O

pytestmark = pytest.mark.gpu


def _arena(gpu, shape_name, n, phi, seed=0, skin=0.5, widths=(64, 64), hyp=(11, 6, 4), precision="auto"):
    import nepx
    from nepx.configs import make_system
    spec = make_system(shape_name, n, phi, seed=seed, skin=skin)
    D = nepx.n_descriptor(*hyp)
    net = nepx.EnergyNet.random(D, list(widths), seed=seed + 100, device=gpu, precision=precision)
    ar = nepx.Arena([spec], nepx.get_shape(shape_name), hyp, net, device=gpu)
    return ar, spec, net


@pytest.mark.parametrize("shape_name,n,phi", [("tet10", 1000, 0.45), ("tet10", 200, 0.1), ("tet10", 4000, 0.45),
                                              ("cube8", 512, 0.3)])
def test_pair_set_bit_exact(gpu, shape_name, n, phi):
    ar, spec, _ = _arena(gpu, shape_name, n, phi)
    ar.init()
    # list within r_list
    f = ar.flags()
    PL = int(f[2])
    rows = ar.lt["slot_row"][:PL].cpu().numpy()
    cols = ar.lt["nbr"][:PL].cpu().numpy()
    got_list = np.stack([rows, cols], 1)
    want_list = O.brute_pairs(spec.x, spec.L, spec.rc + spec.skin)
    assert np.array_equal(got_list, want_list), "Verlet list differs from brute force"
    # per-step filter within rc
    ar.filter_pairs()
    got = ar.pairs().cpu().numpy().astype(np.int64)
    want = O.brute_pairs(spec.x, spec.L, spec.rc)
    print(f"{shape_name} n={n}: list {PL} pairs, active {len(got)} (oracle {len(want)})")
    assert np.array_equal(got, want)


@pytest.mark.parametrize("shape_name,n,hyp,widths,precision",
                         [("tet10", 400, (11, 6, 4), (64, 64), "f32"), ("tet10", 400, (11, 6, 4), (64, 64), "f16x3"),
                          ("cyl12", 200, (17, 6, 6), (128, 128), "f16x3")])
def test_system_forces_match_oracle(gpu, shape_name, n, hyp, widths, precision):
    ar, spec, net = _arena(gpu, shape_name, n, 0.1, seed=4, widths=widths, hyp=hyp, precision=precision)
    ar.init()
    ar.filter_pairs()
    ar.pair_fused()
    ar.gather()
    torch.cuda.synchronize()
    got = ar.system_arrays(0)
    q32 = spec.q.astype(np.float32).astype(np.float64)
    F, T, pairs, r = O.system_forces(spec.x, q32, O.shape_table()[shape_name], spec.L, hyp, net.layers_f64())
    eF = np.max(np.abs(got["F"] - F)) / np.max(np.abs(F))
    eT = np.max(np.abs(got["T"] - T)) / np.max(np.abs(T))
    eU = abs(got["upot"].sum() - r["E"].sum()) / np.sum(np.abs(r["E"]))
    print(f"system forces {shape_name} ({precision}): {len(pairs)} pairs, F {eF:.2e} T {eT:.2e} U {eU:.2e}")
    assert eF < 1e-4 and eT < 1e-4 and eU < 1e-5
    # momentum: Newton's third law holds pair by pair
    assert np.abs(got["F"].sum(0)).max() < 1e-9 * max(1.0, np.abs(got["F"]).sum())


def test_nvt_step_matches_oracle(gpu):
    ar, spec, net = _arena(gpu, "tet10", 300, 0.1, seed=6)
    ar.init()
    th0 = ar.thermo()[0]
    s0 = ar.system_arrays(0)
    ar.step()
    torch.cuda.synchronize()
    s1 = ar.system_arrays(0)
    th1 = ar.thermo()[0]
    # oracle: thermostat from the initial kinetic energies, then one step with the GPU forces
    Kt0 = 0.5 * np.sum(s0["mass"][:, None] * s0["v"] ** 2)
    Kr0 = float(np.sum(O.rot_ke(s0["q"], s0["p"], spec.moi)))
    th = O.thermo_update(dict(xi=-0.0317711, eta=0.108609, xi_r=0.00814227, eta_r=-0.014238), Kt0, Kr0,
                         spec.ndof, spec.rndof, spec.dt, spec.tau, spec.kT)
    assert abs(th0["xi"] - th["xi"]) < 1e-12 and abs(th0["theta"] - th["theta"]) < 1e-12
    new, Kt, Kr = O.nvt_step(dict(x=s0["x"], v=s0["v"], q=s0["q"], p=s0["p"], m=s0["mass"]), s1["F"], s1["T"],
                             spec.dt, spec.kT, spec.tau, spec.L, spec.moi, th)
    for k in "xvqp":
        e = np.max(np.abs(s1[k] - new[k])) / np.max(np.abs(new[k]))
        assert e < 1e-12, (k, e)
    assert abs(th1["Kt"] - Kt) < 1e-10 * Kt and abs(th1["Kr"] - Kr) < 1e-10 * Kr


def test_step_deterministic(gpu):
    outs = []
    for _ in range(2):
        ar, spec, net = _arena(gpu, "tet10", 512, 0.45, seed=8)
        ar.init()
        for _ in range(5):
            ar.step()
        torch.cuda.synchronize()
        outs.append(ar.system_arrays(0))
    for k in "xvqpFT":
        assert np.array_equal(outs[0][k], outs[1][k]), k


def test_tuned_launches_give_same_trajectory(gpu, tmp_path):
    """Skin, gather lanes and fused grid change only FP64 summation order: 30 steps agree to 1e-12."""
    import nepx
    from nepx import tune
    from nepx.configs import make_system
    from nepx.engine import StepRunner
    net = nepx.EnergyNet.random(40, [64, 64], seed=3, device=gpu)
    shape = nepx.get_shape("tet10")
    old = tune.get_launch()
    out = []
    try:
        for t in (tune.Tuning(), tune.Tuning(skin=0.3, gather_lanes=4, fused_blocks_per_cu=2)):
            tune.set_launch(t)
            spec = make_system("tet10", 2000, 0.45, seed=8, skin=t.skin)
            ar = nepx.Arena([spec], shape, (11, 6, 4), net, device=gpu)
            r = StepRunner(ar, use_graph=True)
            r.init()
            r.run(30)
            torch.cuda.synchronize()
            ar.check()
            out.append(ar.system_arrays(0))
        for k in ("x", "q", "v", "p"):
            scale = np.abs(out[0][k]).max()
            assert np.abs(out[0][k] - out[1][k]).max() <= 1e-12 * scale, k
        res = tune.Autotuner(str(tmp_path / "tune.json")).tune(
            [make_system("tet10", 1000, 0.45, seed=9)], shape, (11, 6, 4), net, gpu, steps=10, warmup=3,
            space={"skin": (0.4, 0.5), "gather_lanes": (8, 16), "fused_blocks_per_cu": (1, 2)})
        assert res.best.gather_lanes in (8, 16) and len(res.trials) >= 4
        again = tune.Autotuner(str(tmp_path / "tune.json")).tune(
            [make_system("tet10", 1000, 0.45, seed=9)], shape, (11, 6, 4), net, gpu)
        assert again.cached and again.best == res.best
    finally:
        tune.set_launch(tune.Tuning(**old))


def test_list_overflow_skips_steps_and_recovers(gpu):
    """A list that overflows makes the integrator skip its steps (state frozen); after the host grows
    the list the run continues and ends bitwise where an unconstrained run ends."""
    import nepx
    from nepx.configs import make_system
    from nepx.engine import StepRunner
    net = nepx.EnergyNet.random(40, [64, 64], seed=3, device=gpu)
    shape = nepx.get_shape("tet10")
    ref = nepx.Arena([make_system("tet10", 1500, 0.45, seed=5)], shape, (11, 6, 4), net, device=gpu)
    rr = StepRunner(ref, use_graph=True)
    rr.init()
    rr.run(20)
    torch.cuda.synchronize()
    ref.check()
    ar = nepx.Arena([make_system("tet10", 1500, 0.45, seed=5)], shape, (11, 6, 4), net, device=gpu)
    r = StepRunner(ar, use_graph=True)
    r.init()
    ar._alloc_list(1000)          # far too small: the forced build overflows
    ar.build_list(True)
    r._capture()
    r.run(5)
    torch.cuda.synchronize()
    assert ar.flags()[1] == 1 and ar.device_steps() == 0
    x0 = ar.system_arrays(0)["x"].copy()
    assert r.recover() and ar.cap > 1000
    assert np.array_equal(ar.system_arrays(0)["x"], x0)        # the skipped steps left the state alone
    r.run(20 - ar.device_steps())
    torch.cuda.synchronize()
    ar.check()
    assert ar.device_steps() == 20
    a, b = ref.system_arrays(0), ar.system_arrays(0)
    for k in ("x", "v", "q", "p", "F", "T"):
        assert np.array_equal(a[k], b[k]), k


def test_non_finite_state_stops_only_its_system(gpu):
    """A NaN put into one system's velocities: the device stops that system at the next step (its
    step count stays, check() names it) while the other system of the batch runs on and ends
    bitwise where it ends when run alone."""
    import nepx
    from nepx import _lib
    from nepx.configs import make_system
    from nepx.engine import StepRunner
    net = nepx.EnergyNet.random(40, [64, 64], seed=3, device=gpu)
    shape = nepx.get_shape("tet10")
    mk = lambda: [make_system("tet10", 600, 0.3, seed=41), make_system("tet10", 500, 0.3, kT=1.3, seed=42)]
    alone = nepx.Arena([mk()[1]], shape, (11, 6, 4), net, device=gpu)
    ra = StepRunner(alone, use_graph=True, steps_per_graph=1)
    ra.init()
    ra.run(12)
    torch.cuda.synchronize()
    alone.check()
    ar = nepx.Arena(mk(), shape, (11, 6, 4), net, device=gpu)
    r = StepRunner(ar, use_graph=True)
    r.init()
    r.run(4)
    torch.cuda.synchronize()
    ar.check()
    ar.t["vx"][ar.starts[0] + 17] = float("nan")
    torch.cuda.synchronize()
    r.run(8)
    torch.cuda.synchronize()
    with pytest.raises(_lib.NepxError, match="non-finite"):
        ar.check()
    th = ar.thermo()
    assert ar.faults() == [0] and int(th[0]["steps"]) == 4 and int(th[1]["steps"]) == 12
    assert ar.t["sp"].cpu().numpy().view(_lib.SYSPARAM_DTYPE)[0]["active"] == 0
    a, b = alone.system_arrays(0), ar.system_arrays(1)
    for k in ("x", "v", "q", "p", "F", "T"):
        assert np.array_equal(a[k], b[k]), k


def test_reverse_record_gather_equals_slot_map_gather(gpu):
    """The gather's two ways to the j side of each pair — the fused kernel's reverse-order records,
    read along each particle's reverse row, and the per-pair records reached through the slot ->
    compacted-index map — sum the same values in the same order: forces, torques and energies agree
    bit for bit (tet10 and cube8, list with inactive slots)."""
    import nepx
    from nepx import _lib
    from nepx.configs import make_system
    for shape_name, n, phi in (("tet10", 1500, 0.45), ("cube8", 512, 0.3)):
        spec = make_system(shape_name, n, phi, seed=6)
        net = nepx.EnergyNet.random(40, [64, 64], seed=106, device=gpu)
        ar = nepx.Arena([spec], nepx.get_shape(shape_name), (11, 6, 4), net, device=gpu, reverse_records=True)
        ar.init()
        ar.filter_pairs()
        ar.pair_fused()
        ar.gather()
        torch.cuda.synchronize()
        f = ar.flags()
        assert f[_lib.FLAG_TAG] > 0 and f[_lib.FLAG_RTAG] == f[_lib.FLAG_TAG]
        assert f[_lib.FLAG_PL] > ar.n_pairs()          # some list slots are outside the cutoff this step
        a = ar.system_arrays(0)
        ar.lt["flags"][_lib.FLAG_RTAG:_lib.FLAG_RTAG + 1].fill_(-1)   # no reverse records this pass
        ar.gather()
        torch.cuda.synchronize()
        b = ar.system_arrays(0)
        for k in ("F", "T", "upot"):
            assert np.array_equal(a[k], b[k]), (shape_name, k)


def test_md_step_c_entry_equals_arena_steps(gpu):
    """One C call per step (nepx_md_step: list build, filter, fused pair kernel, gather, integrator)
    gives the same trajectory bit for bit as the Arena's component calls."""
    import ctypes as C
    from nepx import _lib
    a, _, _ = _arena(gpu, "tet10", 700, 0.35, seed=9)
    b, _, _ = _arena(gpu, "tet10", 700, 0.35, seed=9)
    a.init()
    b.init()
    b.prepare_pair()
    stream = torch.cuda.current_stream()
    for _ in range(10):
        a.step()
        _lib.check(_lib.lib.nepx_md_step(C.byref(b.st), C.byref(b.nl), C.byref(b.model), C.byref(b.cnet),
                                         C.c_float(b.clip), _lib.ptr(b.pair_out), _lib.stream_handle(stream)),
                   "nepx_md_step")
    torch.cuda.synchronize()
    a.check()
    b.check()
    x, y = a.system_arrays(0), b.system_arrays(0)
    for k in ("x", "v", "q", "p", "F", "T"):
        assert np.array_equal(x[k], y[k]), k
    assert a.device_steps() == b.device_steps() == 10
