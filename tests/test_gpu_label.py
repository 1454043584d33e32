"""Label potential (soft bead-bead repulsion of 01/SamplerUtils.py:555-631) on the GPU against the FP64
oracle: unreduced-table energies, and forces/torques of a periodic system stepped through the shared
gather and integrator.  Tolerances: energies 1e-5, forces and torques 1e-4 relative (max-norm)."""
import numpy as np
import pytest
import torch

from oracle import nep_oracle as O

pytestmark = pytest.mark.gpu


def test_calc_en_unred_matches_oracle(gpu):
    from nepx import label
    rng = np.random.default_rng(6)
    lo = np.array([-3.142, 0.0, -3.142, 4.3, -3.142, 0.0])     # the unreduced table ranges (get_LHC_simple)
    hi = np.array([3.142, 3.142, 3.142, 7.2, 3.142, 3.142])
    u = lo + rng.random((400, 6)) * (hi - lo)
    E = label.calc_en_unred(u, device=gpu)
    E64 = O.calc_en_unred(u, O.cube_structure_beads())
    e = np.abs(E - E64).max() / np.abs(E64).max()
    print(f"unred energies: {np.count_nonzero(E64)} nonzero of {len(u)}, max rel {e:.2e}")
    assert np.count_nonzero(E64) > 100 and e < 1e-5


def test_bead_potential_system_forces(gpu):
    import nepx
    from nepx import label
    from nepx.configs import make_system
    from nepx.engine import StepRunner
    pot = label.BeadPotential(device=gpu)
    spec = make_system("cube8", 64, 0.018, seed=12, rc=pot.reach, non_overlap=False)
    ar = nepx.Arena([spec], nepx.get_shape("cube8"), (11, 6, 4), None, device=gpu, clip=0.0, potential=pot)
    ar.init()
    ar.filter_pairs()
    ar.pair_fused()
    ar.gather()
    torch.cuda.synchronize()
    got = ar.system_arrays(0)
    q32 = spec.q.astype(np.float32).astype(np.float64)
    pairs = O.brute_pairs(spec.x, spec.L, pot.reach)
    r = O.bead_pair_forces(spec.x, q32, pairs, O.cube_structure_beads(), spec.L)
    F, T = O.scatter_forces(spec.n, pairs, r["F"], r["Ti"], r["Tj"])
    eF = np.abs(got["F"] - F).max() / np.abs(F).max()
    eT = np.abs(got["T"] - T).max() / np.abs(T).max()
    eU = abs(got["upot"].sum() - r["E"].sum()) / r["E"].sum()
    print(f"bead system: {len(pairs)} pairs, {np.count_nonzero(r['E'])} touching, F {eF:.2e} T {eT:.2e} U {eU:.2e}")
    assert np.count_nonzero(r["E"]) > 10 and eF < 1e-4 and eT < 1e-4 and eU < 1e-5
    # and it steps: 50 captured steps, state finite, momentum of the pair forces zero
    run = StepRunner(ar, use_graph=True)
    run._capture()
    run.run(50)
    torch.cuda.synchronize()
    ar.check()
    a = ar.system_arrays(0)
    assert np.all(np.isfinite(a["x"])) and np.all(np.isfinite(a["p"]))
    assert np.abs(a["F"].sum(0)).max() < 1e-9 * max(1.0, np.abs(a["F"]).sum())


def _bead_extended_energy_error(gpu, dt, t_end=0.4, tau=0.1, kT=1.0):
    """Largest deviation of the Nose-Hoover extended energy K + U + N_dof kT (tau^2 xi^2/2 + eta)
    + RN_dof kT (tau^2 xi_r^2/2 + eta_r) from its first value, for 64 cubes under the smooth bead
    potential (every interacting pair is inside the COM cutoff), measured between step two and step
    one (the integrator run part by part; the kinetic energies are taken at the on-step velocities
    with the measurement-only nepx_nvt_init)."""
    import nepx
    from nepx import _lib, label
    from nepx.configs import make_system
    pot = label.BeadPotential(device=gpu)
    spec = make_system("cube8", 64, 0.004, kT=kT, dt=dt, seed=12, rc=pot.reach, non_overlap=False)
    spec.tau = tau
    ar = nepx.Arena([spec], nepx.get_shape("cube8"), (11, 6, 4), None, device=gpu, clip=0.0, potential=pot)
    ar.init()
    H, U = [], []
    for _ in range(int(round(t_end / dt))):
        ar.build_list(False)
        ar.filter_pairs()
        ar.pair_fused()
        ar.gather()
        ar.nvt_part(_lib.NVT_STEP_TWO)
        ar.nvt_init(update_thermostat=False)
        th = ar.thermo()[0]
        H.append(th["Kt"] + th["Kr"] + th["U"] + spec.ndof * kT * (0.5 * tau * tau * th["xi"] ** 2 + th["eta"]) +
                 spec.rndof * kT * (0.5 * tau * tau * th["xi_r"] ** 2 + th["eta_r"]))
        U.append(th["U"])
        ar.nvt_part(_lib.NVT_TRANSLATE)
        ar.nvt_part(_lib.NVT_ROTATE)
        ar.nvt_part(_lib.NVT_THERMO)
    torch.cuda.synchronize()
    ar.check()
    H = np.asarray(H)
    return float(np.abs(H - H[0]).max()), float(H[0]), float(min(U)), float(max(U))


def test_bead_nvt_extended_energy_is_second_order(gpu):
    """The GPU list, bead pair kernel, gather and FP64 integrator over 100 and 200 steps: the
    extended-energy error falls fourfold when dt halves (the FP64 oracle on the same system gives
    1.14 and 0.27, ratio 4.19), and the run exchanges energy (U rises from about 1 to hundreds)."""
    e1, h0, umin, umax = _bead_extended_energy_error(gpu, 0.004)
    e2, _, _, _ = _bead_extended_energy_error(gpu, 0.002)
    print(f"H0 {h0:.3f}  U {umin:.2f}..{umax:.2f}  max |H - H0|: dt 0.004 {e1:.4f}, dt 0.002 {e2:.4f}, "
          f"ratio {e1 / e2:.2f}")
    assert umax - umin > 100.0
    assert e1 < 3e-3 * h0 and 3.3 < e1 / e2 < 5.2


def test_bead_nvt_temperatures_reach_kT(gpu):
    """512 cubes under the smooth bead potential, started at kT 1.5 and coupled to kT 0.8
    (tau 0.1): over the second half of 3,200 captured steps the translational and rotational
    temperatures 2 K_t / N_dof and 2 K_r / RN_dof average kT within 5 %."""
    import nepx
    from nepx import label
    from nepx.configs import make_system
    from nepx.engine import StepRunner
    pot = label.BeadPotential(device=gpu)
    spec = make_system("cube8", 512, 0.004, kT=1.5, dt=0.002, seed=21, rc=pot.reach, non_overlap=False)
    spec.kT, spec.tau = 0.8, 0.1
    ar = nepx.Arena([spec], nepx.get_shape("cube8"), (11, 6, 4), None, device=gpu, clip=0.0, potential=pot)
    r = StepRunner(ar, use_graph=True)
    r.init()
    Tt, Tr = [], []
    th = ar.thermo()[0]
    t0 = (2 * th["Kt"] / spec.ndof, 2 * th["Kr"] / spec.rndof)
    for _ in range(200):
        r.run_exact(16)
        th = ar.thermo()[0]
        Tt.append(2 * th["Kt"] / spec.ndof)
        Tr.append(2 * th["Kr"] / spec.rndof)
    ar.check()
    mt, mr = float(np.mean(Tt[100:])) / spec.kT, float(np.mean(Tr[100:])) / spec.kT
    print(f"bead NVT: T_t {t0[0]:.3f} -> <T_t>/kT {mt:.4f}, T_r {t0[1]:.3f} -> <T_r>/kT {mr:.4f} "
          f"(std {np.std(Tt[100:]) / spec.kT:.3f}, {np.std(Tr[100:]) / spec.kT:.3f})")
    assert t0[0] > 1.2 and abs(mt - 1.0) < 0.05 and abs(mr - 1.0) < 0.05
