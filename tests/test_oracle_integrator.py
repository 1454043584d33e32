"""Pins the oracle's rigid-body NVT integrator with non-zero forces, torques and thermostat coupling.

Functions pinned: nvt_step_two / nvt_step_one (IntegrateStepTwoKernels.py:22-66, NVT.py:36-64,
IntegrateStepOneKernels.py:3-251, NVT.py:67-129), torque_dp (IntegrateStepTwoKernels.py:22-58)
and thermo_update (IntegrateStepTwoKernels.py:74-125).  Each check compares with physics that
the update formulas do not restate:

* constant force: the positions follow uniform acceleration exactly (x = x0 + u t + a t^2/2);
* constant lab-frame torque: the lab angular momentum grows by exactly dt*tau per step, for an
  isotropic and an anisotropic body (torque-free NO_SQUISH rotation conserves it);
* constant friction and force: the steady-state velocity and angular momentum reproduce the
  continuous balance a/xi to within the error x^2/24 (x = xi dt) of a friction-half / kick /
  friction-half splitting (the reverse order, kick-half / friction / kick-half, errs by x^2/12);
* NVE with the oracle's NEP pair energy (itself pinned by finite differences): the energy error
  is second order in dt;
* NVT: the Nose-Hoover extended energy K + U + N_dof kT (tau^2 xi^2/2 + eta) + RN_dof kT
  (tau^2 xi_r^2/2 + eta_r) is conserved to second order in dt under strong coupling.

The last part plants plausible mistakes in the oracle (on the CPU) and shows that each one fails
at least one of these checks.
"""
import math

import numpy as np
import pytest

from oracle import nep_oracle as O

FREE = dict(xi=0.0, eta=0.0, xi_r=0.0, eta_r=0.0, theta=1.0, e_r=1.0)


def _unit_q(rng, n):
    q = rng.standard_normal((n, 4))
    return q / np.linalg.norm(q, axis=1, keepdims=True)


def _lab_angmom(q, p):
    """R(q) s with s = q* (x) p / 2, the body-frame angular momentum."""
    s = 0.5 * O.quat_mul(O.quat_conj(q), p)[:, 1:]
    return O.quat_rotate(q, s)


def _step(st, F, T, dt, L, moi, th):
    mid, dp = O.nvt_step_two(st, F, T, dt, th)
    return O.nvt_step_one(mid, F, dp, dt, L, moi, th)


# ------------------------------------------------------------ closed forms
def check_constant_force():
    rng = np.random.default_rng(30)
    n, dt, steps = 6, 0.01, 40
    m = rng.uniform(0.5, 2.0, n)
    F = rng.standard_normal((n, 3))
    q = _unit_q(rng, n)
    st = dict(x=rng.uniform(-1, 1, (n, 3)), v=rng.standard_normal((n, 3)), q=q, p=np.zeros((n, 4)), m=m)
    x0, u0 = st["x"].copy(), st["v"].copy()
    a = F / m[:, None]
    for _ in range(steps):
        st, _, _ = _step(st, F, np.zeros((n, 3)), dt, 1e3, (0.2, 0.3, 0.4), FREE)
    t = steps * dt
    # the stored velocity trails the positions by half a step: at t = 0 the on-step velocity is u0 + a dt/2
    want = x0 + (u0 + 0.5 * dt * a) * t + 0.5 * a * t * t
    assert np.abs(st["x"] - want).max() < 1e-12, np.abs(st["x"] - want).max()
    assert np.abs(st["v"] - (u0 + a * t)).max() < 1e-12


def check_constant_torque(moi):
    rng = np.random.default_rng(31)
    n, dt, steps = 5, 0.004, 60
    q = _unit_q(rng, n)
    Lb = 0.3 * rng.standard_normal((n, 3))
    p = 2.0 * O.quat_mul(q, np.concatenate([np.zeros((n, 1)), Lb], 1))
    tau = rng.standard_normal((n, 3))
    st = dict(x=np.zeros((n, 3)), v=np.zeros((n, 3)), q=q, p=p, m=np.ones(n))
    L0 = _lab_angmom(q, p)
    for _ in range(steps):
        st, _, _ = _step(st, np.zeros((n, 3)), tau, dt, 1e3, moi, FREE)
    err = np.abs(_lab_angmom(st["q"], st["p"]) - (L0 + steps * dt * tau)).max()
    assert err < 1e-12, err
    # the body did turn: the torque-free parts moved q
    assert np.abs(np.abs(np.sum(st["q"] * q, 1)) - 1.0).max() > 1e-4


def check_friction_kick_splitting():
    """Constant friction xi (theta = e_r = exp(-xi dt/2)) and constant force / torque drive v and
    the lab angular momentum to a steady state; the continuous one is F/(m xi) and tau/xi."""
    x = 0.5                                   # xi * dt
    dt = 0.01
    xi = x / dt
    th = dict(FREE, theta=math.exp(-0.5 * x), e_r=math.exp(-0.5 * x))
    rng = np.random.default_rng(32)
    n = 4
    m = rng.uniform(0.5, 2.0, n)
    F = rng.standard_normal((n, 3))
    tau = rng.standard_normal((n, 3))
    q = _unit_q(rng, n)
    st = dict(x=np.zeros((n, 3)), v=np.zeros((n, 3)), q=q, p=np.zeros((n, 4)), m=m)
    I = 0.35
    for _ in range(150):                      # theta^(2*150) = e^-75: converged
        st, _, _ = _step(st, F, tau, dt, 1e3, (I, I, I), th)
    rv = st["v"] / (F / (m[:, None] * xi))
    rl = _lab_angmom(st["q"], st["p"]) / (tau / xi)
    for r in (rv, rl):
        assert np.abs(r - r.mean()).max() < 1e-12          # one ratio for every component
        dev = abs(float(r.mean()) - 1.0)
        assert dev <= x * x / 24.0, (float(r.mean()), x * x / 24.0)


# ------------------------------------------------------- energy conservation
SH = O.shape_table()["tet10"]
HYP = (11, 6, 4)
BOX, RC = 60.0, 25.0                           # every pair stays inside the cutoff: U is smooth
MOI = (0.08, 0.11, 0.15)


def _cluster():
    rng = np.random.default_rng(5)
    X = np.array([[0.0, 0, 0], [1.6, 0.3, -0.2], [0.2, 1.5, 0.4]])
    Q = _unit_q(rng, 3)
    V = 0.5 * rng.standard_normal((3, 3))
    Lb = 0.3 * rng.standard_normal((3, 3))
    P = 2.0 * O.quat_mul(Q, np.concatenate([np.zeros((3, 1)), Lb], 1))
    return dict(x=X, v=V, q=Q, p=P, m=np.ones(3))


def energy_error(dt, thermostat, t_end=0.8, tau=0.05, kT=1.0):
    """Largest deviation of the (extended) energy from its initial value over t_end, measured at
    the on-step state between step two and step one."""
    layers = O.random_layers(40, [64, 64], seed=3, std_scale=1.0)
    st = _cluster()
    n = len(st["x"])
    ndof, rndof = 3.0 * (n - 1), 3.0 * n

    def forces(s):
        F, T, _, r = O.system_forces(s["x"], s["q"], SH, BOX, HYP, layers, rc=RC)
        return F, T, float(np.sum(r["E"]))

    th = dict(FREE)
    F, T, U = forces(st)
    Kt = 0.5 * float(np.sum(st["m"][:, None] * st["v"] ** 2))
    Kr = float(np.sum(O.rot_ke(st["q"], st["p"], MOI)))
    H = []
    for _ in range(int(round(t_end / dt))):
        if thermostat:
            th = O.thermo_update(th, Kt, Kr, ndof, rndof, dt, tau, kT)
        mid, dp = O.nvt_step_two(st, F, T, dt, th)
        h = 0.5 * float(np.sum(mid["m"][:, None] * mid["v"] ** 2)) + float(np.sum(O.rot_ke(mid["q"], mid["p"], MOI))) + U
        if thermostat:
            h += ndof * kT * (0.5 * tau * tau * th["xi"] ** 2 + th["eta"])
            h += rndof * kT * (0.5 * tau * tau * th["xi_r"] ** 2 + th["eta_r"])
        H.append(h)
        st, Kt, Kr = O.nvt_step_one(mid, F, dp, dt, BOX, MOI, th)
        F, T, U = forces(st)
    H = np.asarray(H)
    return float(np.abs(H - H[0]).max()), st


def check_energy_second_order(thermostat):
    e1, s1 = energy_error(0.004, thermostat)
    e2, _ = energy_error(0.002, thermostat)
    # the run must move: kinetic energy exchanged with U (and with the thermostat)
    assert np.abs(s1["x"] - _cluster()["x"]).max() > 0.05
    assert e1 < 2e-2 and 3.3 < e1 / e2 < 4.8, (e1, e2, e1 / e2)


def test_constant_force_is_uniform_acceleration():
    check_constant_force()


@pytest.mark.parametrize("moi", [(0.4, 0.4, 0.4), (0.1, 0.25, 0.6)])
def test_constant_torque_adds_lab_angular_momentum(moi):
    check_constant_torque(moi)


def test_friction_kick_friction_steady_state():
    check_friction_kick_splitting()


def test_nve_energy_error_second_order():
    check_energy_second_order(False)


def test_nvt_extended_energy_second_order():
    check_energy_second_order(True)


def test_nvt_thermostat_heats_a_cold_cluster():
    """Under strong coupling a cluster started below kT gains kinetic energy: xi turns negative."""
    _, st = energy_error(0.004, True, t_end=0.8)
    K = 0.5 * float(np.sum(st["m"][:, None] * st["v"] ** 2)) + float(np.sum(O.rot_ke(st["q"], st["p"], MOI)))
    c = _cluster()
    K0 = 0.5 * float(np.sum(c["m"][:, None] * c["v"] ** 2)) + float(np.sum(O.rot_ke(c["q"], c["p"], MOI)))
    assert K0 < 0.5 * 15.0 and K > 2.0 * K0


# ------------------------------------------------------------ planted mistakes
def _quat0(v):
    return np.concatenate([np.zeros(v.shape[:-1] + (1,)), v], axis=-1)


def _dp_R_q(q, tau, dt):                     # R(q) in place of R(q*)
    return dt * O.quat_mul(q, _quat0(O.quat_rotate(q, tau)))


def _dp_tau_q(q, tau, dt):                   # (0, tau_b) (x) q in place of q (x) (0, tau_b)
    return dt * O.quat_mul(_quat0(O.quat_rotate(O.quat_conj(q), tau)), q)


def _thermo_no_half(th, Kt, Kr, N_dof, RN_dof, dt, tau, kT):   # coupling dt/tau^2 instead of dt/(2 tau^2)
    c = dt / tau / tau
    out = dict(th)
    for xi, eta, T in (("xi", "eta", 2.0 * Kt / N_dof), ("xi_r", "eta_r", 2.0 * Kr / RN_dof)):
        mid = th[xi] + c * (T / kT - 1.0)
        out[xi] = mid + c * (T / kT - 1.0)
        out[eta] = th[eta] + mid * dt
    out["theta"] = math.exp(-0.5 * out["xi"] * dt)
    out["e_r"] = math.exp(-0.5 * out["xi_r"] * dt)
    return out


def _two_swapped(state, F, T, dt, thermo):   # step two applies (p + dp) e, step one p e + dp
    s = {k: np.array(state[k], dtype=np.float64, copy=True) for k in "xvqpm"}
    s["v"] = (s["v"] + 0.5 * dt * F / s["m"][:, None]) * thermo["theta"]
    dp = O.torque_dp(s["q"], T, dt)
    s["p"] = (s["p"] + dp) * thermo["e_r"]
    return s, dp


def _one_swapped(state, F, dp, dt, L, moi, thermo):
    s = {k: np.array(state[k], dtype=np.float64, copy=True) for k in "xvqpm"}
    s["v"] = s["v"] * thermo["theta"] + 0.5 * dt * F / s["m"][:, None]
    s["x"] = s["x"] + s["v"] * dt
    s["p"] = s["p"] * thermo["e_r"] + dp
    s["p"], s["q"] = O.no_squish(s["p"], s["q"], moi, dt)
    s["q"] /= np.linalg.norm(s["q"], axis=-1, keepdims=True)
    Kt = float(np.sum(0.5 * s["m"] * np.sum(s["v"] ** 2, 1)))
    return s, Kt, float(np.sum(O.rot_ke(s["q"], s["p"], moi)))


def _two_full_kick(state, F, T, dt, thermo):  # a whole dt kick in step two
    s, dp = O.__dict__["_orig_two"](state, F, T, dt, thermo)
    s["v"] = s["v"] + 0.5 * dt * F / s["m"][:, None]
    return s, dp


MUTANTS = {
    "torque_R_q": (dict(torque_dp=_dp_R_q), lambda: check_constant_torque((0.1, 0.25, 0.6))),
    "torque_tau_q": (dict(torque_dp=_dp_tau_q), lambda: check_constant_torque((0.1, 0.25, 0.6))),
    "coupling_no_half": (dict(thermo_update=_thermo_no_half), lambda: check_energy_second_order(True)),
    "step_order_swapped": (dict(nvt_step_two=_two_swapped, nvt_step_one=_one_swapped), check_friction_kick_splitting),
    "full_kick": (dict(nvt_step_two=_two_full_kick), check_constant_force),
}


@pytest.mark.parametrize("name", sorted(MUTANTS))
def test_planted_mistake_fails_its_check(name, monkeypatch):
    patch, check = MUTANTS[name]
    monkeypatch.setitem(O.__dict__, "_orig_two", O.nvt_step_two)
    for k, v in patch.items():
        monkeypatch.setattr(O, k, v)
    with pytest.raises(AssertionError):
        check()
