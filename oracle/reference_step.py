"""The reference algorithm of one MD step in FP64 numpy -- TEST AND CPU-BASELINE INFRASTRUCTURE ONLY.

Used by bench.py's cpu_baseline leg ("port": the reference's own algorithm,
restated) and by tests.  Follows nep_simulator/Sim.py:257-285:
  pairs within the COM cutoff (the KD-tree list of Sim.py:85-92 filtered by
  DescriptorGenerator.py:168-175), the 10-row finite-difference stencil
  (DescriptorGenerator.py:187-255), NEP descriptors of every stencil row,
  the FFNN, forward-difference forces/torques with the +-100 clip and the
  scatter (ForceEvaluator.py:76-146), then integrate_step_two/one (NVT.py:17-129).
"""
import numpy as np

from . import nep_oracle as O


def reference_forces(X, Q, shape, L, hyp, layers, act="tanh", en_min=0.0, en_max=1.0, dx=1e-5, dteta=1e-5,
                     chunk=2000, rc=None):
    rc = shape["cutoff"] if rc is None else rc
    pairs = O.brute_pairs(X, L, rc)
    N = len(X)
    Ns = len(shape["sites"])
    Fi = np.zeros((len(pairs), 3))
    Ti = np.zeros((len(pairs), 3))
    Tj = np.zeros((len(pairs), 3))
    for a in range(0, len(pairs), chunk):
        pc = pairs[a:a + chunk]
        clouds, t = O.pair_clouds(X, Q, pc, shape["sites"], L)
        f, ti, tj = O.fd_pair_forces(clouds, Ns, t, shape["nep_cutoff"], hyp, layers, act, en_min, en_max, dx, dteta)
        Fi[a:a + chunk], Ti[a:a + chunk], Tj[a:a + chunk] = f, ti, tj
    F, T = O.scatter_forces(N, pairs, Fi, Ti, Tj)
    return F, T, pairs


def reference_step(state, thermo, shape, L, hyp, layers, dt, kT, tau, moi, ndof, rndof, act="tanh"):
    """One reference-style step; returns (new_state, new_thermo)."""
    F, T, _ = reference_forces(state["x"], state["q"], shape, L, hyp, layers, act)
    new, Kt, Kr = O.nvt_step(state, F, T, dt, kT, tau, L, moi, thermo)
    th = O.thermo_update(thermo, Kt, Kr, ndof, rndof, dt, tau, kT)
    return new, th
