"""The reference algorithm of one MD step in FP64 numpy -- TEST AND CPU-BASELINE INFRASTRUCTURE ONLY.

Used by bench.py's cpu_baseline leg ("port": the reference's own algorithm,
restated) and by tests.  Follows nep_simulator/Sim.py:207-285:
  a KD-tree half list within cutoff + 2.0 rebuilt every 100 steps
  (Sim.py:85-92,192-198, scipy cKDTree as the reference uses it), pairs within
  the COM cutoff (DescriptorGenerator.py:168-175), the 10-row
  finite-difference stencil (DescriptorGenerator.py:187-255), NEP descriptors
  of every stencil row, the FFNN, forward-difference forces/torques with the
  +-100 clip and the scatter (ForceEvaluator.py:76-146), then
  integrate_step_two/one (NVT.py:17-129).
The pair work of a step can be spread over worker processes (pair chunks, one
BLAS thread each); results do not depend on the split.
"""
import numpy as np

from . import nep_oracle as O

LIST_SKIN = 2.0        # Sim.py:86: Nlist_cutoff = cutoff + 2.0
LIST_EVERY = 100       # Sim.py:87: Nlist_freq


def kdtree_list(X, L, r):
    """Half list i<j within r (Sim.py:192-198: cKDTree over positions shifted by L/2, boxsize L+1e-4)."""
    from scipy.spatial import cKDTree
    tree = cKDTree(np.asarray(X) + 0.5 * L, leafsize=12, boxsize=L + 0.0001)
    return tree.query_pairs(r=r, output_type="ndarray")


def com_filter(X, L, nlist, rc):
    """Candidate pairs whose minimum-image COM distance is below rc (DescriptorGenerator.py:168-175)."""
    if len(nlist) == 0:
        return nlist
    t = O.min_image(X[nlist[:, 1]] - X[nlist[:, 0]], L)
    return nlist[np.sum(t * t, axis=1) < rc * rc]


def fd_chunk(args):
    """Forward-difference forces of one pair chunk: (Fi, Ti, Tj), each (P, 3)."""
    X, Q, pc, shape, L, hyp, layers, act, en_min, en_max, dx, dteta = args
    from threadpoolctl import threadpool_limits
    with threadpool_limits(limits=1):
        clouds, t = O.pair_clouds(X, Q, pc, shape["sites"], L)
        return O.fd_pair_forces(clouds, len(shape["sites"]), t, shape["nep_cutoff"], hyp, layers, act, en_min,
                                en_max, dx, dteta)


def _chunks(pairs, n):
    return [pairs[k:k + n] for k in range(0, len(pairs), n)]


def reference_forces(X, Q, shape, L, hyp, layers, act="tanh", en_min=0.0, en_max=1.0, dx=1e-5, dteta=1e-5,
                     chunk=2000, rc=None, pairs=None, pool=None):
    rc = shape["cutoff"] if rc is None else rc
    pairs = O.brute_pairs(X, L, rc) if pairs is None else pairs
    N = len(X)
    jobs = [(X, Q, pc, shape, L, hyp, layers, act, en_min, en_max, dx, dteta) for pc in _chunks(pairs, chunk)]
    # a worker that stops answering fails the map instead of holding it up
    parts = list(pool.map(fd_chunk, jobs, timeout=600)) if pool is not None else [fd_chunk(j) for j in jobs]
    if parts:
        Fi, Ti, Tj = (np.concatenate([p[k] for p in parts]) for k in range(3))
    else:
        Fi = Ti = Tj = np.zeros((0, 3))
    F, T = O.scatter_forces(N, pairs, Fi, Ti, Tj)
    return F, T, pairs


def reference_step(state, thermo, shape, L, hyp, layers, dt, kT, tau, moi, ndof, rndof, act="tanh"):
    """One reference-style step; returns (new_state, new_thermo)."""
    F, T, _ = reference_forces(state["x"], state["q"], shape, L, hyp, layers, act)
    new, Kt, Kr = O.nvt_step(state, F, T, dt, kT, tau, L, moi, thermo)
    th = O.thermo_update(thermo, Kt, Kr, ndof, rndof, dt, tau, kT)
    return new, th


def reference_run(state, thermo, shape, L, hyp, layers, dt, kT, tau, moi, ndof, rndof, steps, act="tanh",
                  pool=None, chunk=2000):
    """`steps` steps of Sim.run (Sim.py:207-255): KD-tree list every LIST_EVERY steps, per-step COM filter,
    FD stencil forces (pair chunks over `pool`), NVT.  Returns (state, thermo, timings)."""
    import time
    t_list = t_force = t_int = 0.0
    nlist = None
    for k in range(steps):
        t0 = time.perf_counter()
        if k % LIST_EVERY == 0:
            nlist = kdtree_list(state["x"], L, shape["cutoff"] + LIST_SKIN)
        t1 = time.perf_counter()
        pairs = com_filter(state["x"], L, nlist, shape["cutoff"])
        F, T, _ = reference_forces(state["x"], state["q"], shape, L, hyp, layers, act, pairs=pairs, pool=pool,
                                   chunk=chunk)
        t2 = time.perf_counter()
        state, Kt, Kr = O.nvt_step(state, F, T, dt, kT, tau, L, moi, thermo)
        thermo = O.thermo_update(thermo, Kt, Kr, ndof, rndof, dt, tau, kT)
        t3 = time.perf_counter()
        t_list += t1 - t0
        t_force += t2 - t1
        t_int += t3 - t2
    return state, thermo, dict(list_s=t_list, force_s=t_force, integrate_s=t_int)
