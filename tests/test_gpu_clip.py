"""The +-100 per-component clip of the pair forces and torques (ForceEvaluator.py:117-132).

The reference scales the per-pair gradients by the energy range and then clamps every component
with masked assignments (f[f > 100] = 100; f[f < -100] = -100), which leaves NaN components NaN.
These tests choose an energy range that pushes a good share of the components past 100, so the
clip really binds, and compare with the FP64 oracle clamped the same way; tolerance 1e-4 of the
largest component (the pair-kernel budget against FP64).
"""
import ctypes as C

import numpy as np
import pytest
import torch

from oracle import nep_oracle as O
from pairs_util import normwise_rel, oracle_clouds, random_pairs

pytestmark = pytest.mark.gpu

CLIP = 100.0


def _masked_clip(a, c=CLIP):
    """ForceEvaluator.py:125-131 on a numpy copy."""
    a = np.array(a, dtype=np.float64, copy=True)
    a[a > c] = c
    a[a < -c] = -c
    return a


@pytest.fixture(params=[4, 1], ids=["lanes4", "lanes1"])
def pair_lanes(request):
    from nepx import tune
    old = tune.get_launch()
    tune.set_launch(tune.Tuning(**{**old, "pair_lanes": request.param}))
    yield request.param
    tune.set_launch(tune.Tuning(**old))


def test_clip_binds_in_fused_and_split_kernels(gpu, pair_lanes):
    import nepx
    from nepx import ops
    shape = nepx.get_shape("tet10")
    hyp = (11, 6, 4)
    rng = np.random.default_rng(41)
    q, i, j, t = random_pairs(rng, 64, 800, 0.9, 2.4)
    T = dict(qf=torch.from_numpy(q).to(gpu), pi=torch.from_numpy(i).to(gpu), pj=torch.from_numpy(j).to(gpu),
             pt=torch.from_numpy(np.concatenate([t, np.zeros((len(t), 1), np.float32)], 1)).to(gpu))
    model = nepx.make_pair_model(shape, hyp)
    unit = nepx.EnergyNet.random(40, [64, 64], seed=12, device=gpu)
    # FP64 per-pair gradients of the unit-range network
    clouds = oracle_clouds(q, i, j, t, shape.sites)
    G64 = O.descriptor(clouds, shape.nep_cutoff, *hyp)
    y, g = O.ffnn_forward_backward(G64, unit.layers_f64(), "tanh")
    gp = O.descriptor_grad(clouds, g, shape.nep_cutoff, *hyp)
    Ns = shape.n_sites
    raw = np.concatenate([-gp[:, :Ns].sum(1), np.cross(clouds[:, :Ns], -gp[:, :Ns]).sum(1),
                          np.cross(clouds[:, Ns:] - t.astype(np.float64)[:, None, :], -gp[:, Ns:]).sum(1)], 1)
    # energy range that puts the 60th percentile of |component| at the clip
    rng_E = CLIP / np.quantile(np.abs(raw), 0.6)
    en_min = -0.5 * rng_E
    net = nepx.EnergyNet.random(40, [64, 64], seed=12, en_min=en_min, en_max=en_min + rng_E, device=gpu)
    want = _masked_clip(rng_E * raw)
    bound = np.abs(rng_E * raw) > CLIP
    assert 0.3 < bound.mean() < 0.5                      # the clip binds on a good share of components
    fused = ops.pair_fused(model, net, T["qf"], T["pi"], T["pj"], T["pt"], clip=CLIP).cpu().numpy()
    G = ops.descriptor_forward(model, T["qf"], T["pi"], T["pj"], T["pt"])
    E, dG = ops.ffnn_forward_backward(net, G)
    split = ops.descriptor_adjoint(model, T["qf"], T["pi"], T["pj"], T["pt"], dG, E=E, clip=CLIP).cpu().numpy()
    for name, out in (("fused", fused), ("split", split)):
        got = out[:, [0, 1, 2, 4, 5, 6, 8, 9, 10]].astype(np.float64)
        assert np.abs(got).max() <= CLIP
        # components the oracle clamps (away from the boundary) come out exactly +-100
        firm = np.abs(rng_E * raw) > 1.001 * CLIP
        assert np.array_equal(got[firm], np.sign(raw[firm]) * CLIP)
        e = normwise_rel(got, want)
        eE = normwise_rel(out[:, 3], en_min + rng_E * y)
        print(f"clip ({name}, {pair_lanes} lanes/pair): {bound.mean():.2f} of components clamped, "
              f"F/tau err {e:.2e}, E err {eE:.2e}")
        assert e < 1e-4 and eE < 1e-5


def test_fd_forces_clip_keeps_nan(gpu):
    """nepx_fd_forces: forward differences of given stencil-row energies, clamped per component; a
    NaN energy row gives NaN components (not +-100), the other pairs are unaffected."""
    from nepx import _lib
    P, dx, dteta = 257, 1e-3, 2e-3
    rng = np.random.default_rng(5)
    e0 = rng.normal(0, 1, P)
    d = rng.normal(0, 0.2, (9, P))                     # |d|/dx straddles 100
    E = np.concatenate([e0[None], e0[None] + d]).astype(np.float32)
    E[2, 7] = np.nan                                   # pair 7: the y-translation row
    E[0, 11] = np.nan                                  # pair 11: the unperturbed row
    Eg = torch.from_numpy(E.reshape(-1)).to(gpu)
    out = torch.empty((P, 12), dtype=torch.float32, device=gpu)
    _lib.check(_lib.lib.nepx_fd_forces(_lib.ptr(Eg), P, dx, dteta, CLIP, _lib.ptr(out), None), "nepx_fd_forces")
    got = out.cpu().numpy().astype(np.float64)
    E64 = E.astype(np.float64)
    steps = np.array([dx] * 3 + [dteta] * 6)[:, None]
    want = _masked_clip((-(E64[1:] - E64[0]) / steps).T)
    comp = got[:, [0, 1, 2, 4, 5, 6, 8, 9, 10]]
    assert np.isnan(comp[7, 1]) and np.isfinite(np.delete(comp[7], 1)).all()
    assert np.isnan(comp[11]).all()
    ok = np.ones(P, bool)
    ok[[7, 11]] = False
    assert (np.abs(want[ok]) == CLIP).mean() > 0.3
    assert np.abs(comp[ok] - want[ok]).max() < 1e-4 * CLIP
    assert np.array_equal(got[:, 3], E[0].astype(np.float64), equal_nan=True)
