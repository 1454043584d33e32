"""Training-side featurisation: descriptors of arbitrary site clouds with per-site type weights
(nepx_cloud_descriptor, the NEP / pts_to_nep surface of 05/NEP.py and 05/Processors.py) against the
FP64 oracle.  Tolerance: 1e-5 relative, row-wise max-norm."""
import os

import numpy as np
import pytest
import torch

from oracle import nep_oracle as O

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _clouds(rng, P, nd, spread):
    """Random clouds: two bodies of nd/2 points each, the second displaced by up to `spread`."""
    half = nd // 2
    a = rng.normal(0, 0.5, (P, half, 3))
    b = rng.normal(0, 0.5, (P, half, 3)) + rng.uniform(0.5, spread, (P, 1, 1)) * rng.normal(0, 1, (P, 1, 3))
    return np.concatenate([a, b], axis=1).astype(np.float32)


def _rowwise(G, G64):
    return (np.abs(G - G64).max(axis=1) / np.abs(G64).max(axis=1)).max()


@pytest.mark.parametrize("hyp,nd", [((11, 6, 4), 16), ((10, 8, 4), 12), ((17, 6, 6), 24), ((10, 10, 10), 6)])
def test_cloud_descriptor_matches_oracle(gpu, hyp, nd):
    from nepx import ops
    rng = np.random.default_rng(7)
    X = _clouds(rng, 500, nd, 2.0)
    w = rng.uniform(0.5, 2.0, nd).astype(np.float32)
    r_nep = float(np.ceil(np.linalg.norm(X - X.mean(1, keepdims=True), axis=-1).max() * 100) / 100) + 0.5
    G = ops.cloud_descriptor(torch.from_numpy(X).to(gpu), *hyp, r_nep, weights=w).cpu().numpy()
    G64 = O.descriptor(X.astype(np.float64), r_nep, *hyp, weights=w.astype(np.float64))
    e = _rowwise(G, G64)
    G1 = ops.cloud_descriptor(torch.from_numpy(X).to(gpu), *hyp, r_nep).cpu().numpy()
    e1 = _rowwise(G1, O.descriptor(X.astype(np.float64), r_nep, *hyp))
    print(f"clouds {hyp} nd={nd}: weighted {e:.2e} unweighted {e1:.2e}")
    assert e < 1e-5 and e1 < 1e-5


def test_nep_class_on_reference_clouds(gpu):
    """The reference's own cube_v2 pair clouds (01/datadir sample) through NEP.get_g_fast / pts_to_nep."""
    from nepx.NEP import NEP, calculate_cutoff, pts_to_nep
    X = np.load(os.path.join(GOLD, "x12pts_cube_v2_sample.npz"))["clouds"].astype(np.float32)
    rc = calculate_cutoff(X)
    c = X - X.mean(1, keepdims=True)
    assert rc == np.ceil(np.linalg.norm(c, axis=-1).max() * 100) / 100
    nep = NEP(X)
    nep.set_hypers([10, 8, 4, rc])
    nep.set_npts(X.shape[1] // 2)
    nep.set_device(0)
    nep.set_pts_types({"pts_types": list(range(X.shape[1]))})      # types 1..12 -> sqrt weights
    G = nep.get_g_fast()
    w = np.sqrt(np.arange(X.shape[1]) + 1.0)
    assert np.abs(nep.types - w).max() < 1e-6
    e = _rowwise(G, O.descriptor(X.astype(np.float64), rc, 10, 8, 4, weights=w))
    g_plain = nep.get_g()
    e0 = _rowwise(g_plain, O.descriptor(X.astype(np.float64), rc, 10, 8, 4))
    G2, rc2 = pts_to_nep(X, {"nep_hyper": [10, 8, 4, 100], "pts_types": list(range(X.shape[1]))}, 0, None)
    print(f"reference clouds: typed {e:.2e} plain {e0:.2e}")
    assert e < 1e-5 and e0 < 1e-5 and rc2 == rc and np.array_equal(G2, G)


def test_cloud_descriptor_rejects_missing_hypers(gpu):
    from nepx.NEP import NEP
    nep = NEP(np.zeros((4, 6, 3), np.float32) + np.arange(18, dtype=np.float32).reshape(1, 6, 3))
    nep.set_hypers([9, 9, 9, 5.0])
    with pytest.raises(NotImplementedError):
        nep.get_g_fast()
