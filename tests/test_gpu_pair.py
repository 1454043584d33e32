"""Parity of the HIP pair kernels (descriptor, MFMA network, adjoint, fused) with the FP64 oracle.

Tolerances (stated against FP64):
  descriptor G and pair energy E : 1e-5 relative (row-wise max-norm for G)
  dE/dG, forces and torques       : 1e-4 relative (max-norm over the pair set)
"""
import numpy as np
import pytest
import torch

from oracle import nep_oracle as O
from pairs_util import normwise_rel, oracle_clouds, random_pairs, rowwise_rel

pytestmark = pytest.mark.gpu

CASES = [("tet10", (11, 6, 4)), ("cube8", (11, 6, 4)), ("cyl12", (17, 6, 6)), ("cube_v2", (10, 8, 4)),
         ("th_v1", (10, 10, 10))]


def _inputs(gpu, shape_name, n_pairs=777, seed=1):
    import nepx
    shape = nepx.get_shape(shape_name)
    rng = np.random.default_rng(seed)
    scale = shape.nep_cutoff / 2.5
    q, i, j, t = random_pairs(rng, 64, n_pairs, 0.9 * scale, 2.4 * scale)
    dev = gpu
    tens = dict(qf=torch.from_numpy(q).to(dev), pi=torch.from_numpy(i).to(dev), pj=torch.from_numpy(j).to(dev),
                pt=torch.from_numpy(np.concatenate([t, np.zeros((len(t), 1), np.float32)], 1)).to(dev))
    return shape, (q, i, j, t), tens


@pytest.mark.parametrize("shape_name,hyp", CASES)
def test_descriptor_forward_matches_oracle(gpu, shape_name, hyp):
    import nepx
    from nepx import ops
    shape, (q, i, j, t), T = _inputs(gpu, shape_name)
    model = nepx.make_pair_model(shape, hyp)
    G = ops.descriptor_forward(model, T["qf"], T["pi"], T["pj"], T["pt"]).cpu().numpy().astype(np.float64)
    clouds = oracle_clouds(q, i, j, t, shape.sites)
    G64 = O.descriptor(clouds, shape.nep_cutoff, *hyp)
    err = rowwise_rel(G, G64)
    print(f"{shape_name}: G rel err max {err.max():.3e} median {np.median(err):.3e}")
    assert np.all(np.isfinite(G))
    assert err.max() < 1e-5


def test_empty_pair_set(gpu):
    import nepx
    from nepx import ops
    shape = nepx.get_shape("tet10")
    model = nepx.make_pair_model(shape, (11, 6, 4))
    z = torch.zeros(0, dtype=torch.int32, device=gpu)
    G = ops.descriptor_forward(model, torch.zeros(8, 4, device=gpu), z, z, torch.zeros(0, 4, device=gpu))
    assert G.shape == (0, 40)


@pytest.mark.parametrize("precision", ["f32", "f16x3"])
@pytest.mark.parametrize("d_in,widths,act", [(40, [64, 64], "tanh"), (60, [128, 128], "tanh"),
                                             (40, [64, 64], "relu"), (47, [150] * 4, "relu"),
                                             (121, [150] * 4, "relu")])
def test_ffnn_matches_oracle(gpu, d_in, widths, act, precision):
    import nepx
    from nepx import ops
    rng = np.random.default_rng(3)
    P = 1000
    # realistic descriptor magnitudes plus a small-scale block that keeps tanh unsaturated
    G = np.concatenate([rng.uniform(0, 20, (P // 2, d_in)), rng.normal(0, 0.2, (P - P // 2, d_in))]).astype(np.float32)
    net = nepx.EnergyNet.random(d_in, widths, seed=7, act=act, en_min=-2.0, en_max=3.0, device=gpu,
                                precision=precision)
    E, dG = ops.ffnn_forward_backward(net, torch.from_numpy(G).to(gpu))
    y, g = O.ffnn_forward_backward(G.astype(np.float64), net.layers_f64(), act)
    E64 = -2.0 + 5.0 * y
    dG64 = 5.0 * g
    eE = normwise_rel(E.cpu().numpy(), E64)
    eG = normwise_rel(dG.cpu().numpy(), dG64)
    print(f"ffnn {d_in}->{widths} {act} {precision}: E {eE:.3e} dG {eG:.3e}")
    assert eE < 1e-5
    assert eG < 1e-4


@pytest.mark.parametrize("shape_name,hyp", CASES)
def test_adjoint_matches_oracle(gpu, shape_name, hyp):
    import nepx
    from nepx import ops
    shape, (q, i, j, t), T = _inputs(gpu, shape_name, n_pairs=400, seed=5)
    model = nepx.make_pair_model(shape, hyp)
    D = nepx.n_descriptor(*hyp)
    rng = np.random.default_rng(11)
    dG = rng.standard_normal((len(i), D)).astype(np.float32)
    out = ops.descriptor_adjoint(model, T["qf"], T["pi"], T["pj"], T["pt"], torch.from_numpy(dG).to(gpu)).cpu().numpy()
    clouds = oracle_clouds(q, i, j, t, shape.sites)
    gp = O.descriptor_grad(clouds, dG.astype(np.float64), shape.nep_cutoff, *hyp)
    Ns = shape.n_sites
    F = -gp[:, :Ns].sum(1)
    Ti = np.cross(clouds[:, :Ns], -gp[:, :Ns]).sum(1)
    Tj = np.cross(clouds[:, Ns:] - t.astype(np.float64)[:, None, :], -gp[:, Ns:]).sum(1)
    eF = normwise_rel(out[:, 0:3], F)
    eTi = normwise_rel(out[:, 4:7], Ti)
    eTj = normwise_rel(out[:, 8:11], Tj)
    print(f"{shape_name}: F {eF:.3e} tau_i {eTi:.3e} tau_j {eTj:.3e}")
    assert max(eF, eTi, eTj) < 1e-4


@pytest.fixture(params=[4, 1], ids=["lanes4", "lanes1"])
def pair_lanes(request):
    """Fused kernel variant: four lanes per pair (site sums on MFMA) or one lane per pair."""
    from nepx import tune
    old = tune.get_launch()
    tune.set_launch(tune.Tuning(**{**old, "pair_lanes": request.param}))
    yield request.param
    tune.set_launch(tune.Tuning(**old))


@pytest.mark.parametrize("precision", ["f32", "f16x3"])
@pytest.mark.parametrize("shape_name,hyp,widths", [("tet10", (11, 6, 4), [64, 64]), ("cube8", (11, 6, 4), [64, 64]),
                                                   ("cyl12", (17, 6, 6), [128, 128])])
def test_fused_matches_pipeline(gpu, shape_name, hyp, widths, precision, pair_lanes):
    """The fused kernel equals descriptor -> network -> adjoint run as separate kernels to FP32
    rounding (the one-lane kernel sums in the same order; the four-lane kernel sums the sites on
    matrix cores and over lanes, so it is held to 1e-5), and FP64 end to end."""
    import nepx
    from nepx import ops
    shape, (q, i, j, t), T = _inputs(gpu, shape_name, n_pairs=1000, seed=9)
    model = nepx.make_pair_model(shape, hyp)
    D = nepx.n_descriptor(*hyp)
    net = nepx.EnergyNet.random(D, widths, seed=2, device=gpu, precision=precision)
    G = ops.descriptor_forward(model, T["qf"], T["pi"], T["pj"], T["pt"])
    E, dG = ops.ffnn_forward_backward(net, G)
    ref = ops.descriptor_adjoint(model, T["qf"], T["pi"], T["pj"], T["pt"], dG, E=E, clip=100.0).cpu().numpy()
    fused = ops.pair_fused(model, net, T["qf"], T["pi"], T["pj"], T["pt"], clip=100.0).cpu().numpy()
    e = normwise_rel(fused, ref)
    print(f"fused ({pair_lanes} lanes/pair) vs pipeline {shape_name} {precision}: {e:.3e}")
    assert e < (1e-6 if pair_lanes == 1 else 1e-5)
    # and against FP64 end to end (pair energy, analytic forces and both torques)
    clouds = oracle_clouds(q, i, j, t, shape.sites)
    G64 = O.descriptor(clouds, shape.nep_cutoff, *hyp)
    y, g = O.ffnn_forward_backward(G64, net.layers_f64(), "tanh")
    gp = O.descriptor_grad(clouds, g, shape.nep_cutoff, *hyp)
    Ns = shape.n_sites
    Ti = np.cross(clouds[:, :Ns], -gp[:, :Ns]).sum(1)
    Tj = np.cross(clouds[:, Ns:] - t.astype(np.float64)[:, None, :], -gp[:, Ns:]).sum(1)
    eE, eF = normwise_rel(fused[:, 3], y), normwise_rel(fused[:, 0:3], -gp[:, :Ns].sum(1))
    eTi, eTj = normwise_rel(fused[:, 4:7], Ti), normwise_rel(fused[:, 8:11], Tj)
    print(f"  vs FP64: E {eE:.3e} F {eF:.3e} tau_i {eTi:.3e} tau_j {eTj:.3e}")
    assert eE < 1e-5
    assert max(eF, eTi, eTj) < 1e-4


def test_fused_result_independent_of_list_position(gpu, pair_lanes):
    """Each pair's output is bitwise the same wherever the pair sits in the list: the 4x4 block
    products, quad sums and 16-pair network tiles never mix pairs."""
    import nepx
    from nepx import ops
    shape, (q, i, j, t), T = _inputs(gpu, "tet10", n_pairs=517, seed=31)
    model = nepx.make_pair_model(shape, (11, 6, 4))
    net = nepx.EnergyNet.random(40, [64, 64], seed=6, device=gpu)
    out = ops.pair_fused(model, net, T["qf"], T["pi"], T["pj"], T["pt"], clip=100.0).cpu().numpy()
    perm = np.random.default_rng(0).permutation(len(i))
    pt = torch.from_numpy(np.ascontiguousarray(T["pt"].cpu().numpy()[perm])).to(gpu)
    out_p = ops.pair_fused(model, net, T["qf"], T["pi"][torch.from_numpy(perm).to(gpu)].contiguous(),
                           T["pj"][torch.from_numpy(perm).to(gpu)].contiguous(), pt, clip=100.0).cpu().numpy()
    assert np.array_equal(out_p, out[perm])


def test_fused_relu_matches_fp64(gpu, pair_lanes):
    """ReLU networks (the activation of the reference's TheEnergyNet) through both fused kernels."""
    import nepx
    from nepx import ops
    shape, (q, i, j, t), T = _inputs(gpu, "tet10", n_pairs=333, seed=21)
    hyp = (11, 6, 4)
    model = nepx.make_pair_model(shape, hyp)
    net = nepx.EnergyNet.random(40, [64, 64], seed=4, act="relu", en_min=-1.0, en_max=2.0, device=gpu)
    fused = ops.pair_fused(model, net, T["qf"], T["pi"], T["pj"], T["pt"], clip=0.0).cpu().numpy()
    clouds = oracle_clouds(q, i, j, t, shape.sites)
    G64 = O.descriptor(clouds, shape.nep_cutoff, *hyp)
    y, g = O.ffnn_forward_backward(G64, net.layers_f64(), "relu")
    gp = O.descriptor_grad(clouds, 3.0 * g, shape.nep_cutoff, *hyp)
    Ns = shape.n_sites
    Ti = np.cross(clouds[:, :Ns], -gp[:, :Ns]).sum(1)
    eE = normwise_rel(fused[:, 3], -1.0 + 3.0 * y)
    eF = normwise_rel(fused[:, 0:3], -gp[:, :Ns].sum(1))
    eT = normwise_rel(fused[:, 4:7], Ti)
    print(f"relu fused ({pair_lanes} lanes/pair) vs FP64: E {eE:.3e} F {eF:.3e} tau_i {eT:.3e}")
    assert eE < 1e-5
    assert max(eF, eT) < 1e-4
