"""Parity of the HIP pair kernels (descriptor, MFMA network, adjoint, fused) with the FP64 oracle.

Tolerances (stated against FP64):
  descriptor G and pair energy E : 1e-5 relative (row-wise max-norm for G)
  dE/dG, forces and torques       : 1e-4 relative (max-norm over the pair set)
Network precisions f32 and f16x3 are held to these.  The bf16 forms (tanh networks only) have
budgets of their own and their measured errors are printed: bf16x3 (hi/lo bf16 split, ~16-bit
products) BF16X3_TOL; plain bf16 (8-bit mantissas on both MFMA operands) BF16_TOL.
"""
import numpy as np
import pytest
import torch

from oracle import nep_oracle as O
from pairs_util import normwise_rel, oracle_clouds, random_pairs, rowwise_rel

pytestmark = pytest.mark.gpu

# (E, dE/dG or forces and torques) relative to FP64 for the bf16 network forms
BF16_TOL = (5e-2, 1.5e-1)
BF16X3_TOL = (1e-4, 3e-4)


def _tol(precision):
    return {"bf16": BF16_TOL, "bf16x3": BF16X3_TOL}.get(precision, (1e-5, 1e-4))


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


@pytest.mark.parametrize("precision", ["f32", "f16x3", "bf16x3", "bf16"])
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
    if precision in ("bf16", "bf16x3") and act == "relu":
        with pytest.raises(ValueError):     # refused: bf16 rounding flips ReLU units at zero
            nepx.EnergyNet.random(d_in, widths, seed=7, act=act, device=gpu, precision=precision)
        return
    net = nepx.EnergyNet.random(d_in, widths, seed=7, act=act, en_min=-2.0, en_max=3.0, device=gpu,
                                precision=precision)
    E, dG = ops.ffnn_forward_backward(net, torch.from_numpy(G).to(gpu))
    y, g = O.ffnn_forward_backward(G.astype(np.float64), net.layers_f64(), act)
    E64 = -2.0 + 5.0 * y
    dG64 = 5.0 * g
    eE = normwise_rel(E.cpu().numpy(), E64)
    eG = normwise_rel(dG.cpu().numpy(), dG64)
    print(f"ffnn {d_in}->{widths} {act} {precision}: E {eE:.3e} dG {eG:.3e}")
    tE, tG = _tol(precision)
    assert eE < tE
    assert eG < tG


@pytest.mark.parametrize("d_in,widths,act", [(121, [150] * 4, "relu"), (47, [150] * 4, "relu"),
                                             (40, [64, 64], "tanh"), (60, [128, 128], "tanh")])
def test_ffnn_few_pairs_kernel_equals_wide_kernel(gpu, d_in, widths, act):
    """Up to four 16-pair tiles per CU the FP32 network runs one tile per block, the layers' unit
    blocks spread over the block's waves; it gives the one-wave-per-tile kernel's results bit for
    bit.  The capacity argument picks the kernel (241 pairs either way)."""
    import ctypes as C
    import nepx
    from nepx import _lib
    rng = np.random.default_rng(12)
    P = 241
    G = np.concatenate([rng.uniform(0, 20, (P // 2, d_in)), rng.normal(0, 0.2, (P - P // 2, d_in))])
    G = torch.from_numpy(G.astype(np.float32)).to(gpu)
    net = nepx.EnergyNet.random(d_in, widths, seed=5, act=act, en_min=-2.0, en_max=3.0, device=gpu,
                                precision="f32")
    n_dev = torch.tensor([P], dtype=torch.int32, device=gpu)
    cn = net.cstruct()
    out = []
    for maxp in (P, 80000):   # 16 tiles: one per block; 5,000 tiles: one per wave
        E = torch.full((P,), float("nan"), device=gpu)
        dG = torch.full((P, d_in), float("nan"), device=gpu)
        _lib.check(_lib.lib.nepx_ffnn_fwd_bwd(C.byref(cn), _lib.ptr(G), _lib.ptr(n_dev), maxp, _lib.ptr(E),
                                              _lib.ptr(dG), _lib.stream_handle(None)), "nepx_ffnn_fwd_bwd")
        torch.cuda.synchronize()
        out.append((E.cpu().numpy(), dG.cpu().numpy()))
    (E0, g0), (E1, g1) = out
    assert np.all(np.isfinite(E0)) and np.all(np.isfinite(g0))
    assert np.array_equal(E0, E1) and np.array_equal(g0, g1)
    y, g = O.ffnn_forward_backward(G.cpu().numpy().astype(np.float64), net.layers_f64(), act)
    assert normwise_rel(E0, -2.0 + 5.0 * y) < 1e-5 and normwise_rel(g0, 5.0 * g) < 1e-4


def test_ffnn_split_large_inputs_and_energy_range(gpu):
    """The fp16 hi/lo path with descriptor entries up to 4000 (|G| <= Nd^2 bounds real rows far
    below fp16's 65504) and an energy range of 5000: the backward deltas are carried without the
    range, so a large range cannot overflow fp16.  Same 1e-5 / 1e-4 budget against FP64."""
    import nepx
    from nepx import ops
    rng = np.random.default_rng(8)
    big = rng.uniform(0, 4000, (320, 40)) * (rng.uniform(0, 1, (320, 40)) < 0.2)
    G = np.concatenate([big, rng.uniform(0, 0.5, (320, 40))]).astype(np.float32)   # second half unsaturated
    net = nepx.EnergyNet.random(40, [64, 64], seed=9, en_min=-1000.0, en_max=4000.0, device=gpu,
                                precision="f16x3")
    E, dG = ops.ffnn_forward_backward(net, torch.from_numpy(G).to(gpu))
    y, g = O.ffnn_forward_backward(G.astype(np.float64), net.layers_f64(), "tanh")
    eE = normwise_rel(E.cpu().numpy(), -1000.0 + 5000.0 * y)
    eG = normwise_rel(dG.cpu().numpy(), 5000.0 * g)
    print(f"ffnn f16x3 large inputs / range: E {eE:.3e} dG {eG:.3e}")
    assert np.all(np.isfinite(dG.cpu().numpy()))
    assert eE < 1e-5 and eG < 1e-4


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


@pytest.mark.parametrize("shape_name,hyp", CASES)
def test_pass_kernels_equal_pair_kernels(gpu, shape_name, hyp):
    """Few pairs of a multi-pass shape run their l-range passes on separate waves (descriptor and
    adjoint); the results equal the one-lane-per-pair kernels' bit for bit.  The capacity argument
    picks the kernels (400 pairs either way; one-pass shapes take the same kernels twice)."""
    import ctypes as C
    import nepx
    from nepx import _lib
    shape, (q, i, j, t), T = _inputs(gpu, shape_name, n_pairs=400, seed=6)
    model = nepx.make_pair_model(shape, hyp)
    D = nepx.n_descriptor(*hyp)
    P = len(i)
    n_dev = torch.tensor([P], dtype=torch.int32, device=gpu)
    dG = torch.from_numpy(np.random.default_rng(13).standard_normal((P, D)).astype(np.float32)).to(gpu)
    E = torch.from_numpy(np.linspace(-1, 1, P).astype(np.float32)).to(gpu)
    args = (_lib.ptr(T["qf"]), _lib.ptr(T["pi"]), _lib.ptr(T["pj"]), _lib.ptr(T["pt"]), _lib.ptr(n_dev))
    res = []
    for maxp in (P, 100000):   # 7 blocks of 64 pairs with the passes on separate waves; one lane per pair
        G = torch.full((P, D), float("nan"), device=gpu)
        out = torch.full((P, 12), float("nan"), device=gpu)
        s = _lib.stream_handle(None)
        _lib.check(_lib.lib.nepx_descriptor_fwd(C.byref(model), *args, maxp, _lib.ptr(G), s), "nepx_descriptor_fwd")
        _lib.check(_lib.lib.nepx_descriptor_adjoint(C.byref(model), *args, maxp, _lib.ptr(dG), _lib.ptr(E), 100.0,
                                                    _lib.ptr(out), s), "nepx_descriptor_adjoint")
        torch.cuda.synchronize()
        res.append((G.cpu().numpy(), out.cpu().numpy()))
    (G0, o0), (G1, o1) = res
    assert np.all(np.isfinite(G0)) and np.all(np.isfinite(o0))
    assert np.array_equal(G0, G1)
    assert np.array_equal(o0, o1)


@pytest.fixture(params=[4, 1], ids=["lanes4", "lanes1"])
def pair_lanes(request):
    """Fused kernel variant: four lanes per pair (site sums on MFMA) or one lane per pair."""
    from nepx import tune
    old = tune.get_launch()
    tune.set_launch(tune.Tuning(**{**old, "pair_lanes": request.param}))
    yield request.param
    tune.set_launch(tune.Tuning(**old))


@pytest.mark.parametrize("precision", ["f32", "f16x3", "bf16x3", "bf16"])
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
    if pair_lanes == 1:
        assert e < 1e-6
    else:
        # the four-lane kernel's descriptor rows differ from the split kernels' in the last bits;
        # bf16 rounds them to 8 (16 for the split form) mantissa bits, where such a difference can
        # flip a rounding
        assert e < {"bf16": 5e-3, "bf16x3": 1e-4}.get(precision, 1e-5)
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
    tE, tF = _tol(precision)
    assert eE < tE
    assert max(eF, eTi, eTj) < tF


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


def test_bf16_cylinder_network_errors(gpu, pair_lanes):
    """Twisted cylinders (12 helix sites, hypers 17 6 6, D = 60) with the 60 -> 128 -> 128 -> 1 network:
    energies, forces and both torques of the bf16 and bf16x3 network forms against the FP32 network
    and against FP64, printed; each within its budget."""
    import nepx
    from nepx import ops
    shape, (q, i, j, t), T = _inputs(gpu, "cyl12", n_pairs=2000, seed=17)
    hyp = (17, 6, 6)
    model = nepx.make_pair_model(shape, hyp)
    base = nepx.EnergyNet.random(60, [128, 128], seed=3, en_min=-1.0, en_max=4.0, device=gpu, precision="f32")
    clouds = oracle_clouds(q, i, j, t, shape.sites)
    G64 = O.descriptor(clouds, shape.nep_cutoff, *hyp)
    y, g = O.ffnn_forward_backward(G64, base.layers_f64(), "tanh")
    gp = O.descriptor_grad(clouds, 5.0 * g, shape.nep_cutoff, *hyp)
    Ns = shape.n_sites
    want = dict(E=-1.0 + 5.0 * y, F=-gp[:, :Ns].sum(1), Ti=np.cross(clouds[:, :Ns], -gp[:, :Ns]).sum(1),
                Tj=np.cross(clouds[:, Ns:] - t.astype(np.float64)[:, None, :], -gp[:, Ns:]).sum(1))
    outs = {}
    for prec in ("f32", "bf16x3", "bf16"):
        o = ops.pair_fused(model, base.with_precision(prec), T["qf"], T["pi"], T["pj"], T["pt"]).cpu().numpy()
        outs[prec] = dict(E=o[:, 3], F=o[:, 0:3], Ti=o[:, 4:7], Tj=o[:, 8:11])
    for prec in ("bf16x3", "bf16"):
        e64 = {k: normwise_rel(outs[prec][k], want[k]) for k in want}
        e32 = {k: normwise_rel(outs[prec][k], outs["f32"][k].astype(np.float64)) for k in want}
        print(f"cyl12 {prec} ({pair_lanes} lanes/pair): vs FP64 " + " ".join(f"{k} {v:.2e}" for k, v in e64.items())
              + " | vs FP32 " + " ".join(f"{k} {v:.2e}" for k, v in e32.items()))
        tE, tF = _tol(prec)
        assert e64["E"] < tE and max(e64["F"], e64["Ti"], e64["Tj"]) < tF
