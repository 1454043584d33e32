"""Host-side checks (no GPU): the C ABI library, weight packing, GSD I/O, shapes and workloads."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from oracle import nep_oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    from nepx import _lib
    hdr = open(os.path.join(ROOT, "include", "nepx.h")).read()
    names = sorted(set(re.findall(r"^\s*(?:int|int64_t|const char\*)\s+(nepx_\w+)\(", hdr, re.M)))
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(_lib.lib, n)]
    assert not missing, missing
    assert _lib.lib.nepx_version() >= 1


def test_struct_layouts_match_binding():
    from nepx import _lib
    sizes = (C.c_int32 * 8)()
    assert _lib.lib.nepx_struct_sizes(sizes) == 0
    assert list(sizes[:7]) == [C.sizeof(_lib.PairModel), C.sizeof(_lib.Net), _lib.SYSPARAM_DTYPE.itemsize,
                               _lib.THERMO_DTYPE.itemsize, C.sizeof(_lib.State), C.sizeof(_lib.NList),
                               C.sizeof(_lib.Tuning)]


def _emulate_mfma_net(packed, d_in, width, n_hidden, G, act):
    """Run the packed operands through the documented 16x16x4 lane layout on the CPU.

    A operand lane l: A[l%16][l/16]; B operand lane l: B[l/16][l%16];
    accumulator lane l: rows 4*(l/16)+r, column l%16.  Any mistake in the
    packing or in the K-permutation shows up as a wrong network output.
    """
    DP, HP, NL = (d_in + 15) // 16 * 16, (width + 15) // 16 * 16, n_hidden + 1

    def fwd_size(k):
        return (HP // 16) * ((DP if k == 0 else HP) // 4) * 64

    def bwd_size(k):
        return ((DP if k == 0 else HP) // 16) * (HP // 4) * 64

    fwd_off = [sum(fwd_size(j) for j in range(k)) for k in range(NL)]
    bwd0 = sum(fwd_size(j) for j in range(NL))
    bwd_off = [bwd0 + sum(bwd_size(j) for j in range(k)) for k in range(NL)]
    bias_off = [bwd_off[-1] + bwd_size(NL - 1) + k * HP for k in range(NL)]
    wout = bias_off[-1] + HP
    bout = wout + HP
    P = G.shape[0]
    Gp = np.zeros((P, DP))
    Gp[:, :d_in] = G

    def matmul(Wsrc, off, M, K, X):          # returns (M, P) = W (M x K) . X (K x P) with the lane permutation
        out = np.zeros((M, X.shape[1]))
        for mt in range(M // 16):
            for s in range(K // 4):
                for lane in range(64):
                    a = Wsrc[off + (mt * (K // 4) + s) * 64 + lane]
                    row = 16 * mt + lane % 16
                    kidx = 16 * (s // 4) + 4 * (lane // 16) + s % 4
                    out[row] += a * X[kidx]
        return out

    h = Gp.T
    ders = []
    for k in range(NL):
        K = DP if k == 0 else HP
        z = matmul(packed, fwd_off[k], HP, K, h) + packed[bias_off[k]:bias_off[k] + HP][:, None]
        h = np.tanh(z) if act == "tanh" else np.maximum(z, 0)
        ders.append(1 - h * h if act == "tanh" else (h > 0).astype(float))
    y = packed[wout:wout + HP] @ h + packed[bout]
    delta = packed[wout:wout + HP][:, None] * ders[-1]
    for k in range(NL - 1, 0, -1):
        delta = matmul(packed, bwd_off[k], HP, HP, delta) * ders[k - 1]
    dG = matmul(packed, bwd_off[0], DP, HP, delta)
    return y, dG[:d_in].T


@pytest.mark.parametrize("d_in,width,n_hidden,act", [(40, 64, 1, "tanh"), (47, 150, 3, "relu"), (60, 128, 1, "tanh")])
def test_net_packing_reproduces_dense_network(d_in, width, n_hidden, act):
    import nepx
    net = nepx.EnergyNet.random(d_in, [width] * (n_hidden + 1), seed=5, act=act, device=None)
    rng = np.random.default_rng(0)
    G = rng.uniform(0, 2, (5, d_in))
    y, dG = _emulate_mfma_net(net.packed.astype(np.float64), d_in, width, n_hidden, G, act)
    y64, g64 = O.ffnn_forward_backward(G, net.layers_f64(), act)
    assert np.allclose(y, y64, rtol=1e-5, atol=1e-5)
    assert np.allclose(dG, g64, rtol=1e-4, atol=1e-5)


def _bf16_round(x):
    """float32 -> bf16 value (as float64), round to nearest even, by integer bit arithmetic."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def _split16(x, precision="f16x3"):
    """hi and lo parts of float32 x as the device forms them (lo = 0 for plain bf16)."""
    if precision == "f16x3":
        hi = x.astype(np.float16)
        return hi.astype(np.float64), (x - hi.astype(np.float32)).astype(np.float16).astype(np.float64)
    hi = _bf16_round(x)
    if precision == "bf16":
        return hi, np.zeros_like(hi)
    return hi, _bf16_round((x - hi.astype(np.float32)).astype(np.float32))


def _operand_values(buf, precision):
    if precision == "f16x3":
        return buf.view(np.float16).astype(np.float64)
    return (buf.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def _emulate_lowp_net(buf, packed, d_in, width, n_hidden, G, act, precision="f16x3"):
    """Run the 16-bit operands through the 16x16x32 MFMA lane layout on the CPU.

    A operand lane l, element j: A[l%16][8*(l/16) + j]; B operand likewise with
    the column l%16.  K-step s of a layer maps (lane group g, j) to unit
    16*(2s + j/4) + 4g + j%4.  The activations are split into hi/lo parts
    exactly as on the device (fp16: deltas after scaling by 2^10; bf16: no
    scaling, and no lo part for plain bf16) and the products hi*hi + hi*lo +
    lo*hi are summed, so both the operand layout and the precision of the split
    are checked.
    """
    parts = 1 if precision == "bf16" else 2
    up = 1024.0 if precision == "f16x3" else 1.0
    DP, HP, NL = (d_in + 15) // 16 * 16, (width + 15) // 16 * 16, n_hidden + 1
    DPH = (DP + 31) // 32 * 32
    fsz = [parts * (HP // 16) * ((DPH if k == 0 else HP) // 32) * 512 for k in range(NL)]
    bsz = [parts * ((DP if k == 0 else HP) // 16) * (HP // 32) * 512 for k in range(NL)]
    foff = [sum(fsz[:k]) for k in range(NL)]
    boff = [sum(fsz) + sum(bsz[:k]) for k in range(NL)]
    assert len(buf) == sum(fsz) + sum(bsz)
    # fp32 layout: biases and the output layer
    f32 = [(HP // 16) * ((DP if k == 0 else HP) // 4) * 64 for k in range(NL)]
    b32 = [((DP if k == 0 else HP) // 16) * (HP // 4) * 64 for k in range(NL)]
    bias_off = [sum(f32) + sum(b32) + k * HP for k in range(NL)]
    wout = bias_off[-1] + HP
    H = _operand_values(buf, precision)

    def matmul(off, M, K, X, sc):
        ks = K // 32
        lo = (M // 16) * ks * 512
        xh, xl = _split16((X * sc).astype(np.float32), precision)
        out = np.zeros((M, X.shape[1]))
        for mt in range(M // 16):
            for s in range(ks):
                for lane in range(64):
                    for j in range(8):
                        e = off + ((mt * ks + s) * 64 + lane) * 8 + j
                        row = 16 * mt + lane % 16
                        u = 16 * (2 * s + j // 4) + 4 * (lane // 16) + j % 4
                        out[row] += H[e] * xh[u]
                        if parts == 2:
                            out[row] += H[e] * xl[u] + H[e + lo] * xh[u]
        return out / sc

    Gp = np.zeros((G.shape[0], DPH))
    Gp[:, :d_in] = G
    h = Gp.T
    ders = []
    for k in range(NL):
        K = DPH if k == 0 else HP
        z = matmul(foff[k], HP, K, h if k else h, 1.0) + packed[bias_off[k]:bias_off[k] + HP][:, None]
        h = np.tanh(z) if act == "tanh" else np.maximum(z, 0)
        ders.append(1 - h * h if act == "tanh" else (h > 0).astype(float))
    y = packed[wout:wout + HP] @ h + packed[wout + HP]
    delta = packed[wout:wout + HP][:, None] * ders[-1]
    for k in range(NL - 1, 0, -1):
        delta = matmul(boff[k], HP, HP, delta, up) * ders[k - 1]
    dG = matmul(boff[0], DP, HP, delta, up)
    return y, dG[:d_in].T


# error budgets against FP64 of the emulated 16-bit forms (E, dE/dG); plain bf16 rounds every
# operand to 8 mantissa bits
LOWP_TOL = {"f16x3": (1e-5, 1e-4), "bf16x3": (1e-5, 1e-4), "bf16": (2e-2, 5e-2)}


@pytest.mark.parametrize("precision", ["f16x3", "bf16x3", "bf16"])
@pytest.mark.parametrize("d_in,width,n_hidden,act", [(40, 64, 1, "tanh"), (60, 128, 1, "tanh")])
def test_lowp_net_packing_and_precision(d_in, width, n_hidden, act, precision):
    import nepx
    net = nepx.EnergyNet.random(d_in, [width] * (n_hidden + 1), seed=5, act=act, device=None, precision=precision)
    if precision == "f16x3":
        assert nepx.EnergyNet.random(d_in, [width] * (n_hidden + 1), seed=5, act=act, device=None).precision == "f16x3"
    rng = np.random.default_rng(1)
    G = rng.uniform(0, 2, (4, d_in)) * rng.choice([1.0, 1e-3], size=d_in)   # include small descriptor entries
    y, dG = _emulate_lowp_net(net.packed_half, net.packed.astype(np.float64), d_in, width, n_hidden, G, act,
                              precision)
    y64, g64 = O.ffnn_forward_backward(G, net.layers_f64(), act)
    tE, tG = LOWP_TOL[precision]
    eE = np.abs(y - y64).max() / np.abs(y64).max()
    eG = np.abs(dG - g64).max() / np.abs(g64).max()
    print(f"{precision} {d_in}->{width}: E {eE:.2e} dG {eG:.2e}")
    assert eE < tE and eG < tG


def test_bf16_operands_round_to_nearest_even():
    """Host packing rounds each weight to the nearest bf16 (ties to even); the lo part of the split
    form holds the rounded remainder."""
    import nepx
    W = np.zeros((32, 32), dtype=np.float32)
    W[0, :4] = [1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 1.0 + 2.0 ** -9, -3.14159265]
    layers = [(W, np.zeros(32)), (np.eye(32, dtype=np.float32), np.zeros(32)), (np.ones((1, 32)), np.zeros(1))]
    a = nepx.EnergyNet(layers, device=None, precision="bf16").packed_half
    b = nepx.EnergyNet(layers, device=None, precision="bf16x3").packed_half
    va, vb = _operand_values(a, "bf16"), _operand_values(b, "bf16x3")
    # forward operand of layer 0, M tile 0, k-step 0: lane 0 holds row 0, columns j = 0..7
    assert va[0] == 1.0 and va[1] == 1.0 + 4 * 2.0 ** -8 and va[2] == 1.0          # ties to even
    assert va[3] == -3.140625
    lo_off = 2 * 1 * 512                    # the lo part follows (HP/16) * (DPH/32) * 512 hi elements
    assert vb[lo_off + 0] == 2.0 ** -8 and vb[lo_off + 1] == -(2.0 ** -8) and vb[lo_off + 2] == 2.0 ** -9
    assert np.array_equal(_bf16_round(W[0, :4]), va[:4])


def _quad_layout(nang, lmax, n_sites, f, Z, B):
    """Emulate the four-lanes-per-pair kernel's data movement for one pair (csrc/nepx_quad.h):
    the LDS exchange xp[16 ch + 4 i + q], the reads of element q of every chunk, and
    v_mfma_f32_4x4x1_16b_f32 as D[i][j] += a(lane i) b(lane j) held by lane j, register i.
    f: (n_sites, nang+1), Z: (n_sites, NLM) per-site values; B: (nang+1, NLM) adjoint weights.
    Returns A[n][lm] as lanes hold it after the forward, the g_ang row from the per-lane l weights
    and quad sums, and c_lm per site from the adjoint products."""
    from math import factorial
    NA, NLM = nang + 1, lmax * (lmax + 2)
    NC, LC = (NA + 3) // 4, (NLM + 3) // 4
    def l_of(lm):
        l = 1
        while (l + 1) ** 2 - 1 <= lm:
            l += 1
        return l
    def m_of(lm):
        return (lm - (l_of(lm) ** 2 - 1) + 1) // 2
    def split(d):
        lo, hi = l_of(4 * d), l_of(min(4 * d + 3, NLM - 1))
        return 4 * d + 4 if hi == lo else hi * hi - 1
    k2 = lambda l, m: 1.0 if m == 0 else 2.0 * factorial(l - m) / factorial(l + m)
    # forward: Af[c][d][lane j][register i]
    Af = np.zeros((NC, LC, 4, 4))
    for s0 in range(0, n_sites, 4):
        vals = np.zeros((4, 16 * (NC + LC)))
        for q in range(4):                                   # lane q computes site s0 + q
            fv = np.zeros(4 * NC); fv[:NA] = f[s0 + q]
            zv = np.zeros(4 * LC); zv[:NLM] = Z[s0 + q]
            v = np.concatenate([fv, zv])
            for val_idx, x in enumerate(v):                  # value 4 ch + i -> xp[16 ch + 4 i + q]
                ch, i = divmod(val_idx, 4)
                vals[0, 16 * ch + 4 * i + q] = x
        xp = vals[0]
        xf = np.array([[xp[16 * c + 4 * L: 16 * c + 4 * L + 4] for c in range(NC)] for L in range(4)])
        xz = np.array([[xp[16 * (NC + d) + 4 * L: 16 * (NC + d) + 4 * L + 4] for d in range(LC)] for L in range(4)])
        for k in range(4):
            for c in range(NC):
                for d in range(LC):
                    for i in range(4):
                        for j in range(4):
                            Af[c, d, j, i] += xf[i, c, k] * xz[j, d, k]
    A = np.zeros((4 * NC, 4 * LC))
    for c in range(NC):
        for d in range(LC):
            for j in range(4):
                for i in range(4):
                    A[4 * c + i, 4 * d + j] = Af[c, d, j, i]
    # g_ang from per-lane weights and quad sums
    ga = np.zeros((NA, lmax))
    for j in range(4):
        for d in range(LC):
            lm = 4 * d + j
            kk = k2(l_of(lm), m_of(lm)) if lm < NLM else 0.0
            up = lm >= split(d)
            lo, hi = l_of(4 * d), l_of(min(4 * d + 3, NLM - 1))
            for n in range(NA):
                a2 = Af[n // 4, d, j, n % 4] ** 2
                ga[n, lo - 1] += (0.0 if up else kk) * a2
                if lo != hi:
                    ga[n, hi - 1] += (kk if up else 0.0) * a2
    # adjoint: Bf in the forward layout, c^T = B^T f per step of four sites
    Bf = np.zeros((NC, LC, 4, 4))
    for n in range(NA):
        for lm in range(NLM):
            Bf[n // 4, lm // 4, lm % 4, n % 4] = B[n, lm]
    cs = np.zeros((n_sites, 4 * LC))
    for s0 in range(0, n_sites, 4):
        cc = np.zeros((LC, 4, 4))                            # [d][lane j][register i]
        for n in range(NA):
            for d in range(LC):
                for i in range(4):
                    for j in range(4):
                        cc[d, j, i] += Bf[n // 4, d, i, n % 4] * f[s0 + j, n]
        for j in range(4):
            for d in range(LC):
                cs[s0 + j, 4 * d: 4 * d + 4] = cc[d, j]
    return A[:NA, :NLM], ga, cs[:, :NLM]


@pytest.mark.parametrize("nang,lmax,n_sites", [(6, 4, 20), (6, 6, 24), (8, 4, 12), (10, 10, 8)])
def test_quad_kernel_layout_emulation(nang, lmax, n_sites):
    """The four-lane kernel's exchange and 4x4 block products give A = f^T Z, the descriptor's
    g_ang[n][l] = sum_m K_lm^2 A_nlm^2, and c = B^T f per site (checked against the dense forms
    and the FP64 oracle's angular descriptor)."""
    from math import factorial
    rng = np.random.default_rng(nang * 100 + lmax)
    NA, NLM = nang + 1, lmax * (lmax + 2)
    f = rng.standard_normal((n_sites, NA))
    Z = rng.standard_normal((n_sites, NLM))
    B = rng.standard_normal((NA, NLM))
    A, ga, cs = _quad_layout(nang, lmax, n_sites, f, Z, B)
    assert np.allclose(A, f.T @ Z, rtol=1e-12, atol=1e-12)
    want = np.zeros((NA, lmax))
    for l in range(1, lmax + 1):
        for m in range(0, l + 1):
            k2 = 1.0 if m == 0 else 2.0 * factorial(l - m) / factorial(l + m)
            for part in range(2 if m else 1):
                lm = l * l - 1 + (0 if m == 0 else 2 * m - 1 + part)
                want[:, l - 1] += k2 * A[:, lm] ** 2
    assert np.allclose(ga, want, rtol=1e-12)
    assert np.allclose(cs, f @ B, rtol=1e-12, atol=1e-12)


def test_quad_layout_g_ang_is_the_descriptor():
    """With f and Z of a real site cloud the emulated g_ang equals the oracle's angular descriptor."""
    rng = np.random.default_rng(5)
    nrad, nang, lmax, rc = 11, 6, 4, 2.5
    cloud = rng.uniform(-1.2, 1.2, (1, 20, 3))
    G = O.descriptor(cloud, rc, nrad, nang, lmax)[0]
    r = cloud[0] - cloud[0].mean(0)
    rho = np.linalg.norm(r, axis=1) / rc
    u = r / np.linalg.norm(r, axis=1, keepdims=True)
    from scipy.special import eval_chebyt, lpmv
    x = 2 * (rho - 1) ** 2 - 1
    fc = 0.5 * (1 + np.cos(np.pi * rho))
    f = np.stack([(eval_chebyt(n, x) + 1) * 0.5 * fc for n in range(nang + 1)], 1)
    # unnormalised real solid harmonics on the unit sphere: Q_l^m(z) (Condon-Shortley) x Re/Im (x+iy)^m
    NLM = lmax * (lmax + 2)
    Z = np.zeros((20, NLM))
    phi = np.arctan2(u[:, 1], u[:, 0])
    st = np.sqrt(u[:, 0] ** 2 + u[:, 1] ** 2)
    for l in range(1, lmax + 1):
        for m in range(0, l + 1):
            Q = lpmv(m, l, u[:, 2]) / np.where(st > 0, st, 1.0) ** m if m else lpmv(0, l, u[:, 2])
            for part in range(2 if m else 1):
                lm = l * l - 1 + (0 if m == 0 else 2 * m - 1 + part)
                Z[:, lm] = Q * (st ** m) * (np.sin(m * phi) if part else np.cos(m * phi))
    _, ga, _ = _quad_layout(nang, lmax, 20, f, Z, np.zeros((nang + 1, NLM)))
    assert np.allclose(ga.ravel(), G[nrad + 1:], rtol=1e-9, atol=1e-12)


def test_half_net_needs_width_multiple_of_32():
    import nepx
    from nepx import _lib
    assert _lib.lib.nepx_net_packed_halves(47, 150, 3) > 0          # pads to 160
    assert _lib.lib.nepx_net_packed_halves(40, 48, 1) == -1         # pads to 48
    net = nepx.EnergyNet.random(40, [48, 48], seed=1, device=None)
    assert net.precision == "f32"
    with pytest.raises(ValueError):
        nepx.EnergyNet.random(40, [48, 48], seed=1, device=None, precision="f16x3")


def test_net_pack_rejects_bad_input():
    from nepx import _lib
    assert _lib.lib.nepx_net_pack(0, 64, 1, None, None, None) == -1
    assert b"bad network dims" in _lib.lib.nepx_last_error()


def test_gsd_reads_reference_init_and_round_trips(tmp_path):
    from nepx import gsd
    fr = gsd.read_frame(os.path.join(ROOT, "tests", "golden", "th_init_N200.gsd"))
    assert int(fr["particles/N"][0]) == 200
    assert np.allclose(fr["configuration/box"][:3], 33.0)
    assert fr["particles/orientation"].shape == (200, 4)
    assert np.allclose(np.linalg.norm(fr["particles/orientation"], axis=1), 1.0, atol=1e-5)
    assert np.all(fr["particles/mass"] == 1.0) and np.all(fr["particles/typeid"] == 0)   # schema defaults
    path = str(tmp_path / "t.gsd")
    w = gsd.GsdWriter(path, "w")
    x = np.random.default_rng(1).random((7, 3))
    for step in (0, 5):
        w.append({"configuration/step": np.array([step], np.uint64), "particles/N": np.array([7], np.uint32),
                  "log/nepx/position_f64": x + step})
    assert gsd.n_frames(path) == 2
    back = gsd.read_frame(path, 1)
    assert np.array_equal(back["log/nepx/position_f64"], x + 5) and int(back["configuration/step"][0]) == 5
    assert np.array_equal(gsd.read_frame(path, 0)["log/nepx/position_f64"], x)


def test_gsd_rejects_non_gsd(tmp_path):
    from nepx import gsd
    p = tmp_path / "bad.gsd"
    p.write_bytes(b"\0" * 300)
    with pytest.raises(gsd.GsdError):
        gsd.read_frame(str(p))


def test_product_shapes_equal_oracle_shapes():
    import nepx
    T = O.shape_table()
    for name, sh in T.items():
        s = nepx.get_shape(name)
        assert np.allclose(s.sites, sh["sites"]) and s.cutoff == sh["cutoff"] and s.nep_cutoff == sh["nep_cutoff"]
        assert (s.en_min, s.en_max) == (sh["en_min"], sh["en_max"])
    with pytest.raises(ValueError):
        nepx.get_shape("banana")


def test_workload_geometry():
    from nepx.configs import make_system
    s = make_system("tet10", 1000, 0.45, seed=1)
    assert np.isclose(s.L, (1000 * (1 / (6 * np.sqrt(2))) / 0.45) ** (1 / 3))
    assert np.all(np.abs(s.x) <= 0.5 * s.L)
    assert np.allclose(np.linalg.norm(s.q, axis=1), 1.0)
    # p = 2 q (x) (0, L): the body angular momentum recovered by s = q* (x) p / 2 has no scalar part
    sq = 0.5 * O.quat_mul(O.quat_conj(s.q), s.p)
    assert np.allclose(sq[:, 0], 0.0, atol=1e-12)
    c = make_system("cube8", 64, 0.3, seed=2)
    from nepx.configs import _rotmat, cubes_overlap
    for i in range(64):
        for j in range(i + 1, 64):
            t = O.min_image(c.x[j] - c.x[i], c.L)
            if np.linalg.norm(t) < np.sqrt(3):
                assert not cubes_overlap(t, _rotmat(c.q[i]), _rotmat(c.q[j]))


def test_bench_refuses_multi_gpu():
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stdout


def test_interval_allocator_first_fit_and_merge():
    from nepx.serve import _Intervals
    a = _Intervals(10)
    x, y, z = a.alloc(3), a.alloc(4), a.alloc(3)
    assert (x, y, z) == (0, 3, 7) and a.alloc(1) is None
    a.release(3, 4)
    assert a.alloc(5) is None and a.alloc(4) == 3
    a.release(0, 3)
    a.release(3, 4)
    a.release(7, 3)
    assert a.free == [(0, 10)]


def test_tuning_roundtrip_and_validation():
    import ctypes
    from nepx import _lib, tune
    old = tune.get_launch()
    try:
        tune.set_launch(tune.Tuning(gather_lanes=8, fused_blocks_per_cu=2))
        got = tune.get_launch()
        assert got["gather_lanes"] == 8 and got["fused_blocks_per_cu"] == 2
        bad = _lib.Tuning()
        bad.gather_lanes = 12
        assert _lib.lib.nepx_set_tuning(ctypes.byref(bad)) == -1
        assert b"gather_lanes" in _lib.lib.nepx_last_error()
        bad = _lib.Tuning()
        bad.pair_lanes = 2
        assert _lib.lib.nepx_set_tuning(ctypes.byref(bad)) == -1
        assert b"pair_lanes" in _lib.lib.nepx_last_error()
        tune.set_launch(tune.Tuning(gather_lanes=0, fused_blocks_per_cu=0, pair_blocks_per_cu=0,
                                    net_blocks_per_cu=0, pair_lanes=0))
        assert tune.get_launch() == {"fused_blocks_per_cu": 1, "pair_blocks_per_cu": 16, "gather_lanes": 16,
                                     "net_blocks_per_cu": 2, "pair_lanes": 4}
        tune.set_launch(tune.Tuning(pair_lanes=1))
        assert tune.get_launch()["pair_lanes"] == 1
    finally:
        tune.set_launch(tune.Tuning(**old))


def test_fused_kernel_lanes_query():
    """nepx_pair_fused_supported names the fused kernel a model and network get: 4 lanes per pair
    (site sums on matrix cores) by default where the shape allows, 1 when tuned so, 0 without one."""
    import ctypes
    import nepx
    from nepx import _lib, tune
    net = _lib.Net()
    net.d_in, net.width, net.n_hidden, net.act = 40, 64, 1, 0
    tet = nepx.make_pair_model(nepx.get_shape("tet10"), (11, 6, 4))
    old = tune.get_launch()
    try:
        q = lambda m: _lib.lib.nepx_pair_fused_supported(ctypes.byref(m), ctypes.byref(net))
        tune.set_launch(tune.Tuning(**{**old, "pair_lanes": 4}))
        assert q(tet) == 4
        tune.set_launch(tune.Tuning(**{**old, "pair_lanes": 1}))
        assert q(tet) == 1
        ref = nepx.make_pair_model(nepx.get_shape("cube_v2"), (10, 10, 10))   # D = 121: split kernels only
        net.d_in, net.width, net.n_hidden = 121, 150, 3
        assert q(ref) == 0
    finally:
        tune.set_launch(tune.Tuning(**old))


def test_coordinate_descent_finds_separable_minimum(tmp_path):
    from nepx import tune
    target = {"skin": 0.4, "gather_lanes": 8, "fused_blocks_per_cu": 2}

    def cost(t):
        return (10 * abs(t.skin - target["skin"]) + abs(t.gather_lanes - target["gather_lanes"]) / 4
                + abs(t.fused_blocks_per_cu - target["fused_blocks_per_cu"]) + 1.0)

    best, trials = tune.coordinate_descent(cost, tune.DEFAULT_SPACE, tune.Tuning(),
                                           ["skin", "gather_lanes", "fused_blocks_per_cu"])
    assert (best.skin, best.gather_lanes, best.fused_blocks_per_cu) == (0.4, 8, 2)
    assert len(trials) == len({str(t.tuning) for t in trials})        # no candidate measured twice
    at = tune.Autotuner(str(tmp_path / "t.json"))
    at._store("k", best)
    assert at._load()["k"]["gather_lanes"] == 8


def test_label_host_helpers_match_oracle():
    from scipy.spatial.transform import Rotation
    from nepx import label
    rng = np.random.default_rng(5)
    a = rng.uniform(-3.2, 3.2, (50, 3))
    R = label.euler_zxz(a)
    assert np.abs(R - O.euler_zxz(a)).max() < 1e-14
    q = label.matrix_to_quat(R)
    qx = Rotation.from_matrix(R).as_quat()                       # scalar last, sign free
    ref = np.concatenate([qx[:, 3:], qx[:, :3]], axis=1)
    assert np.abs(np.abs(np.sum(q * ref, axis=1)) - 1.0).max() < 1e-12
    assert np.abs(label.cube_structure_beads() - O.cube_structure_beads()).max() == 0.0


REF_MODELS = "/root/reference/nep_simulator/models"


@pytest.mark.skipif(not os.path.isdir(REF_MODELS), reason="reference checkpoints are only in the build container")
@pytest.mark.parametrize("name,hyp", [("model_th_C104", (10, 8, 4)), ("model_th_C100", (10, 10, 10)),
                                      ("model_nep_C104", (10, 8, 4)), ("model_nep_C100", (10, 10, 10))])
def test_reference_checkpoint_packing(name, hyp):
    """ForceEvaluator.read_model on the reference's own TheEnergyNet(150, 3, D) checkpoints
    (ForceEvaluator.py:55-61, loaded with the weights-only loader): the packed MFMA operands, run
    through the documented lane layout, reproduce the checkpoint's module in FP64 (energy and
    dE/dG) on descriptor rows of the reference's initial configuration (th_init_N200.gsd, th_v1)."""
    import torch
    import nepx
    from nepx.ForceEvaluator import ForceEvaluator
    from nepx.GsdHandler import GsdHandler
    D = nepx.n_descriptor(*hyp)
    fe = ForceEvaluator()
    fe.torch_device = None                      # pack on the host only
    fe.setNN_dimensions(150, 3)
    fe.setNd(D)
    fe.read_model(os.path.join(REF_MODELS, name + ".pth"))
    assert fe.net.precision == "f32" and fe.net.act == "relu" and fe.net.n_hidden == 3
    h = GsdHandler(os.path.join(os.path.dirname(__file__), "golden", "th_init_N200.gsd"))
    x, q, L = h.get_central_pos().astype(np.float64), h.getOrientations().astype(np.float64), h.getLx()
    sh = O.shape_table()["th_v1"]
    pairs = O.brute_pairs(x, L, sh["cutoff"])[:12]
    clouds, _ = O.pair_clouds(x, q, pairs, sh["sites"], L)
    G = O.descriptor(clouds, sh["nep_cutoff"], *hyp)
    y, dG = _emulate_mfma_net(fe.net.packed.astype(np.float64), D, 150, 3, G, "relu")
    mod = fe.energy_net.double()
    Gt = torch.from_numpy(G).requires_grad_(True)
    yt = mod(Gt)[:, 0]
    (gt,) = torch.autograd.grad(yt.sum(), Gt)
    y64, g64 = yt.detach().numpy(), gt.numpy()
    assert np.abs(y - y64).max() < 1e-5 * np.abs(y64).max()
    assert np.abs(dG - g64).max() < 1e-4 * np.abs(g64).max()
    # a transposed first layer or a dropped hidden layer would not reproduce the module
    assert np.abs(y64).max() > 0 and np.abs(g64).max() > 0


def test_bf16_forms_refused_for_relu_networks():
    import nepx
    with pytest.raises(ValueError):
        nepx.EnergyNet.random(40, [64, 64], seed=1, act="relu", device=None, precision="bf16")
    with pytest.raises(ValueError):
        nepx.EnergyNet.random(40, [64, 64], seed=1, act="relu", device=None, precision="bf16x3")
    assert nepx.EnergyNet.random(40, [64, 64], seed=1, act="tanh", device=None, precision="bf16").precision == "bf16"


def test_poisson_trace_builds_job_arrays_lazily():
    """A long job trace holds only seeds until a job is looked at; its arrays, built on demand,
    are the same after a release and rebuild, and a deep copy does not build them."""
    import copy
    from nepx.serve import LazySpec, poisson_jobs
    jobs = poisson_jobs(rate=20.0, duration=30.0, seed=2025, shapes=("cube8", "tet10", "cyl12"))
    assert len(jobs) > 400 and all(isinstance(j.spec, LazySpec) and j.spec._spec is None for j in jobs)
    j = jobs[7]
    n = j.n                                  # the size is known without building the arrays
    assert j.spec._spec is None and n > 0
    x0 = j.spec.x.copy()
    assert j.spec._spec is not None and x0.shape == (n, 3)
    c = copy.deepcopy(jobs[8])
    assert c.spec._spec is None
    j.spec.release()
    assert j.spec._spec is None and np.array_equal(j.spec.x, x0)


def test_isa_phase_counter_finds_every_site_trip():
    """The static per-phase instruction counter of the four-lane fused kernel (gfx950 assembly): the
    unrolled iteration of the cube8 instance cuts into 4 forward trips, the network and 3 more
    adjoint trips plus the tail."""
    import subprocess
    import sys
    out = subprocess.run([sys.executable, os.path.join(ROOT, "anisotropic-ml_amd", "isa_phases.py"), "--sites", "8"],
                         capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stderr[-2000:]
    rows = {l.split("|")[1].strip(): [c.strip() for c in l.split("|")[2:-1]]
            for l in out.stdout.splitlines() if l.startswith("| ")}
    assert rows["forward"][-1] == "4" and rows["network"][-1] == "1" and rows["adjoint"][-1] == "3"
    assert int(rows["forward"][1]) == 4 * 48 and int(rows["adjoint"][1]) == 3 * 84   # 4x4x1 products per trip
    total = int(rows["total per 16-pair wave iteration"][0])
    assert 2000 < total < 6000, total
