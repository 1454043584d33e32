"""nepx_pairs_filter_noalloc, the allocation-free variant of nepx_pairs_filter, called through the
C ABI: bitwise the same compacted pairs, translations, slot map and pair count as the filter with
the caller's workspace, on two calls in a row, the second with a smaller system (whatever the first
call left in the library's scratch must not leak into the second)."""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

OUT = ("pair_i", "pair_j", "pair_t", "cidx")


def _arena(gpu, n, phi, seed):
    import nepx
    from nepx.configs import make_system
    net = nepx.EnergyNet.random(40, [64, 64], seed=1, device=gpu)
    ar = nepx.Arena([make_system("tet10", n, phi, seed=seed)], nepx.get_shape("tet10"), (11, 6, 4), net,
                    device=gpu)
    ar.init()
    ar.build_list(True)
    torch.cuda.synchronize()
    return ar


def _snapshot(ar):
    f = ar.flags()
    P, PL = int(f[3]), int(f[2])
    return P, PL, {"pair_i": ar.lt["pair_i"][:P].clone(), "pair_j": ar.lt["pair_j"][:P].clone(),
                   "pair_t": ar.lt["pair_t"][:4 * P].clone(), "cidx": ar.lt["cidx"][:PL].clone()}


def test_noalloc_filter_equals_filter(gpu):
    from nepx import _lib
    for n, phi, seed in ((6000, 0.45, 1), (700, 0.3, 2)):
        print(f"noalloc filter: tet10 n={n} phi={phi} seed={seed}", flush=True)
        ar = _arena(gpu, n, phi, seed)
        ar.filter_pairs()
        torch.cuda.synchronize()
        P, PL, want = _snapshot(ar)
        assert P > 0 and PL >= P
        for k in OUT:
            ar.lt[k].fill_(-7)
        ar.lt["flags"][3] = -1
        nl = _lib.NList.from_buffer_copy(ar.nl)
        nl.blk_count = None                       # the variant takes no workspace from the caller
        _lib.check(_lib.lib.nepx_pairs_filter_noalloc(C.byref(ar.st), C.byref(nl), None),
                   "nepx_pairs_filter_noalloc")
        torch.cuda.synchronize()
        P2, PL2, got = _snapshot(ar)
        assert (P2, PL2) == (P, PL)
        for k in OUT:
            assert torch.equal(got[k], want[k]), k
        # the pairs are the FP64 cut of the positions: spot-check against the host
        x = ar.system_arrays(0)["x"]
        i, j = want["pair_i"].cpu().numpy(), want["pair_j"].cpu().numpy()
        L = ar.specs[0].L
        t = x[j] - x[i]
        t = np.where(t > 0.5 * L, t - L, t)
        t = np.where(t < -0.5 * L, t + L, t)
        assert np.all(np.sum(t * t, 1) < ar.specs[0].rc ** 2)


def test_noalloc_filter_rejects_capacity_beyond_scratch(gpu):
    from nepx import _lib
    ar = _arena(gpu, 500, 0.3, 3)
    nl = _lib.NList.from_buffer_copy(ar.nl)
    nl.cap = (1 << 26) + 1
    assert _lib.lib.nepx_pairs_filter_noalloc(C.byref(ar.st), C.byref(nl), None) == _lib.ECAP
    assert b"scratch" in _lib.lib.nepx_last_error()
