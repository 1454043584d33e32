"""ctypes binding of libnepx.so (declared in include/nepx.h).

The library is required: importing this module without a built library
raises, so no code path can silently fall back to host arithmetic.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("NEPX_LIB", os.path.join(_HERE, "libnepx.so"))

MAX_SITES = 16
OK, EINVAL, EUNSUP, EHIP, ECAP = 0, -1, -2, -3, -4   # return codes of the C ABI
ACT_TANH = 0
ACT_RELU = 1
NET_F32 = 0
NET_F16X3 = 1
NET_BF16 = 2
NET_BF16X3 = 3
NET_PRECISIONS = {"f32": NET_F32, "f16x3": NET_F16X3, "bf16": NET_BF16, "bf16x3": NET_BF16X3}

c_i32p = C.POINTER(C.c_int32)
c_f32p = C.POINTER(C.c_float)
c_f64p = C.POINTER(C.c_double)


class PairModel(C.Structure):
    _fields_ = [("n_sites", C.c_int32), ("nrad", C.c_int32), ("nang", C.c_int32), ("lmax", C.c_int32),
                ("r_nep", C.c_float), ("sites", C.c_float * (3 * MAX_SITES))]


class Net(C.Structure):
    _fields_ = [("d_in", C.c_int32), ("width", C.c_int32), ("n_hidden", C.c_int32), ("act", C.c_int32),
                ("en_min", C.c_float), ("en_max", C.c_float), ("w_packed", C.c_void_p), ("w_half", C.c_void_p),
                ("precision", C.c_int32), ("pad", C.c_int32)]


_STATE_PTRS = ["x", "y", "z", "vx", "vy", "vz", "q0", "q1", "q2", "q3", "p0", "p1", "p2", "p3", "mass",
               "fx", "fy", "fz", "tx", "ty", "tz", "upot", "xb", "yb", "zb", "qf", "sys", "sp", "th",
               "partials", "dp"]


class State(C.Structure):
    _fields_ = [("n", C.c_int32), ("n_sys", C.c_int32)] + [(k, C.c_void_p) for k in _STATE_PTRS]


_NL_PTRS = ["cell_count", "cell_start", "cell_fill", "cell_list", "cell_pos", "cell_of", "row_count", "row_start",
            "rrow_count", "rrow_start", "nbr", "slot_row", "rnbr", "cidx", "blk_count", "flags", "pair_i",
            "pair_j", "pair_t", "rpos", "pair_e", "rrec"]


class NList(C.Structure):
    _fields_ = [("cap", C.c_int64), ("n_cells_total", C.c_int32)] + [(k, C.c_void_p) for k in _NL_PTRS]


class Tuning(C.Structure):
    _fields_ = [("fused_blocks_per_cu", C.c_int32), ("pair_blocks_per_cu", C.c_int32),
                ("gather_lanes", C.c_int32), ("net_blocks_per_cu", C.c_int32), ("pair_lanes", C.c_int32)]


# numpy mirrors of the per-system records
SYSPARAM_DTYPE = np.dtype([("L", "f8"), ("kT", "f8"), ("dt", "f8"), ("tau", "f8"), ("ndof", "f8"),
                           ("rndof", "f8"), ("moi", "f8", 3), ("r_list", "f8"), ("rc", "f8"),
                           ("skin", "f8"), ("pad", "f8", 4), ("p_start", "i4"), ("p_end", "i4"),
                           ("blk_start", "i4"), ("blk_end", "i4"), ("nc", "i4"), ("cell_off", "i4"),
                           ("active", "i4"), ("ipad", "i4", 9)])
THERMO_DTYPE = np.dtype([("xi", "f8"), ("eta", "f8"), ("xi_r", "f8"), ("eta_r", "f8"), ("theta", "f8"),
                         ("e_r", "f8"), ("Kt", "f8"), ("Kr", "f8"), ("U", "f8"), ("max_disp2", "f8"),
                         ("steps", "f8"), ("budget", "f8"), ("pending", "f8"), ("fault", "f8"),
                         ("pad", "f8", 2)])

(FLAG_REBUILD, FLAG_OVERFLOW, FLAG_PL, FLAG_P, FLAG_BUILDS, FLAG_MAXROW, FLAG_FROZEN, FLAG_SOON, FLAG_LONGSYS,
 FLAG_TAG, FLAG_RTAG) = range(11)
MAX_ROW = 2048       # longest neighbour row the list build sorts
NVT_STEP_TWO, NVT_TRANSLATE, NVT_ROTATE, NVT_THERMO = 2, 3, 4, 5
N_FLAGS = 16


class NepxError(RuntimeError):
    pass


class RowLimitError(NepxError):
    """A particle has more neighbours within the list radius than the build sorts (MAX_ROW):
    the system has collapsed.  `system` is the index of one such system (-1: unknown)."""

    def __init__(self, msg, system=-1):
        super().__init__(msg)
        self.system = system


def _load():
    if not os.path.exists(LIB_PATH):
        raise NepxError(f"libnepx.so not found at {LIB_PATH}; build it with "
                        "`python anisotropic-ml_amd/build.py` (the HIP path has no host fallback)")
    lib = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    sig = {
        "nepx_last_error": ([], C.c_char_p),
        "nepx_version": ([], C.c_int),
        "nepx_struct_sizes": ([c_i32p], C.c_int),
        "nepx_set_tuning": ([C.POINTER(Tuning)], C.c_int),
        "nepx_get_tuning": ([C.POINTER(Tuning)], C.c_int),
        "nepx_net_packed_floats": ([C.c_int32, C.c_int32, C.c_int32], C.c_int64),
        "nepx_net_pack": ([C.c_int32, C.c_int32, C.c_int32, C.POINTER(c_f32p), C.POINTER(c_f32p), c_f32p],
                          C.c_int),
        "nepx_net_packed_halves": ([C.c_int32, C.c_int32, C.c_int32], C.c_int64),
        "nepx_net_pack_half": ([C.c_int32, C.c_int32, C.c_int32, C.POINTER(c_f32p), vp], C.c_int),
        "nepx_net_packed_lowp": ([C.c_int32, C.c_int32, C.c_int32, C.c_int32], C.c_int64),
        "nepx_net_pack_lowp": ([C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.POINTER(c_f32p), vp], C.c_int),
        "nepx_descriptor_fwd": ([C.POINTER(PairModel), vp, vp, vp, vp, vp, C.c_int64, vp, vp], C.c_int),
        "nepx_ffnn_fwd_bwd": ([C.POINTER(Net), vp, vp, C.c_int64, vp, vp, vp], C.c_int),
        "nepx_descriptor_adjoint": ([C.POINTER(PairModel), vp, vp, vp, vp, vp, C.c_int64, vp, vp, C.c_float,
                                     vp, vp], C.c_int),
        "nepx_pair_fused": ([C.POINTER(PairModel), C.POINTER(Net), vp, vp, vp, vp, vp, C.c_int64, C.c_float,
                             vp, vp], C.c_int),
        "nepx_pair_fused_list": ([C.POINTER(PairModel), C.POINTER(Net), C.POINTER(State), C.POINTER(NList),
                                  C.c_float, vp, vp], C.c_int),
        "nepx_pair_fused_supported": ([C.POINTER(PairModel), C.POINTER(Net)], C.c_int),
        "nepx_bead_pair_forces": ([vp, C.c_int32, C.c_float, C.c_float, C.c_float, vp, vp, vp, vp, vp, C.c_int64,
                                   C.c_float, vp, vp], C.c_int),
        "nepx_cloud_descriptor": ([vp, C.c_int64, C.c_int32, vp, C.c_int32, C.c_int32, C.c_int32, C.c_float,
                                   vp, vp], C.c_int),
        "nepx_nlist_build": ([C.POINTER(State), C.POINTER(NList), C.c_int32, vp], C.c_int),
        "nepx_nlist_from_pairs": ([C.POINTER(State), C.POINTER(NList), vp, C.c_int64, vp], C.c_int),
        "nepx_pairs_filter": ([C.POINTER(State), C.POINTER(NList), vp], C.c_int),
        "nepx_pairs_filter_noalloc": ([C.POINTER(State), C.POINTER(NList), vp], C.c_int),
        "nepx_forces_gather": ([C.POINTER(State), C.POINTER(NList), vp, vp], C.c_int),
        "nepx_nvt_step": ([C.POINTER(State), C.POINTER(NList), vp], C.c_int),
        "nepx_nvt_init": ([C.POINTER(State), C.POINTER(NList), C.c_int32, vp], C.c_int),
        "nepx_nvt_part": ([C.POINTER(State), C.POINTER(NList), C.c_int32, vp], C.c_int),
        "nepx_stencil_rows": ([vp, vp, vp, vp, C.c_int64, C.c_float, C.c_float, vp, vp, vp, vp, vp], C.c_int),
        "nepx_fd_forces": ([vp, C.c_int64, C.c_float, C.c_float, C.c_float, vp, vp], C.c_int),
        "nepx_md_step": ([C.POINTER(State), C.POINTER(NList), C.POINTER(PairModel), C.POINTER(Net), C.c_float,
                          vp, vp], C.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    sizes = (C.c_int32 * 8)()
    lib.nepx_struct_sizes(sizes)
    expect = [C.sizeof(PairModel), C.sizeof(Net), SYSPARAM_DTYPE.itemsize, THERMO_DTYPE.itemsize,
              C.sizeof(State), C.sizeof(NList), C.sizeof(Tuning)]
    if list(sizes[:7]) != expect:
        raise NepxError(f"struct layout mismatch: library {list(sizes[:7])} vs binding {expect}")
    return lib


lib = _load()
EXPORTED = [n for n in dir(lib) if n.startswith("nepx_")]


def check(rc, what=""):
    if rc != 0:
        msg = lib.nepx_last_error().decode(errors="replace")
        raise NepxError(f"{what or 'nepx call'} failed ({rc}): {msg}")


def ptr(t):
    """Device (or host) address of a contiguous torch tensor, or None."""
    if t is None:
        return None
    assert t.is_contiguous(), "nepx buffers must be contiguous"
    return C.c_void_p(t.data_ptr())


def stream_handle(stream=None):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)
