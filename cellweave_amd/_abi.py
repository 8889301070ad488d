"""ctypes binding of libcwfield.so, the C ABI declared in include/cwfield.h.

There is no fallback: if the HIP library is missing or does not load, import
of the solver fails loudly. Every solve runs through the gfx950 kernels.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcwfield.so")

CW_OK, CW_ERR_CONFIG, CW_ERR_NOT_CONVERGED, CW_ERR_NONFINITE, CW_ERR_DEVICE, CW_ERR_ARG = range(6)
CRIT_MAX_CHANGE, CRIT_REL_RESIDUAL = 0, 1
PREC_FP64, PREC_FP32_REFINE = 0, 1

# Every symbol include/*.h declares, with its ctypes signature.
EXPORTS = {}


class RelaxOpts(C.Structure):
    _fields_ = [
        ("omega", C.c_double),
        ("criterion", C.c_int32),
        ("tolerance", C.c_double),
        ("max_sweeps", C.c_uint64),
        ("precision", C.c_int32),
        ("sweeps_per_pass", C.c_int32),
        ("check_every", C.c_int32),
        ("inner_tolerance", C.c_double),
        ("max_outer", C.c_uint64),
        ("tile_shape", C.c_int32),
        ("use_graph", C.c_int32),
    ]


class RelaxStats(C.Structure):
    _fields_ = [
        ("sweeps", C.c_uint64),
        ("passes", C.c_uint64),
        ("outer_iters", C.c_uint64),
        ("span_voxels", C.c_uint64),
        ("omega", C.c_double),
        ("rel_residual", C.c_double),
        ("seconds", C.c_double),
        ("status", C.c_int32),
        ("sweeps_per_pass", C.c_int32),
        ("tile_shape", C.c_int32),
        ("reserved", C.c_int32),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class AdjointResult(C.Structure):
    _fields_ = [
        ("loss", C.c_double),
        ("dl_dd", C.c_double),
        ("dl_dg", C.c_double),
        ("sweeps", C.c_uint64),
        ("rel_residual", C.c_double),
        ("seconds", C.c_double),
        ("status", C.c_int32),
        ("pad_", C.c_int32),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_ if name != "pad_"}


def _declare():
    P = C.c_void_p
    st = C.c_int
    dp = C.POINTER(C.c_double)
    u8p = C.POINTER(C.c_uint8)
    u32p = C.POINTER(C.c_uint32)
    u64p = C.POINTER(C.c_uint64)
    d = C.c_double
    u32 = C.c_uint32
    u64 = C.c_uint64
    opts = C.POINTER(RelaxOpts)
    stats = C.POINTER(RelaxStats)
    EXPORTS.update({
        "cw_relax_opts_default": (None, [opts]),
        "cw_last_error": (C.c_char_p, []),
        "cw_version": (C.c_char_p, []),
        "cw_field_create": (st, [u32, u32, u32, d, d, d, C.POINTER(P)]),
        "cw_field_destroy": (st, [P]),
        "cw_field_set_source": (st, [P, d]),
        "cw_field_upload": (st, [P, dp, u8p, dp, dp]),
        "cw_field_download": (st, [P, dp]),
        "cw_field_relax": (st, [P, d, u64, u64p]),
        "cw_field_relax_noalloc": (st, [P, d, u64, u64p]),
        "cw_field_relax_ex": (st, [P, opts, stats]),
        "cw_field_residual": (st, [P, dp]),
        "cw_batch_relax": (st, [C.POINTER(P), u32, opts, stats]),
        "cw_campaign_relax": (st, [C.POINTER(P), u32, u32p, C.POINTER(C.c_int32), u32, opts,
                                   stats]),
        "cw_field_fill": (st, [P, d]),
        "cw_batch_time_passes": (st, [C.POINTER(P), u32, opts, u32, dp]),
        "cw_field_adjoint": (st, [P, dp, u32p, u32, opts, dp, C.POINTER(AdjointResult)]),
        "cw_autotune": (st, [C.POINTER(P), u32, opts, u32, C.POINTER(C.c_int32), dp, u32, u32p]),
        "cw_lattice_create": (st, [u32, u32, u32, C.POINTER(P)]),
        "cw_lattice_destroy": (st, [P]),
        "cw_lattice_upload": (st, [P, u32p]),
        "cw_lattice_set": (st, [P, u64p, u32p, u64]),
        "cw_field_apply_uptake": (st, [P, P, u8p, u32, d]),
        "cw_field_apply_uptake_per_cell": (st, [P, P, u8p, dp, u32]),
        "cw_field_apply_background_uptake": (st, [P, P, d]),
        "cw_field_apply_production": (st, [P, P, dp, u32]),
        "cw_field_hold_obstacles_at": (st, [P, P, d]),
        "cw_field_hold_where": (st, [P, u8p, d]),
        "cw_field_hold_bath_at": (st, [P, u8p, d]),
        "cw_field_hold_medium_bath": (st, [P, P, d]),
        "cw_field_release_all": (st, [P]),
        "cw_field_free_count": (st, [P, u64p]),
        "cw_field_inflow": (st, [P, dp]),
        "cw_field_step": (st, [P, d]),
        "cw_field_mean_per_cell": (st, [P, P, u32, dp]),
        "cw_field_download_state": (st, [P, u8p, dp, dp]),
    })


_declare()
_lib = None


def lib():
    """Load libcwfield.so (raises OSError when it is missing or unloadable)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(
                f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; "
                "g.build()'` (hipcc --offload-arch=gfx950)"
            )
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in EXPORTS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def last_error() -> str:
    return lib().cw_last_error().decode()
