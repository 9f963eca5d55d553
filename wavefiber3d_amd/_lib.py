"""ctypes binding of ``libwfpo.so`` (the C-ABI in include/wf_po.h).

The library is the product: there is no CPU fallback.  ``lib()`` raises if the
shared object is missing or does not export every declared entry point.
"""
from __future__ import annotations

import ctypes as ct
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libwfpo.so")

WF_OK = 0
STATUS_NAMES = {
    0: "WF_OK",
    1: "WF_ERR_INVALID_ARG",
    2: "WF_ERR_GRID_TOO_COARSE",
    3: "WF_ERR_OOM",
    4: "WF_ERR_HIP",
    5: "WF_ERR_CANCELLED",
    6: "WF_ERR_UNSUPPORTED",
    7: "WF_ERR_NO_GEOMETRY",
}
WF_MEM_HOST, WF_MEM_DEVICE = 0, 1
WF_PREC_FP32, WF_PREC_BF16X3, WF_PREC_FP32_VALU = 0, 1, 2

EXPORTS = (
    "wf_abi_version", "wf_last_error", "wf_default_opts", "wf_create", "wf_destroy",
    "wf_set_geometry", "wf_solve", "wf_po_solve", "wf_po_solve_fp64_direct", "wf_currents",
    "wf_octree_nodes", "wf_interp", "wf_get_stats", "wf_po_vjp", "wf_geometry_info",
    "wf_get_tune", "wf_set_tune", "wf_autotune", "wf_batch_create", "wf_batch_destroy",
    "wf_batch_add", "wf_batch_solve", "wf_batch_info", "wf_heightfield_points",
    "wf_read_heightfield", "wf_average_grid", "wf_po_solve_fp64_finite", "wf_rough2d_solve",
    "wf_rough2d_solve_noalloc",
)
WF_DEPTH_AUTO = -1


class WfError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


class Surface(ct.Structure):
    _fields_ = [
        ("x", ct.c_void_p), ("y", ct.c_void_p), ("z", ct.c_void_p),
        ("nx", ct.c_void_p), ("ny", ct.c_void_p), ("nz", ct.c_void_p),
        ("area", ct.c_void_p),
        ("n", ct.c_int64),
        ("memspace", ct.c_int32),
    ]


class Incidence(ct.Structure):
    _fields_ = [
        ("theta", ct.c_double), ("phi", ct.c_double), ("lambda_m", ct.c_double),
        ("pol", ct.c_char),
        ("eta_re", ct.c_double), ("eta_im", ct.c_double),
        ("beam_sigma_m", ct.c_double),
    ]


class Rough2dIncidence(ct.Structure):
    _fields_ = [
        ("theta", ct.c_double), ("phi", ct.c_double), ("lambda_m", ct.c_double),
        ("eta_re", ct.c_double), ("eta_im", ct.c_double),
        ("pol", ct.c_char),
    ]


class Grid(ct.Structure):
    _fields_ = [("n_theta", ct.c_int32), ("n_phi", ct.c_int32)]


class Opts(ct.Structure):
    _fields_ = [
        ("struct_size", ct.c_int32),
        ("device", ct.c_int32),
        ("precision", ct.c_int32),
        ("strict_grid", ct.c_int32),
        ("beta", ct.c_double),
        ("box_enlarge", ct.c_double),
        ("lambda_sample_m", ct.c_double),
        ("has_box", ct.c_int32),
        ("box", ct.c_double * 6),
        ("workspace_cap_bytes", ct.c_int64),
        ("cancel_flag", ct.POINTER(ct.c_int32)),
        ("field_chunk", ct.c_int32),
        ("use_graph", ct.c_int32),
        ("solve_streams", ct.c_int32),
        ("ref_dtheta_m", ct.c_double),
        ("ref_dphi_m", ct.c_double),
    ]


KERNEL_KINDS = ("currents", "leaf", "phi_pass", "theta_pass", "trans_table", "epilogue",
                "direct_fp64", "other")
WF_K_COUNT = 8


class Stats(ct.Structure):
    _fields_ = [
        ("ms_total", ct.c_double),
        ("kern_ms", ct.c_double * WF_K_COUNT),
        ("kern_launches", ct.c_int32 * WF_K_COUNT),
        ("kern_work", ct.c_double * WF_K_COUNT),
        ("leaf_pairs", ct.c_int64),
        ("depth", ct.c_int32), ("n_levels", ct.c_int32),
        ("level_nt", ct.c_int32 * 16), ("level_np", ct.c_int32 * 16),
        ("level_nodes", ct.c_int32 * 16),
        ("agg_transitions", ct.c_int32), ("agg_bytes", ct.c_double),
    ]


class Tune(ct.Structure):
    _fields_ = [
        ("leaf_s_per_pair", ct.c_double),
        ("agg_s_per_byte", ct.c_double),
        ("level_s", ct.c_double),
        ("leaf_quad", ct.c_int32),
        ("chunk_solves", ct.c_int32),
        ("agg_kernel", ct.c_int32),
        ("agg_fft_rows", ct.c_int32),
        ("calibrated", ct.c_int32),
        ("agg_fused", ct.c_int32),
        ("theta_tile", ct.c_int32),
        ("phi_rows", ct.c_int32),
        ("beta", ct.c_double),
    ]


_lib = None


def lib(path: str | None = None):
    """Load libwfpo.so; raise loudly if it is missing or incomplete."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(
            f"HIP library {p} is not built; run `python -c 'import __graft_entry__ as g; g.build()'`"
        )
    L = ct.CDLL(p)
    missing = [s for s in EXPORTS if not hasattr(L, s)]
    if missing:
        raise RuntimeError(f"{p} does not export {missing}")
    vp = ct.c_void_p
    L.wf_abi_version.restype = ct.c_int32
    L.wf_last_error.restype = ct.c_char_p
    L.wf_default_opts.argtypes = [ct.POINTER(Opts)]
    L.wf_create.argtypes = [ct.POINTER(vp), ct.POINTER(Opts)]
    L.wf_destroy.argtypes = [vp]
    L.wf_destroy.restype = None
    L.wf_set_geometry.argtypes = [vp, ct.POINTER(Surface), ct.c_int32, vp]
    L.wf_solve.argtypes = [vp, ct.POINTER(Incidence), ct.c_int32, Grid, vp, vp, ct.c_int32, vp]
    L.wf_po_solve.argtypes = [vp, ct.POINTER(Surface), ct.POINTER(Incidence), ct.c_int32, Grid,
                              ct.c_int32, vp, vp, vp]
    L.wf_po_solve_fp64_direct.argtypes = [vp, ct.POINTER(Surface), ct.POINTER(Incidence),
                                          ct.c_int32, Grid, vp, vp, vp]
    L.wf_currents.argtypes = [vp, ct.POINTER(Surface), ct.POINTER(Incidence), ct.c_int32,
                              ct.c_int32, vp, vp]
    L.wf_octree_nodes.argtypes = [vp, vp, vp, vp, vp]
    L.wf_interp.argtypes = [vp, ct.c_int32, ct.c_int32, ct.c_int32, vp, ct.c_int32, ct.c_int32,
                            ct.c_double, vp, vp]
    L.wf_get_stats.argtypes = [vp, ct.POINTER(Stats)]
    L.wf_po_vjp.argtypes = [vp, ct.POINTER(Incidence), ct.c_int32, Grid, vp, vp, vp, vp, vp,
                            ct.c_int32, ct.c_int32, vp]
    L.wf_geometry_info.argtypes = [vp, ct.POINTER(ct.c_int64), ct.POINTER(ct.c_int32),
                                   ct.POINTER(ct.c_double)]
    L.wf_get_tune.argtypes = [vp, ct.POINTER(Tune)]
    L.wf_set_tune.argtypes = [vp, ct.POINTER(Tune)]
    L.wf_autotune.argtypes = [vp, ct.POINTER(Incidence), ct.c_int32, Grid, ct.POINTER(ct.c_int32),
                              vp]
    L.wf_batch_create.argtypes = [ct.POINTER(vp), ct.POINTER(Opts), ct.c_int32]
    L.wf_batch_destroy.argtypes = [vp]
    L.wf_batch_destroy.restype = None
    L.wf_batch_add.argtypes = [vp, ct.POINTER(Surface), ct.c_int32, ct.POINTER(ct.c_int32)]
    L.wf_batch_solve.argtypes = [vp, ct.POINTER(Incidence), ct.c_int32, Grid, vp, ct.c_int32, vp]
    L.wf_batch_info.argtypes = [vp, ct.POINTER(ct.c_int32), vp, vp, vp]
    L.wf_heightfield_points.argtypes = [vp, ct.c_int32, ct.c_int32, ct.c_float, ct.c_float,
                                        ct.c_int32, ct.c_int32, vp, ct.c_int32, vp]
    L.wf_read_heightfield.argtypes = [ct.c_char_p, ct.c_int32, ct.c_int32, vp]
    L.wf_rough2d_solve.argtypes = [vp, vp, ct.c_int32, ct.POINTER(Rough2dIncidence), ct.c_int32,
                                   ct.c_int32, ct.c_double, ct.c_double, vp, ct.c_int32, vp]
    L.wf_rough2d_solve_noalloc.argtypes = [vp, vp, ct.c_int32, ct.POINTER(Rough2dIncidence),
                                           ct.c_int32, ct.c_int32, ct.c_double, ct.c_double, vp,
                                           vp]
    L.wf_po_solve_fp64_finite.argtypes = [vp, ct.POINTER(Surface), ct.POINTER(Incidence),
                                          ct.c_int32, Grid, ct.c_double, vp, vp, vp]
    L.wf_average_grid.argtypes = [vp, ct.c_int32, ct.c_int32, ct.c_int32, ct.c_int32, ct.c_int32,
                                  vp, ct.c_int32, vp]
    for name in EXPORTS:
        if name not in ("wf_abi_version", "wf_last_error", "wf_default_opts", "wf_destroy",
                        "wf_batch_destroy"):
            getattr(L, name).restype = ct.c_int
    if path is None:
        _lib = L
    return L


def check(status: int) -> None:
    if status != WF_OK:
        msg = lib().wf_last_error()
        raise WfError(status, msg.decode() if msg else "")
