"""ctypes binding of libpycwb_amd.so (the C ABI in include/pycwb_amd.h).

The library is built in-tree (``pycwb_amd/libpycwb_amd.so``).  There is no CPU
fallback: importing a GPU entry point without the library, or calling one
without a ROCm device, raises.
"""

from __future__ import annotations

import ctypes as C
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpycwb_amd.so")

PYCWB_OK = 0
PYCWB_ERR_ARG = 1
PYCWB_ERR_HIP = 2
SKIP_MASKED_PIXELS = 1
REGULATOR_EARLY_EXIT = 2
TILE_ENERGY_BOUND = 4
REGULATOR_BOUNDS = 8
DEFAULT_FLAGS = SKIP_MASKED_PIXELS | REGULATOR_EARLY_EXIT | TILE_ENERGY_BOUND | REGULATOR_BOUNDS
WSTATS_PER_CLUSTER = 24
CLASS_STATS_PER_CLUSTER = 16 * 24
TICKETS_PER_GROUP = 16   # uint32 ticket counters per detector group of a batch

_lock = threading.Lock()
_lib = None

vp = C.c_void_p
i32 = C.c_int
i64 = C.c_int64
f32 = C.c_float
f64 = C.c_double
sz = C.c_size_t


class BatchDesc(C.Structure):
    _fields_ = [
        ("n_clusters", i32),
        ("n_sky", i64),
        ("sky_valid_indices", vp),
        ("n_valid", i64),
        ("order", vp),
        ("n_groups", i32),
        ("group_first", i32 * 8),
        ("group_count", i32 * 8),
        ("group_nifo", i32 * 8),
        ("n_pix", vp),
        ("net", vp),
        ("K", vp),
        ("rms_off", vp),
        ("td_off", vp),
        ("rms", vp),
        ("td", vp),
        ("net_nifo", vp),
        ("net_off", vp),
        ("net_En", vp),
        ("FP", vp),
        ("FX", vp),
        ("ml", vp),
        ("sky_order", vp),
        ("flags", i32),
        ("gamma_regulator", f64),
        ("reg0", f32),
        ("netCC", f64),
        ("tickets", vp),
        ("counts", vp),
        ("best", vp),
        ("REG1", vp),
        ("l_max", vp),
        ("sky_stat_max", vp),
        ("reg_f64", vp),
        ("block", i32),
        ("grid", i32),
        ("use_graph", i32), ("weight_stats", vp),
        ("pixel_class", vp), ("class_stats", vp), ("class_rem", vp), ("class_rem_off", vp),
        ("class_bounds", vp), ("class_bounds_bytes", sz),
    ]


class StatsDesc(C.Structure):
    _fields_ = [("n_clusters", C.c_int), ("which", vp), ("n_sky", C.c_int64), ("n_pix", vp),
                ("nifo", vp), ("net", vp), ("K", vp), ("rms_off", vp), ("td_off", vp),
                ("rms", vp), ("td", vp), ("net_off", vp), ("net_En", vp), ("FP", vp),
                ("FX", vp), ("ml", vp), ("l_max", vp), ("REG1", vp), ("reg0", C.c_float),
                ("lk_off", vp), ("lookup", vp), ("xt_off", vp), ("xt", vp),
                ("xgb_rho_mode", C.c_int), ("s32", vp), ("s32_off", vp), ("s64", vp),
                ("s64_off", vp), ("scal", vp), ("S_snr", vp)]


class GradDesc(C.Structure):
    _fields_ = [("n_clusters", C.c_int), ("clusters", vp), ("n_ifo", C.c_int), ("n_pix", vp),
                ("net", vp), ("K", vp), ("rms_off", vp), ("td_off", vp), ("rms", vp),
                ("td", vp), ("geo", vp), ("net_nifo", vp), ("net_En", vp), ("REG1", vp),
                ("reg0", C.c_double), ("theta", vp), ("phi", vp), ("value", vp), ("grad", vp),
                ("grad_td", vp), ("grad_scale", vp)]


class SubnetDesc(C.Structure):
    _fields_ = [("n_which", C.c_int), ("which", vp), ("n_ifo", C.c_int), ("n_sky", C.c_int64),
                ("n_pix", vp), ("net", vp), ("K", vp), ("rms_off", vp), ("td_off", vp),
                ("rms", vp), ("td", vp), ("net_off", vp), ("net_En", vp), ("net_Es", vp),
                ("FP", vp), ("FX", vp), ("ml", vp), ("subcut", C.c_double), ("tiles", vp),
                ("sky", vp), ("lk_off", vp), ("lookup", vp), ("xt_off", vp), ("xt", vp),
                ("scratch", vp), ("scr_off", vp), ("mra", vp), ("sky_order", vp)]


class XtalkDesc(C.Structure):
    _fields_ = [("n_clusters", C.c_int), ("n_pixels", C.c_int64), ("n_slots", C.c_int64),
                ("pix_off", vp), ("nlayer", vp), ("time", vp), ("hash_off", vp),
                ("n_res", C.c_int), ("max_layer", C.c_int), ("layers", vp), ("table", vp),
                ("coeff", vp), ("n_coeff", C.c_int64), ("hash_keys", vp), ("hash_vals", vp),
                ("counts", vp), ("row_off", vp), ("sort_keys", vp), ("block_sums", vp),
                ("lookup", vp), ("xt_off", vp), ("xt", vp), ("xt_capacity", C.c_int64)]


class SelectDesc(C.Structure):
    _fields_ = [("n_ifo", C.c_int), ("n_freq", C.c_int), ("n_time", C.c_int64),
                ("energy", vp), ("n_lag", C.c_int), ("shift_bins", vp), ("veto", vp),
                ("edge_bins", C.c_int), ("ib", C.c_int), ("ie", C.c_int), ("eo", C.c_double),
                ("combined", vp), ("flags", vp), ("live", vp), ("row_off", vp),
                ("lag_off", vp), ("capacity", C.c_int64), ("freq", vp), ("time", vp),
                ("value", vp), ("det_energy", vp), ("det_index", vp)]


class PacketDesc(C.Structure):
    _fields_ = [("maps", vp), ("n_maps", C.c_int), ("T", C.c_int64), ("M1", C.c_int),
                ("jb", C.c_int64), ("je", C.c_int64), ("mL", C.c_int), ("mH", C.c_int),
                ("mean", C.c_double), ("p", C.c_int64 * 9), ("init", C.c_int), ("cur", vp)]


class ClusterDesc(C.Structure):
    _fields_ = [("n_lag", C.c_int), ("lag_off", vp), ("n_pix", C.c_int64), ("freq", vp),
                ("time", vp), ("energy", vp), ("det_energy", vp), ("noise_rms", vp),
                ("n_ifo", C.c_int), ("n_freq", C.c_int), ("n_time", C.c_int), ("kf", C.c_int),
                ("kt", C.c_int), ("dt", C.c_double), ("half_rate", C.c_double),
                ("n_sub", C.c_double), ("work", vp), ("work_bytes", C.c_size_t),
                ("index_map", vp), ("label", vp), ("lag_cluster_off", vp), ("cluster_off", vp),
                ("members", vp),
                ("c_energy", vp), ("c_time", vp), ("c_freq", vp), ("c_subnet", vp),
                ("c_subrho", vp)]


class ThresholdDesc(C.Structure):
    _fields_ = [("maps", vp), ("n_ifo", C.c_int), ("n_freq", C.c_int), ("n_time", C.c_int64),
                ("time_major", C.c_int), ("begin", C.c_int64), ("count", C.c_int64),
                ("hi", C.c_double), ("work", vp), ("work_bytes", C.c_size_t), ("values", vp),
                ("counts", vp), ("sorted", vp), ("chunk_sums", vp)]


class G2gDesc(C.Structure):
    _fields_ = [("flat", vp), ("n", C.c_int64), ("nL", C.c_int64), ("nR", C.c_int64),
                ("work", vp), ("work_bytes", C.c_size_t), ("count", vp), ("sorted", vp),
                ("lo", C.c_double), ("hi", C.c_double), ("chunk_sums", vp), ("n_valid", vp),
                ("alp", C.c_double), ("avr", C.c_double), ("ALP", C.c_double),
                ("scale", C.c_double), ("out", vp)]


# name -> (restype, argtypes)
SIGNATURES = {
    "pycwb_version": (C.c_char_p, []),
    "pycwb_last_error": (C.c_char_p, []),
    "pycwb_sky_scan_workspace_bytes": (sz, [i32, i32, i32]),
    "pycwb_td_relayout": (i32, [i32, i32, i32, vp, vp, vp, sz, vp]),
    "pycwb_class_rem_entries": (i64, [i32]),
    "pycwb_class_bound_scratch_bytes": (sz, [i32, i32]),
    "pycwb_regulator_class_bounds": (i32, [i32, i32, vp, vp, vp, i64, vp, vp, vp, vp, vp, vp, vp]),
    "pycwb_regulator_term_bounds": (i32, [i32, i32, vp, vp, vp, i64, vp, vp, vp, vp]),
    "pycwb_dpf_regulator": (i32, [i32, i32, i64, vp, vp, vp, f64, f64, f32, vp, i64, i32, vp,
                                  vp, vp, sz, vp]),
    "pycwb_sky_scan_plain": (i32, [i32, i32, i64, i32, vp, vp, vp, vp, vp, vp, f64, f64, vp,
                                   i64, vp, i32, vp, vp, vp, vp, sz, vp]),
    "pycwb_batch_create": (i32, [C.POINTER(BatchDesc), C.POINTER(vp)]),
    "pycwb_batch_launch": (i32, [vp, vp]),
    "pycwb_batch_launch_stage": (i32, [vp, i32, vp]),
    "pycwb_batch_destroy": (i32, [vp]),
    "pycwb_batch_run_noalloc": (i32, [C.POINTER(BatchDesc), vp]),
    "pycwb_valu_probe": (i32, [vp, i32, i32, vp]),
    "pycwb_batch_noalloc_max_clusters": (sz, []),
    "pycwb_sky_stats_scratch32": (sz, [i32, i32]),
    "pycwb_sky_stats_scratch64": (sz, [i32, i32]),
    "pycwb_sky_statistics": (i32, [C.POINTER(StatsDesc), vp]),
    "pycwb_wdm_t2w": (i32, [vp, i64, i64, i32, i32, i32, vp, i32, vp, i32, vp, vp]),
    "pycwb_wdm_t2w_shifted": (i32, [vp, i64, vp, i32, i32, vp, i32, vp, i32, vp, vp]),
    "pycwb_sky_grad": (i32, [C.POINTER(GradDesc), vp]),
    "pycwb_td_vectors": (i32, [vp, i64, vp, vp, i64, i32, vp, vp, vp, i32, i32, i32, vp, i64,
                               i32, vp]),
    "pycwb_subnet_tiles": (sz, [i64]),
    "pycwb_subnet_mra_scratch": (sz, [i32, i32]),
    "pycwb_subnet_sky": (i32, [C.POINTER(SubnetDesc), vp]),
    "pycwb_subnet_mra": (i32, [C.POINTER(SubnetDesc), vp]),
    "pycwb_xtalk_scan_blocks": (sz, [i64]),
    "pycwb_xtalk_count": (i32, [C.POINTER(XtalkDesc), vp]),
    "pycwb_xtalk_fill": (i32, [C.POINTER(XtalkDesc), vp]),
    "pycwb_wdm_w2t": (i32, [vp, i64, i32, i32, vp, i32, vp, i32, vp, vp, i64, vp]),
    "pycwb_select_count": (i32, [C.POINTER(SelectDesc), vp]),
    "pycwb_select_emit": (i32, [C.POINTER(SelectDesc), vp]),
    "pycwb_max_energy_shift": (i32, [vp, i64, vp, i32, vp, vp]),
    "pycwb_max_energy_packet": (i32, [C.POINTER(PacketDesc), vp]),
    "pycwb_max_energy_finish": (i32, [vp, i64, i32, i32, vp]),
    "pycwb_cluster_map_init": (i32, [vp, i64, vp]),
    "pycwb_cluster_work_bytes": (i32, [C.POINTER(ClusterDesc), C.POINTER(C.c_size_t)]),
    "pycwb_cluster_pixels": (i32, [C.POINTER(ClusterDesc), vp]),
    "pycwb_threshold_work_bytes": (i32, [C.POINTER(ThresholdDesc), C.POINTER(C.c_size_t)]),
    "pycwb_threshold_stats": (i32, [C.POINTER(ThresholdDesc), vp]),
    "pycwb_g2g_work_bytes": (i32, [C.POINTER(G2gDesc), C.POINTER(C.c_size_t)]),
    "pycwb_g2g_phase": (i32, [C.POINTER(G2gDesc), i32, vp]),
}


def load():
    """Load the library once; raise if it is missing."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; "
                "g.build()'` (hipcc --offload-arch=gfx950); there is no CPU fallback")
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
        return lib


class PycwbError(RuntimeError):
    pass


def check(rc: int, what: str) -> None:
    if rc == PYCWB_OK:
        return
    if rc == PYCWB_ERR_ARG:
        reason = _lib.pycwb_last_error().decode(errors="replace") if _lib is not None else ""
        reason = reason[len(what) + 2:] if reason.startswith(what + ": ") else ""
        raise ValueError(f"{what}: arguments rejected by the C ABI" +
                         (f" ({reason})" if reason else ""))
    detail = ""
    if _lib is not None:
        detail = _lib.pycwb_last_error().decode(errors="replace")
    raise PycwbError(f"{what}: HIP runtime error (code {rc}) {detail}")
