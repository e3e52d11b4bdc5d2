"""ctypes mirror of include/trans_hip.h and the loader of libtrans_hip.so.

The structs are the C-ABI's plain-old-data types; the loader refuses to fall back to
anything else: if the HIP library has not been built, importing the solver fails loudly.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtrans_hip.so")

ML = 7


class TransParams(C.Structure):
    _fields_ = [
        ("kind_spatial", C.c_int), ("kind_roe", C.c_int), ("if_euler", C.c_int),
        ("if_turb_org", C.c_int), ("kind_model", C.c_int), ("kind_dual", C.c_int),
        ("kchld", C.c_int), ("if_rk", C.c_int), ("kitm", C.c_int), ("kind_hybrid", C.c_int),
        ("kind_subhy", C.c_int), ("if_comppress", C.c_int), ("kind_cc", C.c_int),
        ("kind_rotmod", C.c_int), ("if_precondition", C.c_int), ("cfl_sub_step", C.c_int),
        ("ki_ini", C.c_int), ("implicit_fp32", C.c_int),
        ("xm", C.c_double), ("rxm2", C.c_double), ("ppf", C.c_double),
        ("csthlnd", C.c_double), ("csth1", C.c_double), ("ref", C.c_double),
        ("omega1", C.c_double), ("tiny", C.c_double),
        ("vxf", C.c_double), ("vyf", C.c_double), ("vzf", C.c_double),
        ("dt_lusgs_main", C.c_double), ("cfl_sub_start", C.c_double), ("cfl_sub_end", C.c_double),
        ("gama_max", C.c_double), ("gama_min", C.c_double),
        ("prod_lmt", C.c_double), ("visratio", C.c_double), ("fsti", C.c_double),
        ("a1", C.c_double), ("akapa", C.c_double), ("beta_star", C.c_double),
        ("sigk1", C.c_double), ("sigo1", C.c_double), ("beta1", C.c_double), ("gama1", C.c_double),
        ("sigk2", C.c_double), ("sigo2", C.c_double), ("beta2", C.c_double), ("gama2", C.c_double),
        ("p0in", C.c_double), ("t0in", C.c_double), ("pouthub", C.c_double),
        ("ps00", C.c_double), ("ts00", C.c_double),
        ("rkf", C.c_double), ("rof", C.c_double), ("rmiuf", C.c_double),
        ("f_limiter1", C.c_double), ("f_limiter2", C.c_double), ("alscale", C.c_double),
    ]


I3 = C.c_int * 3


class TransPatch(C.Structure):
    _fields_ = [
        ("bctype", C.c_int), ("subbctype", C.c_int), ("pertype", C.c_int),
        ("point_start", I3), ("point_end", I3),
        ("ijkdirect", C.c_int), ("left_or_right", C.c_int),
        ("ijkadd", I3), ("cijkadd", I3), ("cijkadd1", I3), ("cijkadd2", I3),
        ("ceijkadd", I3), ("ceijkadd1", I3), ("ceijkadd2", I3),
        ("target_block", C.c_int), ("target_conn", C.c_int),
        ("transform", I3), ("transform_t", I3), ("rev_matrix", I3), ("lcross", C.c_int),
        ("rotate_angle", C.c_double),
    ]


class TransBlockDesc(C.Structure):
    _fields_ = [("ni", C.c_int), ("nj", C.c_int), ("nk", C.c_int),
                ("npatch", C.c_int), ("patches", C.POINTER(TransPatch))]


class TransFView(C.Structure):
    _fields_ = [("p", C.POINTER(C.c_double)), ("ncomp", C.c_int), ("lo", I3), ("hi", I3)]


class TransGeomView(C.Structure):
    _fields_ = [("sd", TransFView), ("grad", TransFView), ("alagm", TransFView),
                ("vol", TransFView), ("rad", TransFView), ("dst", TransFView),
                ("aleng", TransFView), ("aseng", TransFView), ("yc", TransFView), ("zc", TransFView),
                ("thtc", TransFView),
                ("markwall", C.POINTER(C.c_int) * 6)]


class TransStateView(C.Structure):
    _fields_ = [("v", TransFView), ("pp", TransFView), ("t", TransFView), ("rmiu", TransFView),
                ("wt0", TransFView), ("wt1", TransFView), ("wt2", TransFView)]


class TransResid(C.Structure):
    _fields_ = [("drsa", C.c_double), ("dua", C.c_double),
                ("ima", C.c_int), ("jma", C.c_int), ("kma", C.c_int),
                ("drsk", C.c_double), ("duk", C.c_double),
                ("imk", C.c_int), ("jmk", C.c_int), ("kmk", C.c_int),
                ("drso", C.c_double), ("duo", C.c_double),
                ("imo", C.c_int), ("jmo", C.c_int), ("kmo", C.c_int)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class TransPlaneIntegrals(C.Structure):
    _fields_ = [("mass_flow", C.c_double), ("area", C.c_double), ("p0_mass_avg", C.c_double),
                ("t0_mass_avg", C.c_double), ("p0_abs_mass_avg", C.c_double), ("t0_abs_mass_avg", C.c_double)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


FIELD_D, FIELD_RDS, FIELD_DTM, FIELD_DQDXYZ, FIELD_SHOCK, FIELD_FSCHEME = 0, 1, 2, 3, 4, 5
TRANS_EINVAL, TRANS_EHIP, TRANS_ENOMEM, TRANS_EUNSUPPORTED, TRANS_ESCRATCH, TRANS_ETIMEOUT = -1, -2, -3, -4, -5, -6
TRANS_ENONFINITE = 1  # a residual norm came out NaN (the reference's STOP)
TRANS_WD_STATIC_DOUBLES = 4 << 20


class TransSrvConfig(C.Structure):
    _fields_ = [("hbm_budget", C.c_size_t), ("max_active", C.c_int), ("chunk", C.c_int)]


class TransSrvRequest(C.Structure):
    _fields_ = [("params", TransParams), ("nblk", C.c_int), ("blocks", C.POINTER(TransBlockDesc)),
                ("geom", C.POINTER(TransGeomView)), ("state", C.POINTER(TransStateView)),
                ("target_drop", C.c_double), ("max_iter", C.c_int)]


class TransSrvStatus(C.Structure):
    _fields_ = [("state", C.c_int), ("error", C.c_int), ("iterations", C.c_int),
                ("iterations_to_target", C.c_int), ("drsa_first", C.c_double),
                ("drsa_last", C.c_double), ("t_submit", C.c_double), ("t_admit", C.c_double),
                ("t_done", C.c_double), ("footprint", C.c_size_t)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


SRV_QUEUED, SRV_RUNNING, SRV_CONVERGED, SRV_MAX_ITER, SRV_CANCELLED, SRV_FAILED = range(6)

class TransGeomMetricsOut(C.Structure):
    _fields_ = [(nm, C.POINTER(C.c_double)) for nm in
                ("xc", "yc", "zc", "thtc", "rad", "aleng", "aseng", "vol", "sd", "grad", "alagm")]


class TransTuneEntry(C.Structure):
    _fields_ = [("sweep", C.c_int), ("streams", C.c_int), ("graph", C.c_int), ("ms_per_iter", C.c_double),
                ("tile_a", C.c_int), ("tile_b", C.c_int), ("chunk", C.c_int)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


NSTAT = 23
STAT_NAMES = ("U", "V", "W", "R", "M", "Mu", "k", "P0", "P", "T", "uu", "vv", "ww", "uv", "uw", "vw",
              "uurms", "vvrms", "wwrms", "uvrms", "vwrms", "uwrms", "pprms")

EXPORTS = [
    "trans_hip_create", "trans_hip_destroy", "trans_hip_last_error", "trans_hip_set_params",
    "trans_hip_set_iteration", "trans_hip_set_geometry", "trans_hip_set_state",
    "trans_hip_get_state", "trans_hip_get_field", "trans_hip_spectral_radius", "trans_hip_bc",
    "trans_hip_space", "trans_hip_residual", "trans_hip_lusgs_update", "trans_hip_rk_update",
    "trans_hip_update_u",
    "trans_hip_iterate", "trans_hip_plane_integrals", "trans_hip_sync", "trans_hip_profile", "trans_hip_profile_count",
    "trans_hip_profile_get", "trans_hip_footprint", "trans_hip_device_bytes",
    "trans_hip_write_restart", "trans_hip_read_restart", "trans_hip_stats_enable", "trans_hip_average_fields",
    "trans_hip_get_stats", "trans_hip_set_stats", "trans_hip_write_stats", "trans_hip_read_stats",
    "trans_hip_autotune", "trans_hip_get_tuning", "trans_hip_set_tuning",
    "trans_srv_create", "trans_srv_submit", "trans_srv_progress", "trans_srv_poll", "trans_srv_cancel",
    "trans_srv_result", "trans_srv_release", "trans_srv_destroy", "trans_srv_last_error",
    "trans_geom_wall_distance", "trans_geom_metrics", "trans_geom_wall_distance_static",
    "trans_hip_write_tecplot", "trans_tecplot_write_ascii",
]


def fview(arr: np.ndarray, lo=(-2, -2, -2)) -> TransFView:
    """Describe a Fortran-ordered array A(ncomp, lo:hi, ...) (or A(lo:hi, ...))."""
    assert arr.dtype == np.float64 and arr.flags.f_contiguous
    if arr.ndim == 4:
        ncomp = arr.shape[0]
        shp = arr.shape[1:]
    else:
        ncomp = 1
        shp = arr.shape
    v = TransFView()
    v.p = arr.ctypes.data_as(C.POINTER(C.c_double))
    v.ncomp = ncomp
    v.lo = I3(*lo)
    v.hi = I3(*(lo[d] + shp[d] - 1 for d in range(3)))
    return v


_lib = None


def load_library(path: str = LIB_PATH) -> C.CDLL:
    """Load the HIP library; raises if it is missing (there is no CPU fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
    lib = C.CDLL(path)
    P = C.POINTER
    lib.trans_hip_create.argtypes = [P(TransParams), C.c_int, P(TransBlockDesc), P(C.c_void_p)]
    lib.trans_hip_destroy.argtypes = [C.c_void_p]
    lib.trans_hip_last_error.argtypes = [C.c_void_p]
    lib.trans_hip_last_error.restype = C.c_char_p
    lib.trans_hip_set_params.argtypes = [C.c_void_p, P(TransParams)]
    lib.trans_hip_set_iteration.argtypes = [C.c_void_p, C.c_int]
    lib.trans_hip_set_geometry.argtypes = [C.c_void_p, C.c_int, P(TransGeomView)]
    lib.trans_hip_set_state.argtypes = [C.c_void_p, C.c_int, P(TransStateView)]
    lib.trans_hip_get_state.argtypes = [C.c_void_p, C.c_int, P(TransStateView)]
    lib.trans_hip_get_field.argtypes = [C.c_void_p, C.c_int, C.c_int, P(TransFView)]
    lib.trans_hip_spectral_radius.argtypes = [C.c_void_p]
    lib.trans_hip_bc.argtypes = [C.c_void_p, C.c_int]
    lib.trans_hip_space.argtypes = [C.c_void_p]
    lib.trans_hip_residual.argtypes = [C.c_void_p]
    lib.trans_hip_lusgs_update.argtypes = [C.c_void_p, P(TransResid)]
    lib.trans_hip_rk_update.argtypes = [C.c_void_p, C.c_int, P(TransResid)]
    lib.trans_hip_plane_integrals.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, P(TransPlaneIntegrals)]
    lib.trans_hip_update_u.argtypes = [C.c_void_p]
    lib.trans_hip_iterate.argtypes = [C.c_void_p, C.c_int, P(TransResid)]
    lib.trans_hip_sync.argtypes = [C.c_void_p]
    lib.trans_hip_profile.argtypes = [C.c_void_p, C.c_int]
    lib.trans_hip_profile_count.argtypes = [C.c_void_p]
    lib.trans_hip_profile_get.argtypes = [C.c_void_p, C.c_int, C.c_char_p, C.c_int,
                                          C.POINTER(C.c_double), C.POINTER(C.c_long)]
    lib.trans_hip_footprint.argtypes = [P(TransParams), C.c_int, P(TransBlockDesc), P(C.c_size_t)]
    lib.trans_hip_write_restart.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_char_p]
    lib.trans_hip_read_restart.argtypes = [C.c_void_p, C.c_int, C.c_char_p, P(C.c_int)]
    lib.trans_hip_stats_enable.argtypes = [C.c_void_p, C.c_int, C.c_int]
    lib.trans_hip_average_fields.argtypes = [C.c_void_p]
    lib.trans_hip_get_stats.argtypes = [C.c_void_p, C.c_int, P(TransFView)]
    lib.trans_hip_set_stats.argtypes = [C.c_void_p, C.c_int, P(TransFView)]
    lib.trans_hip_write_stats.argtypes = [C.c_void_p, C.c_int, P(TransFView), P(TransFView), P(TransFView),
                                          C.c_char_p]
    lib.trans_hip_read_stats.argtypes = [C.c_void_p, C.c_int, C.c_char_p]
    lib.trans_hip_write_tecplot.argtypes = [C.c_void_p, C.c_int, P(TransFView), P(TransFView), P(TransFView),
                                            P(C.c_double), C.c_char_p]
    lib.trans_tecplot_write_ascii.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_int, C.c_int, C.c_int,
                                              P(C.c_double), P(C.c_double), P(C.c_double), C.c_int,
                                              P(C.c_char_p), P(C.c_double)]
    lib.trans_hip_autotune.argtypes = [C.c_void_p, C.c_int, P(TransTuneEntry), C.c_int, P(C.c_int)]
    lib.trans_hip_get_tuning.argtypes = [C.c_void_p, P(TransTuneEntry)]
    lib.trans_hip_set_tuning.argtypes = [C.c_void_p, P(TransTuneEntry)]
    lib.trans_hip_device_bytes.argtypes = [C.c_void_p, P(C.c_size_t)]
    lib.trans_geom_wall_distance.argtypes = [P(C.c_double), C.c_long, P(C.c_double), C.c_long, P(C.c_double)]
    lib.trans_geom_wall_distance_static.argtypes = [P(C.c_double), C.c_long, P(C.c_double), C.c_long,
                                                    P(C.c_double)]
    lib.trans_geom_metrics.argtypes = [C.c_int, C.c_int, C.c_int, P(C.c_double), P(C.c_double), P(C.c_double),
                                       C.c_double, P(TransGeomMetricsOut)]
    lib.trans_srv_create.argtypes = [P(TransSrvConfig), P(C.c_void_p)]
    lib.trans_srv_submit.argtypes = [C.c_void_p, P(TransSrvRequest), P(C.c_int)]
    lib.trans_srv_progress.argtypes = [C.c_void_p, C.c_int, P(C.c_int)]
    lib.trans_srv_poll.argtypes = [C.c_void_p, C.c_int, P(TransSrvStatus)]
    lib.trans_srv_cancel.argtypes = [C.c_void_p, C.c_int]
    lib.trans_srv_result.argtypes = [C.c_void_p, C.c_int, C.c_int, P(TransStateView)]
    lib.trans_srv_release.argtypes = [C.c_void_p, C.c_int]
    lib.trans_srv_destroy.argtypes = [C.c_void_p]
    lib.trans_srv_last_error.argtypes = [C.c_void_p]
    lib.trans_srv_last_error.restype = C.c_char_p
    for name in EXPORTS:
        if name not in ("trans_hip_last_error", "trans_srv_last_error"):
            getattr(lib, name).restype = C.c_int
    _lib = lib
    return lib
