"""ctypes binding of libffa_hip.so (C ABI in include/ffa_hip.h).

The structures below mirror the header field by field.  Loading fails loudly
when the library has not been built: there is no CPU fallback for the
tracking hot path.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libffa_hip.so")

FFA_MAX_SEG = 8
FFA_MAX_POLY = 8
FFA_MAX_HARM = 8
FFA_MAX_GAP = 8
FFA_SC_BLOCKS = 1024
SC_NOALLOC_BYTES = 256 << 20

F_INJ = 1
F_ALIVE = 2
F_SCPRE = 4
F_SCPREPRE = 8

_d = ctypes.c_double
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_vp = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)


class FfaParticles(ctypes.Structure):
    _fields_ = [("r", _vp), ("vr", _vp), ("z", _vp), ("vz", _vp), ("fi", _vp), ("fidot", _vp),
                ("rf_phase", _vp), ("flags", _vp), ("esc_pre", _vp), ("esc_prepre", _vp),
                ("n", _i64), ("esc_f32", _i32)]


class FfaMachine(ctypes.Structure):
    _fields_ = [
        ("taylor", _vp), ("taylor_f32", _i32), ("max_order", _i32), ("r_size", _i32), ("fi_size", _i32),
        ("r_min", _d), ("r_step", _d), ("fi_min", _d), ("fi_step", _d),
        ("B0", _d), ("k_value", _d), ("nsectors", _i32), ("n_seg", _i32),
        ("convex", (_d * 5) * FFA_MAX_SEG),
        ("n_spoly", _i32), ("n_sppoly", _i32),
        ("s_poly", _d * FFA_MAX_POLY), ("sp_poly", _d * FFA_MAX_POLY),
        ("n_harm", _i32),
        ("harm_m", _d * FFA_MAX_HARM), ("harm_amp", _d * FFA_MAX_HARM), ("harm_phase", _d * FFA_MAX_HARM),
        ("harm_rmin", _d * FFA_MAX_HARM), ("harm_rmax", _d * FFA_MAX_HARM),
        ("rf_enable", _i32), ("n_gap", _i32), ("acc_etot", _d),
        ("gap_theta", _d * FFA_MAX_GAP),
        ("gap_px", (_d * 4) * FFA_MAX_GAP), ("gap_py", (_d * 4) * FFA_MAX_GAP),
        ("gap_box", (_d * 4) * FFA_MAX_GAP), ("rf_shift", _d * FFA_MAX_GAP),
        ("gap_model", _i32), ("emap_nr", _i32), ("emap_nphi", _i32),
        ("emap_r0", _d), ("emap_dr", _d), ("emap_phi0", _d), ("emap_dphi", _d), ("emap", _vp),
        ("aperture_enable", _i32), ("aperture", _d * 4),
        ("q", _d), ("c", _d), ("m0", _d), ("E0", _d),
    ]


class FfaStep(ctypes.Structure):
    _fields_ = [("t", _d), ("dt", _d), ("dRF", _d), ("v0", _d), ("sc_s", _d),
                ("rk_rate", _d * 3), ("rk_v0", _d * 3)]


class FfaScParams(ctypes.Structure):
    _fields_ = [("ns", _i32), ("nx", _i32), ("nz", _i32), ("sigma_ratio", _d), ("q_macro", _d),
                ("rk", _i32), ("first", _i32), ("s_last", _d), ("bits", _i32), ("cell_counts", _i32), ("charge_grid", _i32), ("field_cubes", _i32),
                ("seo_r", _vp), ("seo_pr", _vp), ("seo_ek_axis", _vp), ("seo_perimeter", _vp),
                ("n_ek", _i32), ("n_fi", _i32), ("ek0", _d), ("dek", _d), ("fi0", _d), ("dfi", _d),
                ("E0_MeV", _d), ("c", _d)]


class FfaScResult(ctypes.Structure):
    _fields_ = [(name, _d) for name in (
        "count", "sum_gamma", "fi_s", "fi_e", "slice_step", "mean_gamma_m1",
        "sum_x", "sum_z", "sum_x2", "sum_z2", "xmin", "xmax", "zmin", "zmax", "dx", "dz",
        "bunch_length", "perimeter", "qscale", "bad_grid")] + [
        ("arc_min", _d * 4), ("arc_max", _d * 4), ("pad", _d * 4)]


class FfaScLayout(ctypes.Structure):
    _fields_ = [(name, ctypes.c_size_t) for name in (
        "total", "grid", "cell_count", "cbuf", "phi", "exz", "px", "ppr0", "pslice",
        "slice_count", "partials", "result", "records", "slice_start", "block_offset", "records_by_row", "row_start",
        "tables")]


MLP_NFREQ = 32
MLP_NPARAM = 58243
MLP_PACKED_BYTES = 236160
MLP_KCHUNK = 512


class FfaMlpDesc(ctypes.Structure):
    _fields_ = [("r_c", _d), ("inv_rh", _d), ("inv_sector", _d), ("zt", _d), ("out_scale", _d * 3), ("out_delta", _d),
                ("m", ctypes.c_float * MLP_NFREQ), ("wa", ctypes.c_float * MLP_NFREQ),
                ("wc", ctypes.c_float * MLP_NFREQ), ("fallback", _i32), ("pad", _i32)]


class FfaAdam(ctypes.Structure):
    _fields_ = [("lr", ctypes.c_float), ("beta1", ctypes.c_float), ("beta2", ctypes.c_float),
                ("eps", ctypes.c_float), ("step", _i32), ("pad", _i32)]


class FfaMlpLayout(ctypes.Structure):
    _fields_ = [("total", ctypes.c_size_t), ("batch_padded", _i64), ("xfm", ctypes.c_size_t),
                ("dzfm", ctypes.c_size_t), ("zs", ctypes.c_size_t), ("partials", ctypes.c_size_t),
                ("loss", ctypes.c_size_t)]


class FfaTaylorGen(ctypes.Structure):
    _fields_ = [("n_orders", _i32), ("n_q", _i32), ("n_poles", _i32), ("layout", _i32), ("out_f32", _i32),
                ("pad", _i32), ("nr", _i64), ("n_phi", _i64), ("theta", _vp), ("poles", _vp), ("sig_poly", _vp),
                ("weights", _vp), ("shift", _vp), ("flutter", _vp), ("out", _vp)]


class FfaOrbit(ctypes.Structure):
    _fields_ = [("state", _vp), ("q_over_p", _vp), ("n", _i64), ("t0", _d), ("t_step", _d), ("h", _d),
                ("nsteps", _i32), ("pad", _i32), ("traj", _vp)]


class FfaSpiralGap(ctypes.Structure):
    _fields_ = [("r_axis", _vp), ("phi_axis", _vp), ("nr", _i32), ("nphi", _i32), ("dr", _d), ("dphi", _d),
                ("r_ref", _d), ("s_ref", _d), ("a", _d), ("r_min", _d), ("r_max", _d), ("U", _d), ("delta", _d),
                ("side", _d), ("n_iter", _i32), ("pad", _i32), ("er", _vp), ("ephi", _vp), ("ez_over_z", _vp),
                ("dist", _vp), ("flag", _vp)]


FFA_TAYLOR_MAX_ORDERS = 8
FFA_TAYLOR_MAX_POLES = 16


class FfaError(RuntimeError):
    pass


_lib = None

# every symbol include/ffa_hip.h declares
EXPORTED = ("ffa_create", "ffa_destroy", "ffa_last_error", "ffa_set_stream", "ffa_set_machine",
            "ffa_inject", "ffa_push_boris_f64", "ffa_push_rk4_f64", "ffa_push_f64_multi",
            "ffa_push_boris_f64_fused", "ffa_push_boris_fused_f32",
            "ffa_sc_workspace_layout", "ffa_sc_solve", "ffa_sc_solve_field", "ffa_push_boris_fused_f32_sc",
            "ffa_reduce_phi", "ffa_bfield_eval",
            "ffa_mlp_workspace_layout", "ffa_mlp_pack", "ffa_mlp_grad", "ffa_mlp_adam", "ffa_mlp_train_step",
            "ffa_mlp_eval", "ffa_mlp_set", "ffa_sort_workspace_size", "ffa_sort_particles", "ffa_detect_crossing",
            "ffa_taylor_maps", "ffa_orbit_track", "ffa_spiral_gap_maps",
            "ffa_sc_stage_workspace_size", "ffa_sc_local_coords", "ffa_sc_divide_arc", "ffa_sc_slices",
            "ffa_sc_deposit", "ffa_sc_green", "ffa_sc_poisson", "ffa_sc_field", "ffa_sc_total_field_1d",
            "ffa_sc_interpolate", "ffa_sc_solve_noalloc")


def load():
    """Load libffa_hip.so (raises when it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FfaError(f"{LIB_PATH} is missing: build it with `make -C pyffa_amd/csrc` "
                       "(or __graft_entry__.build()); the tracking path has no CPU fallback")
    lib = ctypes.CDLL(LIB_PATH)
    lib.ffa_create.argtypes = [ctypes.c_int, _vp, ctypes.POINTER(_vp)]
    lib.ffa_destroy.argtypes = [_vp]
    lib.ffa_last_error.argtypes = [_vp]
    lib.ffa_last_error.restype = ctypes.c_char_p
    lib.ffa_set_stream.argtypes = [_vp, _vp]
    lib.ffa_set_machine.argtypes = [_vp, ctypes.POINTER(FfaMachine)]
    lib.ffa_inject.argtypes = [_vp, ctypes.POINTER(FfaParticles), _i64, _i64]
    lib.ffa_push_boris_f64.argtypes = [_vp, ctypes.POINTER(FfaParticles), ctypes.POINTER(FfaStep)]
    lib.ffa_push_rk4_f64.argtypes = [_vp, ctypes.POINTER(FfaParticles), ctypes.POINTER(FfaStep)]
    lib.ffa_push_f64_multi.argtypes = [_vp, ctypes.POINTER(FfaParticles), ctypes.POINTER(FfaStep),
                                       ctypes.POINTER(_i64), _i32, _i32]
    lib.ffa_push_boris_fused_f32.argtypes = [_vp, ctypes.POINTER(FfaParticles), ctypes.POINTER(FfaStep),
                                             ctypes.POINTER(_i64), _i32]
    lib.ffa_push_boris_f64_fused.argtypes = [_vp, ctypes.POINTER(FfaParticles), ctypes.POINTER(FfaStep),
                                             ctypes.POINTER(_i64), _i32]
    lib.ffa_sc_workspace_layout.argtypes = [_i64, _i32, _i32, _i32, ctypes.POINTER(FfaScLayout)]
    lib.ffa_sc_solve.argtypes = [_vp, ctypes.POINTER(FfaParticles), ctypes.POINTER(FfaScParams), _vp]
    lib.ffa_sc_solve_field.argtypes = [_vp, ctypes.POINTER(FfaParticles), ctypes.POINTER(FfaScParams), _vp]
    lib.ffa_push_boris_fused_f32_sc.argtypes = [_vp, ctypes.POINTER(FfaParticles), ctypes.POINTER(FfaStep),
                                                ctypes.POINTER(_i64), _i32, ctypes.POINTER(FfaScParams), _vp]
    lib.ffa_sc_solve_noalloc.argtypes = [_vp, ctypes.POINTER(FfaParticles), ctypes.POINTER(FfaScParams)]
    lib.ffa_reduce_phi.argtypes = [_vp, ctypes.POINTER(FfaParticles), _vp]
    lib.ffa_bfield_eval.argtypes = [_vp, _vp, _vp, _vp, _i64, _vp, _i32, _d]
    lib.ffa_mlp_workspace_layout.argtypes = [_i64, ctypes.POINTER(FfaMlpLayout)]
    lib.ffa_mlp_pack.argtypes = [_vp, _vp, _vp]
    lib.ffa_mlp_grad.argtypes = [_vp, ctypes.POINTER(FfaMlpDesc), _vp, _vp, _vp, _i64, _vp, _vp, _vp]
    lib.ffa_mlp_adam.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp, ctypes.POINTER(FfaAdam)]
    lib.ffa_mlp_train_step.argtypes = [_vp, ctypes.POINTER(FfaMlpDesc), _vp, _vp, _vp, _vp, _vp, _vp, _i64, _vp,
                                       ctypes.POINTER(FfaAdam), _vp]
    lib.ffa_mlp_eval.argtypes = [_vp, ctypes.POINTER(FfaMlpDesc), _vp, _vp, _i64, _vp, _i32]
    lib.ffa_mlp_set.argtypes = [_vp, ctypes.POINTER(FfaMlpDesc), _vp]
    lib.ffa_detect_crossing.argtypes = [_vp, _vp, _vp, _vp, _i64, _d, _vp, _vp]
    lib.ffa_sort_workspace_size.argtypes = [_i64, ctypes.POINTER(ctypes.c_size_t)]
    lib.ffa_sort_particles.argtypes = [_vp, ctypes.POINTER(FfaParticles), ctypes.POINTER(FfaParticles), _i64, _i32,
                                       _vp, _vp, _vp]
    lib.ffa_taylor_maps.argtypes = [_vp, ctypes.POINTER(FfaTaylorGen)]
    lib.ffa_orbit_track.argtypes = [_vp, ctypes.POINTER(FfaOrbit)]
    lib.ffa_spiral_gap_maps.argtypes = [_vp, ctypes.POINTER(FfaSpiralGap)]
    _sz = ctypes.c_size_t
    lib.ffa_sc_stage_workspace_size.argtypes = [_i64, _i32, _i32, _i32, ctypes.POINTER(_sz)]
    lib.ffa_sc_local_coords.argtypes = [_vp, _vp, _vp, _i64, _d, _i32, _vp, _vp, _vp, _sz]
    lib.ffa_sc_divide_arc.argtypes = [_vp, _vp, _i64, _vp, _vp, _sz]
    lib.ffa_sc_slices.argtypes = [_vp, _vp, _vp, _vp, _vp, _i64, _d, _d, _i32, _vp, _vp, _vp, _vp, _i64, _vp, _vp,
                                  _vp, _vp, _sz]
    lib.ffa_sc_deposit.argtypes = [_vp, _vp, _vp, _i64, _vp, _i32, _i32, _i32, _d, _d, _d, _d, _d, _d, _i32, _vp,
                                   _vp, _sz]
    lib.ffa_sc_green.argtypes = [_vp, _i32, _i32, _d, _d, _vp, _vp, _sz]
    lib.ffa_sc_poisson.argtypes = [_vp, _vp, _i32, _i32, _i32, _d, _d, _vp, _vp, _sz]
    lib.ffa_sc_field.argtypes = [_vp, _vp, _i32, _i32, _i32, _d, _d, _d, _d, _vp, _vp, _vp, _sz]
    lib.ffa_sc_total_field_1d.argtypes = [_vp, _vp, _vp, _i32, _vp]
    lib.ffa_sc_interpolate.argtypes = [_vp, _vp, _vp, _i32, _i32, _i32, _vp, _vp, _vp, _vp, _vp, _i64, _d, _d, _d, _d,
                                       _d, _vp, _vp, _vp, _vp, _i64]
    for name in EXPORTED:
        if name != "ffa_last_error":
            getattr(lib, name).restype = ctypes.c_int
    _lib = lib
    return lib


class Context:
    """Owns one ffa context bound to a device and a HIP stream."""

    def __init__(self, device: int = 0, stream: int | None = None):
        self.lib = load()
        h = _vp()
        rc = self.lib.ffa_create(int(device), _vp(stream or 0), ctypes.byref(h))
        if rc != 0:
            raise FfaError(f"ffa_create failed ({rc}): {self.lib.ffa_last_error(None).decode()}")
        self.h = h

    def check(self, rc: int, what: str):
        if rc != 0:
            raise FfaError(f"{what} failed ({rc}): {self.lib.ffa_last_error(self.h).decode()}")

    def close(self):
        if getattr(self, "h", None):
            self.lib.ffa_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def sc_layout(n: int, ns: int, nz: int, nx: int) -> FfaScLayout:
    lib = load()
    L = FfaScLayout()
    rc = lib.ffa_sc_workspace_layout(_i64(n), _i32(ns), _i32(nz), _i32(nx), ctypes.byref(L))
    if rc != 0:
        raise FfaError(f"ffa_sc_workspace_layout failed ({rc}): {lib.ffa_last_error(None).decode()}")
    return L


def mlp_layout(n: int) -> FfaMlpLayout:
    lib = load()
    L = FfaMlpLayout()
    rc = lib.ffa_mlp_workspace_layout(_i64(n), ctypes.byref(L))
    if rc != 0:
        raise FfaError(f"ffa_mlp_workspace_layout failed ({rc}): {lib.ffa_last_error(None).decode()}")
    return L
