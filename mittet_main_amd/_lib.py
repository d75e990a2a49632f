"""ctypes binding of libmittet_fdtd.so (include/mittet_fdtd.h).

The library must exist (built by ``mittet_main_amd.build``); there is no
fallback: a missing or unloadable library raises immediately.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# MITTET_FDTD_LIB selects another build of the same library (A/B comparisons)
LIB_PATH = os.environ.get("MITTET_FDTD_LIB") or os.path.join(HERE, "libmittet_fdtd.so")

FWI3D_OK = 0
ERRORS = {-1: "EINVAL", -2: "ENOMEM", -3: "EDEVICE", -4: "ESTATE", -5: "EBUDGET"}
NPROFILES = 12


class Fwi3dError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class GridParams(C.Structure):
    _fields_ = [("ix", C.c_int), ("iy", C.c_int), ("iz", C.c_int), ("ncpml", C.c_int),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
                ("dt", C.c_double), ("omega0", C.c_double), ("mu0", C.c_double),
                ("c1", C.c_double), ("c2", C.c_double), ("cmax", C.c_double),
                ("rcoef", C.c_double), ("nn", C.c_double), ("order", C.c_double),
                ("ma", C.c_double), ("kappa_max", C.c_double), ("alpha_max", C.c_double),
                ("opt_to_max", C.c_double * 3)]


class ShotParams(C.Structure):
    _fields_ = [("it", C.c_int), ("src_comp", C.c_int),
                ("src_i", C.c_int), ("src_j", C.c_int), ("src_k", C.c_int),
                ("signal", C.POINTER(C.c_double)),
                ("nrec", C.c_int), ("rec_ijk", C.POINTER(C.c_int)),
                ("nf_vol", C.c_int), ("omega_vol", C.POINTER(C.c_double)),
                ("ring", C.c_int), ("zchunk", C.c_int)]


_SIGS = {
    "fwi3d_last_error": (C.c_char_p, []),
    "fwi3d_version": (C.c_char_p, []),
    "fwi3d_lattice_time_3d": (C.c_double, [C.c_double] * 7 + [C.POINTER(C.c_double)]),
    "fwi3d_fmax_check": (C.c_int, [C.c_double] * 8 + [C.POINTER(C.c_double)]),
    "fwi3d_stable_dt": (C.c_double, [C.POINTER(GridParams), C.c_int, C.c_void_p, C.c_void_p]),
    "fwi3d_model_create": (C.c_int, [C.POINTER(GridParams), C.c_void_p, C.POINTER(C.c_void_p)]),
    "fwi3d_model_destroy": (C.c_int, [C.c_void_p]),
    "fwi3d_model_get_profile": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p]),
    "fwi3d_media_coeff_3d": (C.c_int, [C.c_void_p, C.c_void_p]),
    "fwi3d_shot_create": (C.c_int, [C.c_void_p, C.POINTER(ShotParams), C.POINTER(C.c_void_p)]),
    "fwi3d_shot_destroy": (C.c_int, [C.c_void_p]),
    "fwi3d_shot_bytes": (C.c_longlong, [C.POINTER(GridParams), C.POINTER(ShotParams)]),
    "fwi3d_volume_dft_h_bytes": (C.c_longlong, [C.POINTER(GridParams), C.POINTER(ShotParams)]),
    "fwi3d_set_zero_eh": (C.c_int, [C.c_void_p]),
    "fwi3d_e_field_cpml42": (C.c_int, [C.c_void_p, C.c_int]),
    "fwi3d_epml_abcs4": (C.c_int, [C.c_void_p]),
    "fwi3d_read_source_3d": (C.c_int, [C.c_void_p, C.c_int]),
    "fwi3d_h_field_cpml42": (C.c_int, [C.c_void_p, C.c_int]),
    "fwi3d_hpml_abcs4": (C.c_int, [C.c_void_p]),
    "fwi3d_output": (C.c_int, [C.c_void_p, C.c_int]),
    "fwi3d_step_n": (C.c_int, [C.c_void_p, C.c_int]),
    "fwi3d_run": (C.c_int, [C.c_void_p]),
    "fwi3d_shot_sync": (C.c_int, [C.c_void_p]),
    "fwi3d_profile_steps": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
    "fwi3d_shot_step": (C.c_int, [C.c_void_p]),
    "fwi3d_shot_diverged": (C.c_int, [C.c_void_p]),
    "fwi3d_diag_copy": (C.c_int, [C.c_longlong, C.c_int, C.c_void_p]),
    "fwi3d_get_field": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
    "fwi3d_set_field": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
    "fwi3d_get_traces": (C.c_int, [C.c_void_p, C.c_void_p]),
    "fwi3d_get_volume_dft": (C.c_int, [C.c_void_p, C.c_void_p]),
    "fwi3d_enable_volume_dft_h": (C.c_int, [C.c_void_p]),
    "fwi3d_get_volume_dft_h": (C.c_int, [C.c_void_p, C.c_void_p]),
    "fwi3d_receiver_dft": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p]),
    "fwi3d_receiver_dft_noalloc": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_int,
                                             C.c_void_p]),
    "fwi3d_f_to_d_h": (C.c_int, [C.c_double, C.c_int, C.c_void_p, C.c_longlong, C.c_void_p]),
    "fwi3d_source_spectrum": (C.c_int, [C.c_double, C.c_double, C.c_int, C.c_void_p, C.c_int,
                                        C.c_void_p, C.c_void_p]),
    "fwi3d_laplace_to_freq": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                        C.c_void_p, C.c_void_p]),
    "fwi3d_model_create_variant": (C.c_int, [C.POINTER(GridParams), C.c_int, C.c_void_p,
                                             C.POINTER(C.c_void_p)]),
    "fwi3d_model_set_mu": (C.c_int, [C.c_void_p, C.c_void_p]),
    "fwi3d_shot_add_source": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int]),
    "fwi3d_dft_rows": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_void_p]),
    "fwi3d_laplace_fft": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "fwi3d_convolve_gj": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p]),
    "fwi3d_grad_create": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_void_p)]),
    "fwi3d_grad_create_levels": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_void_p)]),
    "fwi3d_grad_plan_bytes": (C.c_longlong, [C.c_void_p, C.c_int]),
    "fwi3d_grad_plan": (C.c_int, [C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                  C.POINTER(C.c_longlong)]),
    "fwi3d_grad_destroy": (C.c_int, [C.c_void_p]),
    "fwi3d_grad_bytes": (C.c_longlong, [C.c_void_p]),
    "fwi3d_grad_forward": (C.c_int, [C.c_void_p]),
    "fwi3d_grad_backward": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "fwi3d_compat_create": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_int,
                                      C.c_int, C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p)]),
    "fwi3d_compat_destroy": (C.c_int, [C.c_void_p]),
    "fwi3d_compat_omegas": (C.c_int, [C.c_void_p, C.c_void_p]),
    "fwi3d_compat_reset": (C.c_int, [C.c_void_p]),
    "fwi3d_compat_source": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_void_p]),
    "fwi3d_compat_green": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p]),
    "fwi3d_compat_gamma": (C.c_int, [C.c_void_p, C.c_void_p]),
    "fwi3d_compat_phi": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
    "fwi3d_compat_receivers": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "fwi3d_compat_alpha": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]),
    "fwi3d_compat_update": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_double,
                                      C.c_int, C.c_void_p, C.c_void_p]),
    "fwi3d_batch_create": (C.c_int, [C.c_int, C.c_int, C.c_void_p, C.POINTER(C.c_void_p)]),
    "fwi3d_batch_destroy": (C.c_int, [C.c_void_p]),
    "fwi3d_batch_add": (C.c_int, [C.c_void_p, C.c_void_p]),
    "fwi3d_batch_remove": (C.c_int, [C.c_void_p, C.c_void_p]),
    "fwi3d_batch_step": (C.c_int, [C.c_void_p, C.c_int]),
    "fwi3d_batch_sync": (C.c_int, [C.c_void_p]),
    "fwi3d_batch_size": (C.c_int, [C.c_void_p]),
}

_LIB = None


def load() -> C.CDLL:
    """Loads the HIP library; raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m mittet_main_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def exported_symbols():
    return list(_SIGS)


def check(rc: int) -> None:
    if rc != FWI3D_OK:
        msg = load().fwi3d_last_error()
        raise Fwi3dError(rc, msg.decode() if msg else "")
