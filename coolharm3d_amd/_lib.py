"""ctypes binding of the C ABI in include/hgpu.h.

The product path loads ``libhgpu.so`` (HIP kernels) and raises if it is missing
or if no GPU is present; there is no fallback.  ``load(hostcheck=True)`` loads the
host build of the same kernel bodies, which only tests select explicitly.
"""
from __future__ import annotations

import ctypes
import os

PKG = os.path.dirname(os.path.abspath(__file__))
GPU_LIB = os.path.join(PKG, "libhgpu.so")
HOSTCHECK_LIB = os.path.join(PKG, "libhgpu_hostcheck.so")

c_double_p = ctypes.POINTER(ctypes.c_double)
c_int_p = ctypes.POINTER(ctypes.c_int)
c_ll_p = ctypes.POINTER(ctypes.c_longlong)
_vp, _ll, _d = ctypes.c_void_p, ctypes.c_longlong, ctypes.c_double


class HgpuConfig(ctypes.Structure):
    _fields_ = [
        ("n1", ctypes.c_int), ("n2", ctypes.c_int), ("n3", ctypes.c_int),
        ("metric", ctypes.c_int), ("coords", ctypes.c_int), ("recon", ctypes.c_int),
        ("flux", ctypes.c_int), ("ct_para", ctypes.c_int), ("cooling", ctypes.c_int),
        ("rescale_r", ctypes.c_int), ("bc", ctypes.c_int), ("track_nfail", ctypes.c_int),
        ("startx", ctypes.c_double * 4), ("dx", ctypes.c_double * 4),
        ("gam", ctypes.c_double), ("cour", ctypes.c_double), ("t_final", ctypes.c_double),
        ("a", ctypes.c_double), ("R0", ctypes.c_double), ("h_slope", ctypes.c_double),
        ("th_beg", ctypes.c_double), ("th_length", ctypes.c_double),
        ("r_isco", ctypes.c_double), ("h_o_r", ctypes.c_double),
        ("rhomin", ctypes.c_double), ("uumin", ctypes.c_double),
        ("rhopower", ctypes.c_double), ("uupower", ctypes.c_double),
        ("precision", ctypes.c_int),
        ("bbh_m1", ctypes.c_double), ("bbh_m2", ctypes.c_double), ("bbh_sep", ctypes.c_double),
        ("bbh_phi0", ctypes.c_double), ("bbh_inspiral", ctypes.c_int),
        ("bc_nbr", ctypes.c_int), ("gpos", ctypes.c_int * 3), ("no_confirm_iters", ctypes.c_int),
        ("diag3_exponent", ctypes.c_int), ("diag3_factor", ctypes.c_double),
    ]


class HgpuPnConfig(ctypes.Structure):
    _fields_ = [
        ("m1", ctypes.c_double), ("m2", ctypes.c_double), ("b0", ctypes.c_double), ("phi0", ctypes.c_double),
        ("inspiral", ctypes.c_int), ("R0", ctypes.c_double), ("th_beg", ctypes.c_double),
        ("th_length", ctypes.c_double), ("h_slope", ctypes.c_double),
        ("n1", ctypes.c_int), ("n2", ctypes.c_int), ("n3", ctypes.c_int),
        ("startx", ctypes.c_double * 4), ("dx", ctypes.c_double * 4),
        ("coords", ctypes.c_int), ("diag3_exponent", ctypes.c_int), ("diag3_factor", ctypes.c_double),
    ]


# exported symbols and their signatures (restype, argtypes)
SIGNATURES = {
    "hgpu_last_error": (ctypes.c_char_p, []),
    "hgpu_create": (ctypes.c_int, [ctypes.POINTER(HgpuConfig), ctypes.POINTER(ctypes.c_void_p)]),
    "hgpu_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "hgpu_dims": (ctypes.c_int, [ctypes.c_void_p, c_int_p]),
    "hgpu_upload_prims": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, c_double_p]),
    "hgpu_download_prims": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, c_double_p]),
    "hgpu_geometry_counts": (ctypes.c_int, [ctypes.c_void_p, c_ll_p, c_ll_p]),
    "hgpu_upload_geometry": (ctypes.c_int, [ctypes.c_void_p, c_double_p, c_double_p]),
    "hgpu_download_geometry": (ctypes.c_int, [ctypes.c_void_p, c_double_p, c_double_p]),
    "hgpu_set_time": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_double]),
    "hgpu_get_time": (ctypes.c_int, [ctypes.c_void_p, c_double_p, c_double_p, c_double_p]),
    "hgpu_fixup_bounds": (ctypes.c_int, [ctypes.c_void_p]),
    "hgpu_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "hgpu_substep": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "hgpu_advance": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "hgpu_get_boundary_flags": (ctypes.c_int, [ctypes.c_void_p, c_int_p, c_int_p, c_int_p]),
    "hgpu_fixup_stage": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "hgpu_end_step": (ctypes.c_int, [ctypes.c_void_p, c_double_p]),
    "hgpu_set_dt": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double]),
    "hgpu_download_box": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, c_int_p, c_int_p, c_double_p, c_int_p]),
    "hgpu_upload_box": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, c_int_p, c_int_p, c_double_p, c_int_p]),
    "hgpu_set_graph": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "hgpu_set_metric_eval": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "hgpu_get_metric_eval": (ctypes.c_int, [ctypes.c_void_p, c_int_p]),
    "hgpu_device_bytes": (ctypes.c_int, [ctypes.c_void_p, c_ll_p]),
    "hgpu_get_m_tot": (ctypes.c_int, [ctypes.c_void_p, c_double_p]),
    "hgpu_autotune": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, c_double_p]),
    "hgpu_set_launch_shape": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "hgpu_get_launch_shape": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, c_int_p, c_int_p, c_int_p]),
    "hgpu_download_pflag": (ctypes.c_int, [ctypes.c_void_p, c_int_p]),
    "hgpu_download_gamma": (ctypes.c_int, [ctypes.c_void_p, c_double_p]),
    "hgpu_download_nfail": (ctypes.c_int, [ctypes.c_void_p, c_int_p]),
    "hgpu_get_nfailtot": (ctypes.c_int, [ctypes.c_void_p, c_ll_p]),
    "hgpu_reset_nfail": (ctypes.c_int, [ctypes.c_void_p]),
    "hgpu_get_status": (ctypes.c_int, [ctypes.c_void_p, c_int_p, c_double_p]),
    "hgpu_get_crash": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_longlong),
                                      ctypes.POINTER(ctypes.c_longlong)]),
    "hgpu_download_fluxes": (ctypes.c_int, [ctypes.c_void_p, c_double_p]),
    "hgpu_divb_max": (ctypes.c_int, [ctypes.c_void_p, c_double_p]),
    "hgpu_sync": (ctypes.c_int, [ctypes.c_void_p]),
    "hgpu_profile_enable": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "hgpu_profile_read": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, c_double_p, c_ll_p]),
    "hgpu_con2prim_batch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_longlong, ctypes.c_void_p, ctypes.c_void_p,
                                            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    # device plumbing
    "hgpu_dev_count": (ctypes.c_int, [c_int_p]),
    "hgpu_dev_alloc": (ctypes.c_int, [_ll, ctypes.POINTER(ctypes.c_void_p)]),
    "hgpu_dev_free": (ctypes.c_int, [_vp]),
    "hgpu_dev_copy": (ctypes.c_int, [_vp, _vp, _ll, ctypes.c_int, _vp]),
    "hgpu_dev_copy_kernel": (ctypes.c_int, [_vp, _vp, _ll, _vp]),
    "hgpu_stream_sync": (ctypes.c_int, [_vp]),
    "hgpu_event_create": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p)]),
    "hgpu_event_destroy": (ctypes.c_int, [_vp]),
    "hgpu_event_record": (ctypes.c_int, [_vp, _vp]),
    "hgpu_event_elapsed": (ctypes.c_int, [_vp, _vp, ctypes.POINTER(ctypes.c_float)]),
    # binary-black-hole metric
    "hgpu_pn_geometry": (ctypes.c_int, [_vp, ctypes.POINTER(HgpuPnConfig), _d, ctypes.c_int, _vp]),
    "hgpu_pn_connection": (ctypes.c_int, [_vp, ctypes.POINTER(HgpuPnConfig), _d, _vp]),
    # batched con-to-prim and its network (pointers are addresses: ints)
    "hgpu_c2p_samples": (ctypes.c_int, [_vp, _ll, ctypes.c_ulonglong, _d, _d, _vp, _vp, _vp, _vp]),
    "hgpu_c2p_workspace_bytes": (_ll, [_ll, ctypes.c_int]),
    "hgpu_c2p_solve": (ctypes.c_int, [_vp, _ll, _d, _d, _vp, _vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int, _vp, _ll, _vp, _vp, _vp]),
    "hgpu_mlp_sizes": (ctypes.c_int, [c_int_p]),
    "hgpu_mlp_features": (ctypes.c_int, [_vp, _ll, _d, _d, _vp, _vp, _vp, _vp, _vp]),
    "hgpu_mlp_forward": (ctypes.c_int, [_vp, _ll, _vp, _vp, _vp]),
    "hgpu_mlp_workspace_bytes": (_ll, [_ll]),
    "hgpu_mlp_grad": (ctypes.c_int, [_vp, _ll, _vp, _vp, _vp, _vp, _ll, _vp, _vp]),
    "hgpu_mlp_init": (ctypes.c_int, [_vp, ctypes.c_ulonglong, _ll, _vp, _vp, _vp]),
    "hgpu_mlp_train": (ctypes.c_int, [_vp, _ll, _vp, _vp, ctypes.c_ulonglong, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_float, _vp, _ll, _vp, _vp]),
}

_cache = {}


class HgpuError(RuntimeError):
    pass


def load(hostcheck: bool = False):
    path = HOSTCHECK_LIB if hostcheck else GPU_LIB
    if path in _cache:
        return _cache[path]
    if not os.path.exists(path):
        raise HgpuError(f"{os.path.basename(path)} is not built; run `python -m coolharm3d_amd.build` "
                        "(or __graft_entry__.build()) first")
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _cache[path] = lib
    return lib


def check(lib, rc, what):
    if rc != 0:
        msg = lib.hgpu_last_error().decode(errors="replace")
        raise HgpuError(f"{what} failed ({rc}): {msg}")
