"""ctypes binding of libqbx_mi355x.so (the C ABI declared in include/qbx_mi355x.h).

The library is built in-tree by ``__graft_entry__.build()`` (or ``make -C
pytential_amd/csrc``).  There is no fallback: if the shared object is missing
or fails to load, every entry point raises.
"""
from __future__ import annotations

import ctypes as ct
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libqbx_mi355x.so")

QBX_OK = 0
QBX_ERR_TARGET_ASSOC_FAILED = 3
LAPLACE2D, HELMHOLTZ2D, LAPLACE3D, HELMHOLTZ3D = 0, 1, 2, 3
TSQBX_S, TSQBX_D, TSQBX_SP = 0, 1, 2
SOURCE_S, SOURCE_D = 0, 1
# QBX_TUNE_DRIVE_STREAMS (include/qbx_mi355x.h): 1 = qbx_drive_fmm on the caller's stream only
TUNE_DRIVE_STREAMS = 5

P = ct.c_void_p
I = ct.c_int
D = ct.c_double
L = ct.c_long
IP = ct.POINTER(ct.c_int)


class KernelDesc(ct.Structure):
    _fields_ = [("kernel", I), ("k", D), ("source_kind", I), ("qbx_order", I),
                ("nlevels", I), ("level_orders", IP), ("nrhs", I),
                ("target_deriv", I), ("target_dir", P), ("target_specific_qbx", I),
                ("real_strengths", I)]


class QBXError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"libqbx_mi355x error {code}: {msg}")
        self.code = code


_lib = None

_SIGS = {
    "qbx_ctx_create": [ct.POINTER(P), I],
    "qbx_ctx_destroy": [P],
    "qbx_ctx_release_graphs": [P],
    "qbx_ctx_set_tuning": [P, I, I],
    "qbx_ctx_get_tuning": [P, I, IP],
    "qbx_version": [],
    "qbx_tree_build": [P, P, I, I, P, I, P, P, P, P, I, D, I, ct.POINTER(P)],
    "qbx_tree_destroy": [P],
    "qbx_tree_sizes": [P, IP, ct.POINTER(D)],
    "qbx_tree_export": [P, ct.c_char_p, P, L, P],
    "qbx_tree_field_bytes": [P, ct.c_char_p],
    "qbx_traverse": [P, P, P, I, I, I, ct.POINTER(P)],
    "qbx_trav_destroy": [P],
    "qbx_trav_sizes": [P, IP],
    "qbx_trav_export": [P, ct.c_char_p, P, L, P],
    "qbx_trav_field_bytes": [P, ct.c_char_p],
    "qbx_target_assoc": [P, P, I, I, P, P, I, P, P, I, P, P, P, P, IP],
    "qbx_refine_flags_for_failed": [P, P, I, I, P, P, I, P, I, P, P, P],
    "qbx_geo_build": [P, P, P, P, I, P, P, P, ct.POINTER(P)],
    "qbx_geo_destroy": [P],
    "qbx_geo_sizes": [P, IP],
    "qbx_geo_export": [P, ct.c_char_p, P, L, P],
    "qbx_geo_field_bytes": [P, ct.c_char_p],
    "qbx_expansion_size": [ct.POINTER(KernelDesc), P],
    "qbx_expansion_level_offsets": [ct.POINTER(KernelDesc), P, ct.POINTER(L)],
    "qbx_ncoeff_qbx": [ct.POINTER(KernelDesc)],
    "qbx_reorder_sources": [P, P, P, I, P, P, P, P, P],
    "qbx_upsample": [P, P, I, I, I, I, P, P, P],
    "qbx_form_multipoles": [P, P, P, P, ct.POINTER(KernelDesc), P, P, P],
    "qbx_coarsen_multipoles": [P, P, P, P, ct.POINTER(KernelDesc), P],
    "qbx_eval_direct": [P, P, P, P, P, ct.POINTER(KernelDesc), P, P, P, I],
    "qbx_multipole_to_local": [P, P, P, P, ct.POINTER(KernelDesc), P, P],
    "qbx_eval_multipoles": [P, P, P, P, P, ct.POINTER(KernelDesc), P, P, I],
    "qbx_form_locals": [P, P, P, P, ct.POINTER(KernelDesc), P, P, P, I],
    "qbx_refine_locals": [P, P, P, P, ct.POINTER(KernelDesc), P],
    "qbx_eval_locals": [P, P, P, P, P, ct.POINTER(KernelDesc), P, P, I],
    "qbx_form_global_qbx_locals": [P, P, P, P, P, ct.POINTER(KernelDesc), P, P, P, I],
    "qbx_translate_box_multipoles_to_qbx_local": [P, P, P, P, P, ct.POINTER(KernelDesc),
                                                  P, P, I],
    "qbx_translate_box_local_to_qbx_local": [P, P, P, P, P, ct.POINTER(KernelDesc), P, P, I],
    "qbx_eval_qbx_expansions": [P, P, P, P, ct.POINTER(KernelDesc), P, P],
    "qbx_finalize_potentials": [P, P, P, P, I, P, P],
    "qbx_drive_fmm": [P, P, P, P, P, ct.POINTER(KernelDesc), P, P, P],
    "qbx_drive_fmm_noalloc": [P, P, P, P, P, ct.POINTER(KernelDesc), P, P, P],
    "qbx_eval_target_specific_qbx_locals": [P, P, P, P, P, I, I, D, D, I, P, P, P, P],
    "qbx_eval_target_specific_qbx_locals_nrhs": [P, P, P, P, P, I, I, I, I, P, P, P, I],
    "qbx_drive_fmm_adjoint": [P, P, P, P, P, P, P, P, P],
    "qbx_adjoint_stage": [P, P, P, P, P, P, I, P, P, P],
    "qbx_direct_qbx": [P, P, P, P, ct.POINTER(KernelDesc), P, P, P],
    "qbx_p2p_all": [P, P, ct.POINTER(KernelDesc), I, P, P, P, I, P, P],
    "qbx_lpot_matrix": [P, P, ct.POINTER(KernelDesc), I, I, I, P, P, P, P, I, P, P, P, L, P],
    "qbx_element_geometry": [P, P, I, I, I, P, I, P, P, P, P, D, P, P, P, P, P],
    "qbx_check_expansion_disks_undisturbed": [P, P, I, I, P, I, P, P, P, D, P, P],
    "qbx_check_source_quadrature_resolution": [P, P, I, I, P, I, P, P, I, P, P, P],
    "qbx_helmholtz3d_p2p": [P, P, I, P, P, I, P, D, P, P],
    "qbx_helmholtz3d_p2p_adjoint": [P, P, I, P, I, P, D, P, P],
    "qbx_upsample_adjoint": [P, P, I, I, I, I, P, P, P],
}


def exported_symbols():
    return sorted(_SIGS) + ["qbx_last_error"]


def lib():
    """Load the shared library (raises if it is missing: no CPU fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing; build it with __graft_entry__.build() "
                "(hipcc --offload-arch=gfx950)")
        lb = ct.CDLL(LIB_PATH)
        lb.qbx_last_error.restype = ct.c_char_p
        lb.qbx_last_error.argtypes = []
        for name, args in _SIGS.items():
            fn = getattr(lb, name)
            fn.argtypes = args
            fn.restype = L if (name.endswith("_field_bytes")
                               or name == "qbx_expansion_size") else I
        _lib = lb
    return _lib


def check(rc):
    if rc != QBX_OK:
        raise QBXError(rc, lib().qbx_last_error().decode())
    return rc


def call(name, *args):
    return check(getattr(lib(), name)(*args))
