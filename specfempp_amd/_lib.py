"""ctypes binding of libspecfempp_amd.so (include/specfempp_amd.h).

The library is built in-tree by specfempp_amd/build.py.  Loading fails loudly
when it is missing: there is no CPU fallback for any operator.
"""

from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libspecfempp_amd.so")

SMX_FIELD_FORWARD, SMX_FIELD_ADJOINT, SMX_FIELD_BACKWARD = 0, 1, 2
SMX_U, SMX_V, SMX_A, SMX_MASS_INVERSE = 0, 1, 2, 3
SMX_OK, SMX_ERR_ARG, SMX_ERR_HIP, SMX_ERR_STATE, SMX_ERR_MESH = 0, -1, -2, -3, -4
SEIS_TYPES = {"displacement": 0, "velocity": 1, "acceleration": 2, "pressure": 3}
SMX_STF_RICKER, SMX_STF_DIRAC, SMX_STF_HEAVISIDE = 0, 1, 2
SMX_SRC_FORCE, SMX_SRC_MOMENT = 0, 1

# Every symbol include/specfempp_amd.h declares.
EXPORTS = [
    "smx_last_error", "smx_version", "smx_ctx_create2d", "smx_ctx_destroy",
    "smx_ctx_info", "smx_ctx_layout", "smx_ctx_t0", "smx_initialize", "smx_predictor",
    "smx_corrector", "smx_update_wavefields", "smx_compute_seismograms",
    "smx_run_forward", "smx_synchronize", "smx_get_seismograms", "smx_get_field",
    "smx_set_field", "smx_get_ibool", "smx_field_device_ptr", "smx_last_run_stats",
    "smx_time_stiffness", "smx_set_model", "smx_run_forward_checkpointed",
    "smx_run_adjoint", "smx_get_gradient", "smx_last_adjoint_stats",
    "smx_ctx_create2d_opts", "smx_ctx_create2d_batch", "smx_ctx_shots", "smx_ctx_compact",
    "smx_ctx_create3d", "smx_ctx_create3d_opts", "smx_ctx_geometry", "smx_ctx_components",
    "smx_update_wavefields_noalloc", "smx_predictor_medium", "smx_corrector_medium",
    "smx_update_wavefields_medium", "smx_ctx_degrees_of_freedom", "smx_set_model_device",
    "smx_get_gradient_device", "smx_ctx_set_launch",
]

# Medium tags (specfem::element::medium_tag, include/enumerations/medium.hpp:18-28).
MEDIUM_TAGS = {"elastic_psv": 0, "elastic_sh": 1, "elastic_psv_t": 2, "acoustic": 3,
               "poroelastic": 4, "electromagnetic_te": 5, "elastic": 6, "elastic_spin": 7,
               "electromagnetic": 8}


class SmxMesh2D(ctypes.Structure):
    _fields_ = [("nspec", ctypes.c_int), ("ngnod", ctypes.c_int), ("nnodes", ctypes.c_int),
                ("coord", ctypes.c_void_p), ("knods", ctypes.c_void_p),
                ("kappa", ctypes.c_void_p), ("mu", ctypes.c_void_p), ("rho", ctypes.c_void_p),
                ("nabs", ctypes.c_int), ("abs_ispec", ctypes.c_void_p),
                ("abs_type", ctypes.c_void_p), ("nfree", ctypes.c_int),
                ("free_ispec", ctypes.c_void_p), ("free_type", ctypes.c_void_p)]


class SmxSources2D(ctypes.Structure):
    _fields_ = [("nsrc", ctypes.c_int), ("x", ctypes.c_void_p), ("z", ctypes.c_void_p),
                ("angle", ctypes.c_void_p), ("f0", ctypes.c_void_p),
                ("tshift", ctypes.c_void_p), ("factor", ctypes.c_void_p),
                ("stf", ctypes.c_void_p), ("kind", ctypes.c_void_p),
                ("moment", ctypes.c_void_p)]


class SmxReceivers2D(ctypes.Structure):
    _fields_ = [("nrec", ctypes.c_int), ("x", ctypes.c_void_p), ("z", ctypes.c_void_p),
                ("ntypes", ctypes.c_int), ("types", ctypes.c_void_p),
                ("nstep_between_samples", ctypes.c_int)]


class SmxMesh3D(ctypes.Structure):
    _fields_ = [("nspec", ctypes.c_int), ("nnodes", ctypes.c_int), ("coord", ctypes.c_void_p),
                ("knods", ctypes.c_void_p), ("kappa", ctypes.c_void_p), ("mu", ctypes.c_void_p),
                ("rho", ctypes.c_void_p), ("nabs", ctypes.c_int), ("abs_ispec", ctypes.c_void_p),
                ("abs_face", ctypes.c_void_p)]


class SmxSources3D(ctypes.Structure):
    _fields_ = [("nsrc", ctypes.c_int), ("x", ctypes.c_void_p), ("y", ctypes.c_void_p),
                ("z", ctypes.c_void_p), ("kind", ctypes.c_void_p), ("comp", ctypes.c_void_p),
                ("f0", ctypes.c_void_p), ("tshift", ctypes.c_void_p), ("factor", ctypes.c_void_p)]


class SmxReceivers3D(ctypes.Structure):
    _fields_ = [("nrec", ctypes.c_int), ("x", ctypes.c_void_p), ("y", ctypes.c_void_p),
                ("z", ctypes.c_void_p), ("ntypes", ctypes.c_int), ("types", ctypes.c_void_p),
                ("nstep_between_samples", ctypes.c_int)]


class SmxOptions(ctypes.Structure):
    _fields_ = [("tile_elems", ctypes.c_int), ("graph_steps", ctypes.c_int),
                ("geometry_from_nodes", ctypes.c_int), ("waves_per_eu", ctypes.c_int)]


class SmxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


_lib = None


def load():
    """Loads the in-tree library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -m specfempp_amd.build` "
            "(the HIP path has no CPU fallback)")
    lib = ctypes.CDLL(LIB_PATH)
    i, d, P = ctypes.c_int, ctypes.c_double, ctypes.c_void_p
    IP = ctypes.POINTER(ctypes.c_int)
    DP = ctypes.POINTER(ctypes.c_double)
    sig = {
        "smx_last_error": (ctypes.c_char_p, []),
        "smx_version": (ctypes.c_char_p, []),
        "smx_ctx_create2d": (i, [ctypes.POINTER(SmxMesh2D), ctypes.POINTER(SmxSources2D),
                                 ctypes.POINTER(SmxReceivers2D), d, i, i, ctypes.POINTER(P)]),
        "smx_ctx_destroy": (i, [P]),
        "smx_ctx_info": (i, [P, IP, IP, IP, IP]),
        "smx_ctx_components": (i, [P, IP]),
        "smx_ctx_layout": (i, [P, IP, IP, IP, IP]),
        "smx_ctx_t0": (d, [P]),
        "smx_initialize": (i, [P]),
        "smx_predictor": (i, [P, i, IP]),
        "smx_corrector": (i, [P, i, IP]),
        "smx_update_wavefields": (i, [P, i, i, IP]),
        "smx_update_wavefields_noalloc": (i, [P, P, P, P, i, IP]),
        "smx_predictor_medium": (i, [P, i, i, IP]),
        "smx_corrector_medium": (i, [P, i, i, IP]),
        "smx_update_wavefields_medium": (i, [P, i, i, i, IP]),
        "smx_ctx_degrees_of_freedom": (i, [P, ctypes.POINTER(ctypes.c_longlong)]),
        "smx_set_model_device": (i, [P, P, P, P]),
        "smx_get_gradient_device": (i, [P, P]),
        "smx_compute_seismograms": (i, [P, i, i]),
        "smx_run_forward": (i, [P, i, i]),
        "smx_ctx_set_launch": (i, [P, i, i]),
        "smx_synchronize": (i, [P]),
        "smx_get_seismograms": (i, [P, P]),
        "smx_get_field": (i, [P, i, i, P]),
        "smx_set_field": (i, [P, i, i, P]),
        "smx_get_ibool": (i, [P, P]),
        "smx_field_device_ptr": (i, [P, i, i, ctypes.POINTER(P)]),
        "smx_last_run_stats": (i, [P, DP, DP, IP]),
        "smx_time_stiffness": (i, [P, i, DP]),
        "smx_set_model": (i, [P, P, P, P]),
        "smx_run_forward_checkpointed": (i, [P, i]),
        "smx_run_adjoint": (i, [P, P]),
        "smx_get_gradient": (i, [P, P]),
        "smx_last_adjoint_stats": (i, [P, DP, DP]),
        "smx_ctx_create2d_opts": (i, [ctypes.POINTER(SmxMesh2D), ctypes.POINTER(SmxSources2D),
                                      ctypes.POINTER(SmxReceivers2D), d, i, i,
                                      ctypes.POINTER(SmxOptions), ctypes.POINTER(P)]),
        "smx_ctx_create2d_batch": (i, [i, ctypes.POINTER(SmxMesh2D), ctypes.POINTER(SmxSources2D),
                                       ctypes.POINTER(SmxReceivers2D), d, i, i,
                                       ctypes.POINTER(SmxOptions), ctypes.POINTER(P)]),
        "smx_ctx_shots": (i, [P, IP, P, P, P]),
        "smx_ctx_compact": (i, [P, IP, IP]),
        "smx_ctx_geometry": (i, [P, IP, ctypes.POINTER(ctypes.c_longlong)]),
        "smx_ctx_create3d": (i, [ctypes.POINTER(SmxMesh3D), ctypes.POINTER(SmxSources3D),
                                 ctypes.POINTER(SmxReceivers3D), d, i, i, ctypes.POINTER(P)]),
        "smx_ctx_create3d_opts": (i, [ctypes.POINTER(SmxMesh3D), ctypes.POINTER(SmxSources3D),
                                      ctypes.POINTER(SmxReceivers3D), d, i, i,
                                      ctypes.POINTER(SmxOptions), ctypes.POINTER(P)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status):
    if status != 0:
        raise SmxError(status, load().smx_last_error().decode())
    return status
