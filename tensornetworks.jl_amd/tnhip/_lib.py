"""ctypes binding of libtnhip.so (declared in include/tnhip.h).

The library is the only compute path: if it is missing, importing this module
raises ImportError and nothing falls back to the CPU.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

_HERE = Path(__file__).resolve().parent
LIB_PATH = Path(os.environ.get("TNHIP_LIB", _HERE / "libtnhip.so"))

if not LIB_PATH.exists():
    raise ImportError(
        f"libtnhip.so not found at {LIB_PATH}; build it with "
        "`python tensornetworks.jl_amd/build.py` (hipcc, gfx950)"
    )

lib = C.CDLL(str(LIB_PATH))

# dtype / precision / status codes (include/tnhip.h)
F64, F32, C128, C64 = 0, 1, 2, 3
PREC_F64, PREC_F32, PREC_BF16X3 = 0, 1, 2
STATUS = {
    0: "TNH_OK",
    -1: "TNH_ERR_ARG",
    -2: "TNH_ERR_HIP",
    -3: "TNH_ERR_SHAPE",
    -4: "TNH_ERR_NOMEM",
    -5: "TNH_ERR_NUMERIC",
    -6: "TNH_ERR_NOCONV",
    -7: "TNH_ERR_UNSUPPORTED",
}

ctx_p = C.c_void_p
buf_p = C.c_void_p
i64p = C.POINTER(C.c_int64)


class DmrgParams(C.Structure):
    _fields_ = [
        ("krylovdim", C.c_int),
        ("maxiter", C.c_int),
        ("lanczos_tol", C.c_double),
        ("cutoff", C.c_double),
        ("maxdim", C.c_int),
        ("mindim", C.c_int),
        ("precision", C.c_int),
        ("nsites", C.c_int),
        ("nonhermitian", C.c_int),
    ]


class QjmcProblem(C.Structure):
    _fields_ = [
        ("N", C.c_int),
        ("d", C.c_int),
        ("nlayers", C.c_int),
        ("layer_count", C.POINTER(C.c_int)),
        ("gate_site", C.POINTER(C.c_int)),
        ("gate_range", C.POINTER(C.c_int)),
        ("gate_offset", C.POINTER(C.c_int64)),
        ("gate_data", C.POINTER(C.c_double)),
        ("njump", C.c_int),
        ("jump_site", C.POINTER(C.c_int)),
        ("jump_coeff", C.POINTER(C.c_double)),
        ("jump_op", C.POINTER(C.c_double)),
        ("nobs", C.c_int),
        ("obs_site", C.POINTER(C.c_int)),
        ("obs_op", C.POINTER(C.c_double)),
        ("init_state", C.POINTER(C.c_double)),
        ("dt", C.c_double),
        ("nsteps", C.c_int),
        ("savesteps", C.c_int),
        ("cutoff", C.c_double),
        ("maxdim", C.c_int),
        ("mindim", C.c_int),
    ]


# name -> (restype, argtypes); every entry point of include/tnhip.h
SIGNATURES = {
    "tnh_version": (C.c_char_p, []),
    "tnh_ctx_create": (C.c_int, [C.c_int, C.POINTER(ctx_p)]),
    "tnh_ctx_destroy": (C.c_int, [ctx_p]),
    "tnh_last_error": (C.c_char_p, [ctx_p]),
    "tnh_ctx_synchronize": (C.c_int, [ctx_p]),
    "tnh_profile_enable": (C.c_int, [ctx_p, C.c_int]),
    "tnh_profile_reset": (C.c_int, [ctx_p]),
    "tnh_profile_read": (C.c_int, [ctx_p, C.c_char_p, i64p, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "tnh_gemm_autotune": (C.c_int, [ctx_p, C.c_int]),
    "tnh_gemm_tuning_report": (C.c_int, [ctx_p, C.c_char_p, C.c_int64, i64p]),
    "tnh_buf_create": (C.c_int, [ctx_p, C.c_int, C.c_int, i64p, C.POINTER(buf_p)]),
    "tnh_buf_upload": (C.c_int, [ctx_p, C.c_void_p, C.c_int, C.c_int, C.c_int, i64p, C.POINTER(buf_p)]),
    "tnh_buf_download": (C.c_int, [ctx_p, buf_p, C.c_void_p, C.c_int]),
    "tnh_buf_wrap": (C.c_int, [ctx_p, C.c_void_p, C.c_int, C.c_int, i64p, C.POINTER(buf_p)]),
    "tnh_buf_free": (C.c_int, [ctx_p, buf_p]),
    "tnh_buf_shape": (C.c_int, [buf_p, C.POINTER(C.c_int), C.POINTER(C.c_int), i64p]),
    "tnh_buf_device_ptr": (C.c_void_p, [buf_p]),
    "tnh_buf_convert": (C.c_int, [ctx_p, buf_p, C.c_int, C.POINTER(buf_p)]),
    "tnh_gemm": (C.c_int, [ctx_p, C.c_int, C.c_int, C.c_int64, C.c_int64, C.c_int64, C.c_double, buf_p, C.c_int64, buf_p, C.c_int64, C.c_double, buf_p, C.c_int64, C.c_int]),
    "tnh_buf_norm": (C.c_int, [ctx_p, buf_p, C.POINTER(C.c_double)]),
    "tnh_buf_scale": (C.c_int, [ctx_p, buf_p, C.c_double, C.c_double]),
    "tnh_apply_site_op": (C.c_int, [ctx_p, C.POINTER(buf_p), buf_p]),
    "tnh_heff_apply": (C.c_int, [ctx_p, buf_p, buf_p, buf_p, buf_p, buf_p, buf_p, C.c_double, C.c_double, C.c_int]),
    "tnh_heff_apply_noalloc": (C.c_int, [ctx_p, buf_p, buf_p, buf_p, buf_p, buf_p, buf_p, C.c_double, C.c_double, C.c_int]),
    "tnh_env_left": (C.c_int, [ctx_p, buf_p, buf_p, buf_p, C.POINTER(buf_p), C.c_int]),
    "tnh_env_right": (C.c_int, [ctx_p, buf_p, buf_p, buf_p, C.POINTER(buf_p), C.c_int]),
    "tnh_env_calculate": (C.c_int, [ctx_p, buf_p, buf_p, buf_p, buf_p, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "tnh_eigsolve_ground": (C.c_int, [ctx_p, buf_p, buf_p, buf_p, buf_p, buf_p, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int)]),
    "tnh_svd_trunc": (C.c_int, [ctx_p, buf_p, C.c_double, C.c_int, C.c_int, C.c_int, C.POINTER(buf_p), C.POINTER(buf_p), C.POINTER(buf_p), C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_int)]),
    "tnh_replacesites2": (C.c_int, [ctx_p, buf_p, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, C.POINTER(buf_p), C.POINTER(buf_p), C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int]),
    "tnh_two_site_theta": (C.c_int, [ctx_p, buf_p, buf_p, C.POINTER(buf_p)]),
    "tnh_mps_movecenter": (C.c_int, [ctx_p, C.c_int, C.POINTER(buf_p), C.c_int, C.c_int, C.c_double, C.c_int, C.c_int]),
    "tnh_applygate2": (C.c_int, [ctx_p, C.POINTER(buf_p), C.POINTER(buf_p), buf_p, C.c_int, C.c_double, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double)]),
    "tnh_applygate_sites": (C.c_int, [ctx_p, C.c_int, C.POINTER(buf_p), buf_p, C.c_int, C.c_double, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double)]),
    "tnh_replacesites_sites": (C.c_int, [ctx_p, C.c_int, buf_p, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, C.POINTER(buf_p), C.POINTER(C.c_int), C.POINTER(C.c_double)]),
    "tnh_inner_oplist": (C.c_int, [ctx_p, C.c_int, C.POINTER(buf_p), C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "tnh_inner_oplist_pair": (C.c_int, [ctx_p, C.c_int, C.POINTER(buf_p), C.POINTER(buf_p), C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "tnh_inner_mpo": (C.c_int, [ctx_p, C.c_int, C.POINTER(buf_p), C.POINTER(buf_p), C.POINTER(buf_p), C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "tnh_mpo_from_oplist": (C.c_int, [ctx_p, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_double, C.c_int, C.c_int, C.POINTER(buf_p)]),
    "tnh_mpo_compress": (C.c_int, [ctx_p, C.c_int, C.POINTER(buf_p), C.c_double, C.c_int, C.c_int]),
    "tnh_mpo_apply": (C.c_int, [ctx_p, C.c_int, C.POINTER(buf_p), C.POINTER(buf_p), C.c_double, C.c_int, C.c_int, C.POINTER(buf_p)]),
    "tnh_mpo_mpo_product": (C.c_int, [ctx_p, C.c_int, C.POINTER(buf_p), C.POINTER(buf_p), C.c_double, C.c_int, C.c_int, C.POINTER(buf_p)]),
    "tnh_qjmc_run": (C.c_int, [ctx_p, C.POINTER(QjmcProblem), C.c_int, C.POINTER(C.c_double), C.c_int64, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_int)]),
    "tnh_qjmc_run_batched": (C.c_int, [ctx_p, C.POINTER(QjmcProblem), C.c_int, C.POINTER(C.c_double), C.c_int64, C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_int)]),
    "tnh_mpo_expect_grad":(C.c_int, [ctx_p, C.c_int, C.POINTER(buf_p), C.POINTER(buf_p), C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(buf_p)]),
    "tnh_dmrg_create":(C.c_int, [ctx_p, C.c_int, C.POINTER(buf_p), C.c_int, C.POINTER(buf_p), C.POINTER(C.c_double), C.POINTER(C.c_void_p)]),
    "tnh_dmrg_add_projector":(C.c_int, [C.c_void_p, C.POINTER(buf_p), C.c_double]),
    "tnh_dmrg_destroy": (C.c_int, [C.c_void_p]),
    "tnh_dmrg_sweep": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(DmrgParams), C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_double)]),
    "tnh_dmrg_energy": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_double)]),
    "tnh_dmrg_last_ritz": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "tnh_dmrg_site": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(buf_p)]),
    "tnh_dmrg_center": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "tnh_dmrg_bonddims": (C.c_int, [C.c_void_p, i64p]),
}

MISSING = []
for _name, (_res, _args) in SIGNATURES.items():
    try:
        _f = getattr(lib, _name)
    except AttributeError:
        MISSING.append(_name)
        continue
    _f.restype = _res
    _f.argtypes = _args


class TnhError(RuntimeError):
    """A non-zero status of a libtnhip call, with tnh_last_error's message."""

    def __init__(self, status: int, message: str, fn: str = ""):
        self.status = status
        self.message = message
        super().__init__(f"{fn}: {STATUS.get(status, status)}: {message}")


def version() -> str:
    return lib.tnh_version().decode()
