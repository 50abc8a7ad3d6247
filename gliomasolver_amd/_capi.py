"""ctypes binding of libgliomahip.so (include/gliomahip.h).

The library is built in-tree (gliomasolver_amd/libgliomahip.so).  Loading
fails loudly when it is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
import re

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libgliomahip.so")
HEADER_PATH = os.path.join(os.path.dirname(_PKG), "include", "gliomahip.h")

GH_OK = 0
GH_REAL_F32 = 0
GH_REAL_F64 = 1
FIELD_PHI, FIELD_PW, FIELD_PG, FIELD_PCSF, FIELD_DPHIDT = 0, 1, 2, 3, 4

_vp = C.c_void_p
_i = C.c_int
_d = C.c_double
_i64 = C.c_int64
_u64 = C.c_uint64
_pd = C.POINTER(C.c_double)
_pi = C.POINTER(C.c_int)
_pi64 = C.POINTER(C.c_int64)

# name -> (restype, argtypes)
SIGNATURES = {
    "gh_ctx_create": (_i, [_i, _i, C.POINTER(_vp)]),
    "gh_ctx_destroy": (_i, [_vp]),
    "gh_last_error": (C.c_char_p, []),
    "gh_sync": (_i, [_vp]),
    "gh_grid_init_uniform": (_i, [_vp, _i]),
    "gh_grid_nblocks": (_i, [_vp, _pi]),
    "gh_grid_blocks": (_i, [_vp, _vp, _vp]),
    "gh_grid_configure": (_i, [_vp, _i, _i, _i]),
    "gh_grid_init_leaves": (_i, [_vp, _i, _vp, _vp]),
    "gh_fwt_details": (_i, [_vp, _vp]),
    "gh_refine": (_i, [_vp, _d, _i, _pi]),
    "gh_compress": (_i, [_vp, _d, _i, _pi]),
    "gh_grid_refine_blocks": (_i, [_vp, _i, _vp, _pi]),
    "gh_grid_compress_blocks": (_i, [_vp, _i, _vp, _pi]),
    "gh_adapt_profile": (_i, [_vp, _vp, _i]),
    "gh_adapt_counters": (_i, [_vp, _vp, _i]),
    "gh_set_field_blocks": (_i, [_vp, _i, _vp]),
    "gh_get_field_blocks": (_i, [_vp, _i, _vp]),
    "gh_uq_resample": (_i, [_vp, _i, _vp]),
    "gh_plan_neighbours": (_i, [_i, _vp, _i, _pi, _vp]),
    "gh_plan_refine": (_i, [_i, _vp, _vp, _i, _pi, _vp]),
    "gh_plan_compress": (_i, [_i, _vp, _vp, _i, _pi, _vp]),
    "gh_set_anatomy": (_i, [_vp, _i, _vp, _vp, _vp]),
    "gh_ic_from_anatomy": (_i, [_vp, _i, _i, _i, _vp, _vp, _vp, _pd]),
    "gh_set_field_dense": (_i, [_vp, _i, _i, _vp]),
    "gh_get_field_dense": (_i, [_vp, _i, _i, _vp]),
    "gh_schedule": (_i, [_i, _d, _d, _d, _i, _d, _pd, _pd, _pd, _pi64]),
    "gh_rd_prepare": (_i, [_vp, _d, _d, _d]),
    "gh_rd_rhs": (_i, [_vp]),
    "gh_rd_update": (_i, [_vp, _d]),
    "gh_rd_steps": (_i, [_vp, _i64, _d, _pd]),
    "gh_rd_set_tile": (_i, [_vp, _i]),
    "gh_rd_autotune": (_i, [_vp, _d, _i, _pi]),
    "gh_rd_schedule_info": (_i, [_vp, _pi, _pi]),
    "gh_rd_zero_flux_blocks": (_i, [_vp, _pi]),
    "gh_rd_thresholds": (_i, [_i, _i64, _vp, _vp, _vp, _vp]),
    "gh_lik_set_data": (_i, [_vp, _i, _vp, _vp, _vp, _vp, _i]),
    "gh_likelihood": (_i, [_vp, _pd, _i, _i, _i, _pd, _pi64]),
    "gh_likelihood_dense": (_i, [_vp, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _i, _pd, _i, _i, _pd, _pi64]),
    "gh_likelihood_noalloc": (_i, [_vp, _pd, _i, _i, _i, _pd, _pi64]),
    "gh_rd_adjoint": (_i, [_vp, _i64, _d, _i, _pd, _i, _i, _i, _pd, _d, _pd, _pd]),
    "gh_checkpoint_save": (_i, [_vp, _i]),
    "gh_checkpoint_load": (_i, [_vp, _i]),
    "gh_batch_create": (_i, [_i, _i, C.POINTER(_vp)]),
    "gh_batch_destroy": (_i, [_vp]),
    "gh_batch_add_anatomy": (_i, [_vp, _i, _i, _i, _i, _vp, _vp, _vp, _pi]),
    "gh_batch_set_observations": (_i, [_vp, _i, _vp, _vp, _vp, _vp, _i]),
    "gh_batch_set_samples": (_i, [_vp, _i, _vp, _vp, _d]),
    "gh_batch_run": (_i, [_vp, _pd]),
    "gh_batch_likelihood": (_i, [_vp, _vp, _i, _i, _i, _vp]),
    "gh_batch_set_adjoint_samples": (_i, [_vp, _i, _vp, _vp, _vp, _d, _d]),
    "gh_batch_adjoint": (_i, [_vp, _i, _vp, _i, _i, _i, _vp, _vp, _pd]),
    "gh_batch_sample_info": (_i, [_vp, _i, _pd, _pd, _pd, _pi64]),
    "gh_batch_get_phi": (_i, [_vp, _i, _vp]),
    "gh_batch_get_anatomy": (_i, [_vp, _i, _vp, _vp]),
    "gh_tmcmc_create": (_i, [_i, _vp, _vp, _i, _i, _i, _i, _d, _d, _u64, C.POINTER(_vp)]),
    "gh_tmcmc_destroy": (_i, [_vp]),
    "gh_tmcmc_candidates": (_i, [_vp, _pi, _vp]),
    "gh_tmcmc_submit": (_i, [_vp, _i, _vp]),
    "gh_tmcmc_state": (_i, [_vp, _pi, _pd, _pd, _pd, _pi]),
    "gh_tmcmc_population": (_i, [_vp, _pi, _vp, _vp]),
    "gh_tmcmc_next_exponent": (_i, [_i, _vp, _d, _d, _pd, _pd]),
    "gh_glioma_params": (_i, [_i, _vp, _i, _vp, _vp]),
    "gh_glioma_fitness": (_i, [_i, _vp, _vp]),
    "gh_phantom": (_i, [_i, _u64, _vp, _vp, _vp, _pi]),
    "gh_gaussian_ic": (_i, [_i, _i, _vp, _vp, _pd, _d, _d, _vp]),
}


class GHError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__("gliomahip error %d: %s" % (code, msg))
        self.code = code


_lib = None


def load() -> C.CDLL:
    """Load libgliomahip.so once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("libgliomahip.so not built at %s (run python -m gliomasolver_amd.build)" % LIB_PATH)
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int) -> None:
    if rc != GH_OK:
        msg = load().gh_last_error()
        raise GHError(rc, msg.decode() if msg else "")


def declared_symbols() -> list:
    """Function names declared in include/gliomahip.h."""
    txt = open(HEADER_PATH).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(gh_\w+)\s*\(", txt, flags=re.M)))
