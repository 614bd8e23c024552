"""ctypes bindings of libtensors_amd.so (the C ABI declared in include/tensors_amd.h).

The library is built in-tree by ``make -C tensors_amd/csrc`` (``__graft_entry__.build``).
Loading fails loudly when it is missing: there is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtensors_amd.so")

# tg_dtype
F64, F32, I32, I64 = 0, 1, 2, 3
# tg_status
OK, ERR_INVALID, ERR_ALIAS, ERR_NRHS, ERR_DTYPE, ERR_HIP, ERR_WORKSPACE, ERR_SHAPE = 0, -1, -2, -3, -4, -5, -6, -7

STATUS_NAMES = {
    OK: "TG_OK", ERR_INVALID: "TG_ERR_INVALID", ERR_ALIAS: "TG_ERR_ALIAS", ERR_NRHS: "TG_ERR_NRHS",
    ERR_DTYPE: "TG_ERR_DTYPE", ERR_HIP: "TG_ERR_HIP", ERR_WORKSPACE: "TG_ERR_WORKSPACE",
    ERR_SHAPE: "TG_ERR_SHAPE",
}

i32, i64, u64, dbl, vp = C.c_int32, C.c_int64, C.c_uint64, C.c_double, C.c_void_p

# name -> (restype, argtypes)
SIGNATURES = {
    "tg_last_error": (C.c_char_p, []),
    "tg_version": (C.c_char_p, []),
    "tg_csr_job_pointers": (C.c_int, [i32, vp, C.c_int, i32, vp, vp]),
    "tg_csr_default_blocks": (i32, [i32, i64, i32]),
    "tg_csr_spmm": (C.c_int, [i32, i32, i64, vp, C.c_int, vp, vp, C.c_int, dbl, vp, i64, dbl, vp, i64,
                              i32, i32, vp, i32, vp]),
    "tg_csr_plan_create": (C.c_int, [i32, i32, i64, vp, C.c_int, vp, vp, vp]),
    "tg_csr_plan_destroy": (C.c_int, [vp]),
    "tg_csr_plan_slices": (i32, [vp, C.c_int, i32]),
    "tg_csr_plan_set_values": (C.c_int, [vp, vp, C.c_int, vp]),
    "tg_csr_plan_uses_split": (i32, [vp, vp, C.c_int, i32]),
    "tg_csr_plan_long_rows": (i32, [vp]),
    "tg_csr_spmm_plan": (C.c_int, [vp, vp, C.c_int, dbl, vp, i64, dbl, vp, i64, i32, i32, vp]),
    "tg_tune_set": (C.c_int, [C.c_char_p, i32]),
    "tg_tune_get": (i32, [C.c_char_p]),
    "tg_axpby": (C.c_int, [i32, i32, C.c_int, dbl, vp, i64, dbl, vp, i64, vp]),
    "tg_cg_workspace_bytes": (C.c_size_t, [i32, i32, C.c_int, i32]),
    "tg_cg_solve": (C.c_int, [i32, i64, vp, C.c_int, vp, vp, C.c_int, i32, vp, i32, i32, dbl, vp, i64, dbl, vp,
                              i64, dbl, i32, vp, C.c_size_t, vp, vp, vp, vp]),
    "tg_cg_residuals": (C.c_int, [vp, C.c_size_t, i32, vp, vp, vp]),
    "tg_cg_solve_noalloc": (C.c_int, [i32, i64, vp, C.c_int, vp, vp, C.c_int, i32, vp, i32, i32, dbl, vp, i64, dbl,
                                      vp, i64, dbl, i32, vp, vp, vp, vp]),
    "tg_gmres_workspace_bytes": (C.c_size_t, [i32, i32, i32, C.c_int]),
    "tg_gmres_workspace_parts": (C.c_int, [vp, C.c_size_t, i32, i32, i32, C.c_int, vp, vp, vp, vp]),
    "tg_gmres_solve": (C.c_int, [i32, i64, vp, C.c_int, vp, vp, C.c_int, i32, vp, i32, i32, i32, dbl, vp, i64,
                                 dbl, vp, i64, dbl, i32, i32, vp, C.c_size_t, vp, vp, vp, vp, vp]),
    "tg_cg_solve_op": (C.c_int, [i32, vp, vp, vp, vp, C.c_int, i32, i32, dbl, vp, i64, dbl, vp, i64, dbl, i32,
                                 vp, C.c_size_t, vp, vp, vp, vp]),
    "tg_gmres_solve_op": (C.c_int, [i32, vp, vp, vp, vp, C.c_int, i32, i32, i32, dbl, vp, i64, dbl, vp, i64,
                                    dbl, i32, i32, vp, C.c_size_t, vp, vp, vp, vp, vp]),
    "tg_bsr_spmm": (C.c_int, [i32, i32, i64, i32, i32, vp, C.c_int, vp, vp, C.c_int, i32, dbl, vp, i32, dbl, vp,
                              i32, i32, vp]),
    "tg_bsr_fill_lower": (C.c_int, [i32, vp, C.c_int, vp, i32, vp, C.c_int, vp]),
    "tg_mm_info": (C.c_int, [C.c_char_p, vp, vp, vp, vp, vp]),
    "tg_mm_read": (C.c_int, [C.c_char_p, i64, vp, vp, vp, i32]),
    "tg_mm_write": (C.c_int, [C.c_char_p, i32, i32, vp, C.c_int, vp, vp, C.c_int, i32]),
    "tg_csr_from_triples_workspace_bytes": (C.c_size_t, [i64, i32, i32]),
    "tg_csr_from_triples": (C.c_int, [i64, vp, vp, vp, C.c_int, i32, i32, i32, i32, vp, C.c_int, vp, vp, i64,
                                      vp, C.c_size_t, vp, vp]),
    "tg_csr_transpose_workspace_bytes": (C.c_size_t, [i32, i32, i64]),
    "tg_csr_transpose": (C.c_int, [i32, i32, i64, vp, C.c_int, vp, vp, C.c_int, vp, vp, vp, vp, C.c_size_t, vp]),
    "tg_csr_spgemm_expanded_count": (C.c_int, [i32, vp, vp, vp, C.c_int, vp, vp, C.c_size_t, vp]),
    "tg_csr_spgemm_workspace_bytes": (C.c_size_t, [i32, i32, i64, C.c_int]),
    "tg_csr_spgemm": (C.c_int, [i32, i32, i32, vp, vp, vp, vp, vp, vp, C.c_int, C.c_int, i64, vp, vp, vp, vp,
                                C.c_size_t, vp, vp]),
    "tg_cg_mixed_workspace_bytes": (C.c_size_t, [i32, i64, i32, i32, i32]),
    "tg_cg_solve_mixed": (C.c_int, [i32, i64, vp, C.c_int, vp, vp, vp, i32, vp, i32, i32, i32, dbl, dbl, vp, i64,
                                    dbl, vp, i64, dbl, i32, vp, C.c_size_t, vp, vp, vp, vp, vp]),
    "tg_gmres_mixed_workspace_bytes": (C.c_size_t, [i32, i64, i32, i32, i32, i32]),
    "tg_gmres_solve_mixed": (C.c_int, [i32, i64, vp, C.c_int, vp, vp, vp, i32, vp, i32, i32, i32, i32, dbl, vp,
                                       i64, dbl, vp, i64, dbl, i32, vp, C.c_size_t, vp, vp, vp, vp, vp]),
    "tg_gbatch_create": (C.c_int, [vp, vp]),
    "tg_gbatch_destroy": (C.c_int, [vp]),
    "tg_gbatch_add": (C.c_int, [vp, i32, i64, vp, vp, vp, vp, i32, vp, vp, dbl, i32, i32, vp]),
    "tg_gbatch_step": (C.c_int, [vp, i32]),
    "tg_gbatch_finished": (C.c_int, [vp, vp, i32, vp]),
    "tg_gbatch_result": (C.c_int, [vp, i64, vp, vp, vp, vp, vp]),
    "tg_gbatch_release": (C.c_int, [vp, i64]),
    "tg_gbatch_active": (C.c_int, [vp, vp, vp]),
    "tg_gbatch_request_bytes": (C.c_size_t, [i32, i32, i32]),
    "tg_batch_create": (C.c_int, [vp, vp]),
    "tg_batch_destroy": (C.c_int, [vp]),
    "tg_batch_add_cg": (C.c_int, [vp, i32, i64, vp, vp, vp, vp, i32, vp, vp, dbl, i32, i32, vp]),
    "tg_batch_step": (C.c_int, [vp, i32]),
    "tg_batch_finished": (C.c_int, [vp, vp, i32, vp]),
    "tg_batch_result": (C.c_int, [vp, i64, vp, vp, vp, vp, vp]),
    "tg_batch_release": (C.c_int, [vp, i64]),
    "tg_batch_active": (C.c_int, [vp, vp, vp]),
    "tg_batch_request_bytes": (C.c_size_t, [i32, i64, i32, i32, i32]),
    "tg_server_create": (C.c_int, [i64, i32, i32, vp]),
    "tg_server_destroy": (C.c_int, [vp]),
    "tg_server_submit_cg": (C.c_int, [vp, i32, i64, vp, vp, vp, vp, i32, vp, vp, dbl, i32, i32, vp]),
    "tg_server_submit_gmres": (C.c_int, [vp, i32, i64, vp, vp, vp, i32, vp, vp, dbl, i32, i32, vp]),
    "tg_server_wait": (C.c_int, [vp, vp, i32, vp, dbl]),
    "tg_server_result": (C.c_int, [vp, i64, vp, vp, vp, vp, vp, vp]),
    "tg_server_release": (C.c_int, [vp, i64]),
    "tg_server_stats": (C.c_int, [vp, vp, vp, vp, vp]),
    "tg_server_request_bytes": (C.c_size_t, [i32, i32, i64, i32, i32, i32, i32]),
    "tg_probe_copy": (C.c_int, [vp, vp, i64, vp]),
    "tg_probe_read": (C.c_int, [vp, i64, vp, vp]),
    "tg_gen_uniform": (C.c_int, [u64, u64, i64, C.c_int, vp]),
    "tg_gen_random_csr_rowptr": (i64, [i64, dbl, i32, i32, u64, vp]),
    "tg_gen_random_csr_fill": (C.c_int, [i64, i64, vp, u64, vp, vp, vp]),
    "tg_gen_laplacian_3d": (i64, [i32, dbl, vp, vp, vp]),
    "tg_gen_weighted_grid": (i64, [i32, i32, i32, i32, dbl, u64, vp, vp, vp]),
    "tg_gen_convection_diffusion_2d": (i64, [i32, dbl, dbl, vp, vp, vp]),
}

APPLY_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p)

_lib = None


class TensorsError(RuntimeError):
    """A tg_* call returned a non-zero status."""

    def __init__(self, status: int, where: str, message: str):
        self.status = status
        super().__init__(f"{where}: {STATUS_NAMES.get(status, status)}: {message}")


def lib() -> C.CDLL:
    """The loaded library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `make -C tensors_amd/csrc` "
                "(or __graft_entry__.build()); there is no CPU fallback")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def last_error() -> str:
    return lib().tg_last_error().decode()


def check(status: int, where: str) -> int:
    if status < 0:
        raise TensorsError(int(status), where, last_error())
    return status


def call(name: str, *args) -> int:
    """Call tg_<name> and raise TensorsError on a negative status."""
    return check(getattr(lib(), name)(*args), name)
