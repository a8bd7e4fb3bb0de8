"""Loader for the native library libzxh_amd.so (the C ABI of include/zxh_amd.h).

There is no fallback: if the library is missing or no GPU is visible, calls
fail loudly.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libzxh_amd.so")

ZXH_OK = 0
ZXH_C128 = 0
ZXH_C64 = 1
ZXH_POLICY_REFERENCE = 0
ZXH_POLICY_FUSED = 1

_sz = C.c_size_t
_u64 = C.c_uint64
_dp = C.POINTER(C.c_double)
_u64p = C.POINTER(C.c_uint64)
_vp = C.c_void_p
_i32p = C.POINTER(C.c_int32)

# (name, restype, argtypes); every function returns int status unless noted
_SIGS = [
    ("zxh_last_error", C.c_char_p, []),
    ("zxh_init", C.c_int, []),
    ("zxh_finalize", C.c_int, []),
    ("zxh_active", C.c_int, []),
    ("zxh_rank", C.c_int, []),
    ("zxh_nprocs", C.c_int, []),
    ("zxh_create", C.c_int, [_sz, C.c_int, C.c_int, C.c_int, C.POINTER(_vp)]),
    ("zxh_destroy", C.c_int, [_vp]),
    ("zxh_set_seed", C.c_int, [_vp, _u64]),
    ("zxh_clear_seed", C.c_int, [_vp]),
    ("zxh_clear_gates", C.c_int, [_vp]),
    ("zxh_barrier", C.c_int, [_vp]),
    ("zxh_rz", C.c_int, [_vp, _sz, C.c_double]),
    ("zxh_crz", C.c_int, [_vp, _sz, _sz, C.c_double]),
    ("zxh_cp", C.c_int, [_vp, _sz, _sz, C.c_double]),
    ("zxh_p", C.c_int, [_vp, _sz, C.c_double]),
    ("zxh_z", C.c_int, [_vp, _sz]),
    ("zxh_x", C.c_int, [_vp, _sz]),
    ("zxh_cx", C.c_int, [_vp, _sz, _sz]),
    ("zxh_h", C.c_int, [_vp, _sz]),
    ("zxh_u3", C.c_int, [_vp, _sz, C.c_double, C.c_double, C.c_double]),
    ("zxh_rx", C.c_int, [_vp, _sz, C.c_double]),
    ("zxh_gate", C.c_int, [_vp, C.c_int, _sz, _sz, C.c_double, C.c_double, C.c_double]),
    ("zxh_execute", C.c_int, [_vp]),
    ("zxh_measure", C.c_int, [_vp, _sz]),
    ("zxh_measured_count", C.c_int, [_vp, C.POINTER(_sz)]),
    ("zxh_get_results", C.c_int, [_vp, _u64p, _sz]),
    ("zxh_get_real_results", C.c_int, [_vp, _u64p, _sz]),
    ("zxh_get_state", C.c_int, [_vp, _dp, _sz]),
    ("zxh_required_M", C.c_int, [_vp, C.POINTER(_sz)]),
    ("zxh_num_qubits", C.c_int, [_vp, C.POINTER(_sz)]),
    ("zxh_set_policy", C.c_int, [_vp, C.c_int]),
    ("zxh_set_seg_bits", C.c_int, [_vp, _sz]),
    ("zxh_get_phys_state", C.c_int, [_vp, _vp, _sz, C.POINTER(_sz)]),
    ("zxh_set_phys_state", C.c_int, [_vp, _vp, _sz]),
    ("zxh_state_device_ptr", C.c_int, [_vp, C.POINTER(_vp), C.POINTER(_sz)]),
    ("zxh_set_timing", C.c_int, [_vp, C.c_int]),
    ("zxh_get_timing", C.c_int, [_vp, C.c_int, _u64p, _dp, _u64p]),
    ("zxh_reset_timing", C.c_int, [_vp]),
    ("zxh_dev_alloc", C.c_int, [_sz, C.POINTER(_vp)]),
    ("zxh_dev_free", C.c_int, [_vp]),
    ("zxh_dev_upload", C.c_int, [_vp, _vp, _sz]),
    ("zxh_dev_download", C.c_int, [_vp, _vp, _sz]),
    ("zxh_dev_sync", C.c_int, []),
    ("zxh_k_u3_batch", C.c_int, [_vp, C.c_int, _sz, _vp, _sz]),
    ("zxh_k_diag", C.c_int, [_vp, C.c_int, _sz, _u64p, _u64p, _dp, _sz]),
    ("zxh_k_diag_noalloc", C.c_int, [_vp, C.c_int, _sz, _u64p, _u64p, _dp, _sz]),
    ("zxh_k_fused_pass", C.c_int, [_vp, C.c_int, _sz, _vp, _sz, _u64p, _u64p, _dp, _sz]),
    ("zxh_window_table", C.c_int, [_sz, _u64p, _u64p, _dp, _sz, _u64p, C.POINTER(C.c_int), _dp]),
    ("zxh_tile_layer_plan", C.c_int, [C.c_int, _sz, C.c_int, C.POINTER(C.c_uint32), _sz, C.POINTER(C.c_uint32),
                                      C.POINTER(_sz)]),
    ("zxh_k_x", C.c_int, [_vp, C.c_int, _sz, _u64, _u64]),
    ("zxh_k_cx", C.c_int, [_vp, C.c_int, _sz, _u64, _u64, _u64]),
    ("zxh_k_cp", C.c_int, [_vp, C.c_int, _sz, _u64, _u64, C.c_double]),
    ("zxh_k_mcx", C.c_int, [_vp, C.c_int, _sz, _u64, _sz]),
    ("zxh_k_sample", C.c_int, [_vp, C.c_int, _sz, _sz, _u64, _sz, _u64p]),
    ("zxh_k_sample_thresholds", C.c_int, [_vp, C.c_int, _sz, _sz, _dp, _sz, _u64p]),
    ("zxh_k_prob_total", C.c_int, [_vp, C.c_int, _sz, _dp]),
    ("zxh_k_gather", C.c_int, [_vp, C.c_int, _sz, _u64p, _sz, _vp]),
    ("zxh_k_copy", C.c_int, [_vp, _vp, _sz]),
    ("zxh_k_scale", C.c_int, [_vp, C.c_int, _sz, C.c_double, C.c_double]),
    ("zxh_k_expect_z", C.c_int, [_vp, C.c_int, _sz, _i32p, _i32p, _dp, _sz, _dp]),
    ("zxh_k_block5", C.c_int, [_vp, C.c_int, _sz, C.POINTER(C.c_uint32), _dp]),
    ("zxh_k_reset", C.c_int, [_vp, C.c_int, _sz]),
    ("zxh_k_compare", C.c_int, [_vp, C.c_int, _vp, C.c_int, _sz, _dp]),
    ("zxh_batch_set_geometry", C.c_int, [_vp, _sz, C.c_int]),
    ("zxh_k_measure", C.c_int, [_vp, C.c_int, _sz, _sz, _sz, _u64p, C.POINTER(C.c_uint8), C.c_int, _u64, _sz,
                                _u64p, _u64p]),
    ("zxh_dev_realloc", C.c_int, [_vp, _sz, _sz, C.POINTER(_vp)]),
    ("zxh_dev_set_zero", C.c_int, [_vp, C.c_int, _sz, _sz]),
    ("zxh_dev_mem_set", C.c_int, [_vp, C.c_int, C.c_double, C.c_double]),
    ("zxh_ansatz_create", C.c_int, [_sz, C.c_int, _vp, _sz, _sz, C.POINTER(_vp)]),
    ("zxh_ansatz_destroy", C.c_int, [_vp]),
    ("zxh_ansatz_set_observable", C.c_int, [_vp, _i32p, _i32p, _dp, _sz]),
    ("zxh_ansatz_forward", C.c_int, [_vp, _dp, _dp]),
    ("zxh_ansatz_backward", C.c_int, [_vp, _dp, _dp, _dp]),
    ("zxh_ansatz_state", C.c_int, [_vp, C.POINTER(_vp), C.POINTER(_sz)]),
    ("zxh_ansatz_passes", C.c_int, [_vp, C.POINTER(_sz)]),
    ("zxh_ansatz_set_timing", C.c_int, [_vp, C.c_int]),
    ("zxh_ansatz_get_timing", C.c_int, [_vp, C.c_int, C.POINTER(_u64), _dp, C.POINTER(_u64)]),
    ("zxh_ansatz_reset_timing", C.c_int, [_vp]),
    ("zxh_batch_create", C.c_int, [C.c_int, C.POINTER(_vp)]),
    ("zxh_batch_destroy", C.c_int, [_vp]),
    ("zxh_batch_add", C.c_int, [_vp, _sz, _vp, _sz, _sz, _i32p, _i32p, _dp, _sz]),
    ("zxh_batch_finalize", C.c_int, [_vp]),
    ("zxh_batch_run", C.c_int, [_vp, _dp, _dp, _dp]),
    ("zxh_batch_info", C.c_int, [_vp, C.POINTER(_sz), C.POINTER(_sz), C.POINTER(_sz)]),
    ("zxh_batch_set_info", C.c_int, [_vp, _sz, C.POINTER(C.c_uint), C.POINTER(_sz)]),
    ("zxh_k_apply_z", C.c_int, [_vp, _vp, C.c_int, _sz, _i32p, _i32p, _dp, _sz, _dp]),
    ("zxh_tune_set", C.c_int, [C.c_char_p, C.c_int]),
    ("zxh_tune_get", C.c_int, [C.c_char_p, C.POINTER(C.c_int)]),
    ("zxh_k_set_timing", C.c_int, [C.c_int]),
    ("zxh_k_get_timing", C.c_int, [C.c_int, _u64p, _dp, _u64p]),
    ("zxh_k_reset_timing", C.c_int, []),
]

EXPORTED = [s[0] for s in _SIGS]

_lib = None


class ZXHError(RuntimeError):
    """A failure reported by the native engine (the reference aborts here)."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"[zxh code {code}] {msg}")
        self.code = code


def load() -> C.CDLL:
    """Load the native library (no GPU needed to load; calls need one)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"native library not built: {LIB_PATH} (run python -m zxh_adae_amd.build_lib)")
    lib = C.CDLL(LIB_PATH)
    for name, res, args in _SIGS:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int) -> None:
    if rc != ZXH_OK:
        msg = load().zxh_last_error()
        raise ZXHError(rc, msg.decode() if msg else "unknown error")
