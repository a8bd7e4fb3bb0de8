"""ctypes bindings for the oracle (oracle/liboracle.so) and the reference
checker build (oracle/_ref/libzxhref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, always as the checker (or the timed CPU
baseline), never as the product path.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import List, Optional, Sequence, Tuple

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libzxhref.so")

# gate codes (reference gate_type_t order, plus RX as a recording alias)
BARRIER, Z, RZ, P, CRZ, CP, X, CX, H, U3, RX = 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11

_u64 = C.c_uint64
_sz = C.c_size_t
_dp = C.POINTER(C.c_double)
_u64p = C.POINTER(C.c_uint64)


class U3Desc(C.Structure):
    _fields_ = [("packed_sel", C.c_uint64), ("delta", C.c_uint64), ("u", C.c_double * 8)]


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def _u64ptr(a: np.ndarray):
    return a.ctypes.data_as(_u64p)


_oracle = None


def oracle_lib() -> C.CDLL:
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"oracle library missing: {ORACLE_SO} (run make -C oracle)")
        lib = C.CDLL(ORACLE_SO)
        lib.zo_create.restype = C.c_void_p
        lib.zo_create.argtypes = [_sz, C.c_int, C.c_int]
        lib.zo_destroy.argtypes = [C.c_void_p]
        lib.zo_set_seed.argtypes = [C.c_void_p, _u64]
        lib.zo_clear_seed.argtypes = [C.c_void_p]
        lib.zo_clear_gates.argtypes = [C.c_void_p]
        lib.zo_add_gate.argtypes = [C.c_void_p, C.c_int, _sz, _sz, C.c_double, C.c_double, C.c_double]
        lib.zo_execute.argtypes = [C.c_void_p]
        lib.zo_measure.argtypes = [C.c_void_p, _sz]
        lib.zo_measured_count.restype = _sz
        lib.zo_measured_count.argtypes = [C.c_void_p]
        lib.zo_get_results.argtypes = [C.c_void_p, _u64p, _sz]
        lib.zo_get_real_results.argtypes = [C.c_void_p, _u64p, _sz]
        lib.zo_get_state.argtypes = [C.c_void_p, _dp]
        lib.zo_get_phys_state.argtypes = [C.c_void_p, _dp, _sz, C.POINTER(_sz)]
        lib.zo_required_M.restype = _sz
        lib.zo_required_M.argtypes = [C.c_void_p]
        lib.zo_num_qubits.restype = _sz
        lib.zo_num_qubits.argtypes = [C.c_void_p]
        lib.zo_global_phase.argtypes = [C.c_void_p, _dp, _dp]
        lib.zo_set_seg_min_i.argtypes = [C.c_void_p, _sz]
        lib.zo_get_affine.argtypes = [C.c_void_p, _u64p, C.POINTER(C.c_uint8), C.POINTER(_sz)]
        lib.zo_mix_seed.restype = _u64
        lib.zo_mix_seed.argtypes = [_u64, _u64]
        lib.zo_mt64_nth.restype = _u64
        lib.zo_mt64_nth.argtypes = [_u64, _sz]
        lib.zo_thresholds.argtypes = [_u64, _sz, C.c_double, C.c_double, _dp]
        lib.zo_sample_state.argtypes = [_dp, _sz, _sz, _u64, _sz, _u64p]
        lib.zo_sample_thresholds.argtypes = [_dp, _sz, _sz, _dp, _sz, _u64p]
        lib.zo_prob_total.restype = C.c_double
        lib.zo_prob_total.argtypes = [_dp, _sz]
        lib.zo_k_u3_batch.argtypes = [_dp, _sz, C.POINTER(U3Desc), _sz]
        lib.zo_k_x.argtypes = [_dp, _sz, _u64, _u64]
        lib.zo_k_cx.argtypes = [_dp, _sz, _u64, _u64, _u64]
        lib.zo_k_cp.argtypes = [_dp, _sz, _u64, _u64, C.c_double]
        lib.zo_k_diag.argtypes = [_dp, _sz, _u64p, _u64p, _dp, _sz]
        lib.zo_u3_matrix.argtypes = [C.c_double, C.c_double, C.c_double, _dp]
        _oracle = lib
    return _oracle


def ref_available() -> bool:
    return os.path.exists(REF_SO)


_ref = None


def ref_lib() -> C.CDLL:
    global _ref
    if _ref is None:
        if not ref_available():
            raise RuntimeError(f"reference checker build missing: {REF_SO} (run make -C oracle ref)")
        lib = C.CDLL(REF_SO)
        lib.zr_create.restype = C.c_void_p
        lib.zr_create.argtypes = [_sz, C.c_int, C.c_int]
        lib.zr_destroy.argtypes = [C.c_void_p]
        lib.zr_set_seed.argtypes = [C.c_void_p, _u64]
        lib.zr_gate.argtypes = [C.c_void_p, C.c_int, _sz, _sz, C.c_double, C.c_double, C.c_double]
        lib.zr_execute.argtypes = [C.c_void_p]
        lib.zr_required_M.restype = _sz
        lib.zr_required_M.argtypes = [C.c_void_p]
        lib.zr_get_state.argtypes = [C.c_void_p, _dp]
        lib.zr_sampling.argtypes = [C.c_void_p, _sz, C.POINTER(C.c_uint8)]
        lib.zr_sample_state.argtypes = [_dp, _sz, _u64, _sz, _u64p]
        _ref = lib
    return _ref


Gate = Tuple  # (type, cq, q, theta, lambda, phi)


class OracleSim:
    """The oracle's ZXH object (restatement of the reference's CPU path)."""

    def __init__(self, n: int, disable_x: bool = False, eager_expand_all: bool = False,
                 seg_min_i: Optional[int] = None):
        self.lib = oracle_lib()
        self.n = n
        self.h = self.lib.zo_create(n, int(disable_x), int(eager_expand_all))
        if seg_min_i is not None:
            self.lib.zo_set_seg_min_i(self.h, seg_min_i)

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.zo_destroy(self.h)
            self.h = None

    def gate(self, t, cq=0, q=0, theta=0.0, lam=0.0, phi=0.0):
        rc = self.lib.zo_add_gate(self.h, t, cq, q, theta, lam, phi)
        if rc:
            raise ValueError(f"oracle gate rejected (code {rc})")

    def gates(self, gs: Sequence[Gate]):
        for g in gs:
            self.gate(*g)

    def set_seed(self, s):
        self.lib.zo_set_seed(self.h, s)

    def execute(self):
        rc = self.lib.zo_execute(self.h)
        if rc:
            raise RuntimeError(f"oracle execute failed ({rc})")

    def get_state(self) -> np.ndarray:
        out = np.zeros(2 << self.n, dtype=np.float64)
        rc = self.lib.zo_get_state(self.h, _dptr(out))
        if rc:
            raise RuntimeError(f"oracle get_state failed ({rc})")
        return out.view(np.complex128)

    def phys_state(self) -> np.ndarray:
        cap = 1 << self.lib.zo_required_M(self.h) if False else None
        m = _sz(0)
        # physical width is at most N
        buf = np.zeros(2 << self.n, dtype=np.float64)
        rc = self.lib.zo_get_phys_state(self.h, _dptr(buf), 1 << self.n, C.byref(m))
        if rc:
            raise RuntimeError(f"oracle phys state failed ({rc})")
        return buf[: 2 << m.value].view(np.complex128).copy()

    def global_phase(self) -> complex:
        re, im = C.c_double(), C.c_double()
        self.lib.zo_global_phase(self.h, C.byref(re), C.byref(im))
        return complex(re.value, im.value)

    def affine(self):
        rows = np.zeros(max(self.n, 1), dtype=np.uint64)
        b = np.zeros(max(self.n, 1), dtype=np.uint8)
        m = _sz(0)
        self.lib.zo_get_affine(self.h, _u64ptr(rows), b.ctypes.data_as(C.POINTER(C.c_uint8)), C.byref(m))
        return rows[: self.n], b[: self.n], m.value

    def measure(self, shots: int):
        rc = self.lib.zo_measure(self.h, shots)
        if rc:
            raise RuntimeError(f"oracle measure failed ({rc})")

    def real_results(self, shots: int) -> np.ndarray:
        out = np.zeros(max(shots, 1), dtype=np.uint64)
        rc = self.lib.zo_get_real_results(self.h, _u64ptr(out), shots)
        if rc:
            raise RuntimeError(f"oracle results failed ({rc})")
        return out[:shots]

    def sampling_bits(self, shots: int) -> np.ndarray:
        """Sampling(shots) as a (shots, N) uint8 array of virtual bits."""
        self.measure(shots)
        wc = (self.n + 63) // 64
        packed = np.zeros(max(shots * wc, 1), dtype=np.uint64)
        rc = self.lib.zo_get_results(self.h, _u64ptr(packed), shots)
        if rc:
            raise RuntimeError(f"oracle results failed ({rc})")
        packed = packed[: shots * wc].reshape(shots, wc)
        out = np.zeros((shots, self.n), dtype=np.uint8)
        for q in range(self.n):
            out[:, q] = (packed[:, q >> 6] >> np.uint64(q & 63)) & np.uint64(1)
        return out

    def required_M(self) -> int:
        return self.lib.zo_required_M(self.h)


class RefSim:
    """The compiled reference's ZXH object (checker in this container only)."""

    def __init__(self, n: int, disable_x: bool = False, eager_expand_all: bool = False):
        self.lib = ref_lib()
        self.n = n
        self.h = self.lib.zr_create(n, int(disable_x), int(eager_expand_all))

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.zr_destroy(self.h)
            self.h = None

    def gate(self, t, cq=0, q=0, theta=0.0, lam=0.0, phi=0.0):
        self.lib.zr_gate(self.h, t, cq, q, theta, lam, phi)

    def gates(self, gs):
        for g in gs:
            self.gate(*g)

    def set_seed(self, s):
        self.lib.zr_set_seed(self.h, s)

    def execute(self):
        self.lib.zr_execute(self.h)

    def get_state(self) -> np.ndarray:
        out = np.zeros(2 << self.n, dtype=np.float64)
        self.lib.zr_get_state(self.h, _dptr(out))
        return out.view(np.complex128)

    def sampling_bits(self, shots: int) -> np.ndarray:
        out = np.zeros(max(shots * self.n, 1), dtype=np.uint8)
        self.lib.zr_sampling(self.h, shots, out.ctypes.data_as(C.POINTER(C.c_uint8)))
        return out[: shots * self.n].reshape(shots, self.n)

    def required_M(self) -> int:
        return self.lib.zr_required_M(self.h)


def oracle_sample_state(psi: np.ndarray, seg_bits: int, stream_seed: int, shots: int) -> np.ndarray:
    lib = oracle_lib()
    psi = np.ascontiguousarray(psi, dtype=np.complex128)
    m = int(psi.size).bit_length() - 1
    assert psi.size == 1 << m
    out = np.zeros(max(shots, 1), dtype=np.uint64)
    rc = lib.zo_sample_state(_dptr(psi.view(np.float64)), m, seg_bits, stream_seed, shots, _u64ptr(out))
    if rc:
        raise RuntimeError(f"oracle sampler failed ({rc})")
    return out[:shots]


def oracle_sample_thresholds(psi: np.ndarray, seg_bits: int, thr: np.ndarray) -> np.ndarray:
    """The reference's segment walk on caller-given sorted thresholds."""
    lib = oracle_lib()
    psi = np.ascontiguousarray(psi, dtype=np.complex128)
    thr = np.ascontiguousarray(thr, dtype=np.float64)
    m = int(psi.size).bit_length() - 1
    out = np.zeros(max(thr.size, 1), dtype=np.uint64)
    rc = lib.zo_sample_thresholds(_dptr(psi.view(np.float64)), m, seg_bits, _dptr(thr), thr.size, _u64ptr(out))
    if rc:
        raise RuntimeError(f"oracle sampler failed ({rc})")
    return out[: thr.size]


def prob_total(psi: np.ndarray) -> float:
    psi = np.ascontiguousarray(psi, dtype=np.complex128)
    m = int(psi.size).bit_length() - 1
    return float(oracle_lib().zo_prob_total(_dptr(psi.view(np.float64)), m))


def ref_sample_state(psi: np.ndarray, seed: int, shots: int) -> np.ndarray:
    """Reference sampler on the given state (seed is the user seed; the
    reference mixes it).  Segment size follows ZXHSIM_SEG_MIN_I of this process."""
    lib = ref_lib()
    psi = np.ascontiguousarray(psi, dtype=np.complex128)
    m = int(psi.size).bit_length() - 1
    out = np.zeros(max(shots, 1), dtype=np.uint64)
    lib.zr_sample_state(_dptr(psi.view(np.float64)), m, seed, shots, _u64ptr(out))
    return out[:shots]


def mix_seed(base: int, salt: int = 0) -> int:
    return int(oracle_lib().zo_mix_seed(base, salt))


def thresholds(stream_seed: int, cnt: int, begin: float, end: float) -> np.ndarray:
    out = np.zeros(max(cnt, 1), dtype=np.float64)
    rc = oracle_lib().zo_thresholds(stream_seed, cnt, begin, end, _dptr(out))
    if rc:
        raise RuntimeError("threshold stream rejected")
    return out[:cnt]


def u3_matrix(theta: float, lam: float, phi: float) -> np.ndarray:
    out = np.zeros(8, dtype=np.float64)
    oracle_lib().zo_u3_matrix(theta, lam, phi, _dptr(out))
    return out.view(np.complex128).reshape(2, 2)
