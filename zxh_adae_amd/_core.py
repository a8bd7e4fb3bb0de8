"""Drop-in counterpart of ZXH-Sim's compiled ``_core`` module.

Same module functions and ZXH methods, argument order and result shapes as
the reference's pybind11 module (src/zxh/src/pybind/bindings.cpp:79-118):
``init/finalize/active/rank/nprocs/PrintRes`` and ``ZXH(n, disable_x=False,
eager_expand_all=False)`` with ``set_seed .. num_qubits``.  Every call goes
through the C ABI of libzxh_amd.so to the HIP engine; the reference aborts the
process on errors, this module raises ``ZXHError`` instead.

Extensions (no reference counterpart) are keyword-only or separately named:
``precision="c64"``, ``policy="reference"`` (one HBM pass per flushed batch
instead of the default fused passes), ``get_state_array``, ``sample_real``, ``phys_state``,
``set_phys_state``, timing accessors.
"""
from __future__ import annotations

import ctypes as C
from typing import Dict, List, Sequence

import numpy as np

from . import _lib
from ._lib import ZXHError, check

__all__ = ["init", "finalize", "active", "rank", "nprocs", "PrintRes", "ZXH", "ZXHError"]
__version__ = "0.1.0"

_PREC = {"c128": _lib.ZXH_C128, "complex128": _lib.ZXH_C128, "c64": _lib.ZXH_C64, "complex64": _lib.ZXH_C64}

# kernel family ids of zxh_get_timing
TIMING_FAMILIES = {"diag": 0, "u3_tile": 1, "u3_pair": 2, "perm": 3, "cp": 4, "sample": 5, "reset": 6, "copy": 7,
                   "scale": 8, "fused": 9, "obs": 10, "adjoint": 11, "block": 12, "u3_upper": 13,
                   "tree": 14}
_POLICY = {"fused": _lib.ZXH_POLICY_FUSED, "reference": _lib.ZXH_POLICY_REFERENCE}


def init() -> None:
    check(_lib.load().zxh_init())


def finalize() -> None:
    check(_lib.load().zxh_finalize())


def active() -> bool:
    return bool(_lib.load().zxh_active())


def rank() -> int:
    return int(_lib.load().zxh_rank())


def nprocs() -> int:
    return int(_lib.load().zxh_nprocs())


def PrintRes(res: Sequence[Sequence[bool]], n: int) -> None:
    """Counts per bitstring, qubit n-1 first, sorted by key (zxh.cpp:842-871)."""
    counts: Dict[str, int] = {}
    for item in res:
        s = "".join("1" if (i < len(item) and item[i]) else "0" for i in range(n - 1, -1, -1))
        counts[s] = counts.get(s, 0) + 1
    for k in sorted(counts):
        print(f"{k}: {counts[k]}")


class ZXH:
    def __init__(self, n: int, disable_x: bool = False, eager_expand_all: bool = False, *,
                 precision: str = "c128", policy: str = "fused"):
        if precision not in _PREC:
            raise ValueError(f"unknown precision {precision!r}")
        if policy not in _POLICY:
            raise ValueError(f"unknown policy {policy!r}")
        self._lib = _lib.load()
        self._n = int(n)
        self._prec = _PREC[precision]
        h = C.c_void_p()
        check(self._lib.zxh_create(self._n, int(bool(disable_x)), int(bool(eager_expand_all)), self._prec,
                                   C.byref(h)))
        self._h = h
        check(self._lib.zxh_set_policy(self._h, _POLICY[policy]))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._lib.zxh_destroy(h)
            self._h = None

    # ---- seed and gate recording (zxh.cpp:216-310) ----
    def set_seed(self, seed: int) -> None:
        check(self._lib.zxh_set_seed(self._h, int(seed) & 0xFFFFFFFFFFFFFFFF))

    def clear_seed(self) -> None:
        check(self._lib.zxh_clear_seed(self._h))

    def clear_gates(self) -> None:
        check(self._lib.zxh_clear_gates(self._h))

    def Barrier(self) -> None:
        check(self._lib.zxh_barrier(self._h))

    def Rz(self, q: int, theta: float) -> None:
        check(self._lib.zxh_rz(self._h, q, theta))

    def CRz(self, cq: int, q: int, theta: float) -> None:
        check(self._lib.zxh_crz(self._h, cq, q, theta))

    def CP(self, cq: int, q: int, theta: float) -> None:
        check(self._lib.zxh_cp(self._h, cq, q, theta))

    def P(self, q: int, theta: float) -> None:
        check(self._lib.zxh_p(self._h, q, theta))

    def Z(self, q: int) -> None:
        check(self._lib.zxh_z(self._h, q))

    def X(self, q: int) -> None:
        check(self._lib.zxh_x(self._h, q))

    def CX(self, cq: int, q: int) -> None:
        check(self._lib.zxh_cx(self._h, cq, q))

    def H(self, q: int) -> None:
        check(self._lib.zxh_h(self._h, q))

    def U3(self, q: int, theta: float, lam: float, phi: float) -> None:
        """U3 with the reference's argument order (theta, lambda, phi)."""
        check(self._lib.zxh_u3(self._h, q, theta, lam, phi))

    def Rx(self, q: int, theta: float) -> None:
        check(self._lib.zxh_rx(self._h, q, theta))

    def gate(self, code: int, cq: int = 0, q: int = 0, theta: float = 0.0, lam: float = 0.0,
             phi: float = 0.0) -> None:
        """Record one gate by its code (reference gate_type_t order)."""
        check(self._lib.zxh_gate(self._h, code, cq, q, theta, lam, phi))

    def gates(self, seq) -> None:
        for g in seq:
            self.gate(*g)

    # ---- execution and results ----
    def execute(self) -> None:
        check(self._lib.zxh_execute(self._h))

    def measure(self, shots: int) -> None:
        check(self._lib.zxh_measure(self._h, shots))

    def measured_count(self) -> int:
        out = C.c_size_t()
        check(self._lib.zxh_measured_count(self._h, C.byref(out)))
        return out.value

    def _packed(self, shots: int) -> np.ndarray:
        wc = (self._n + 63) // 64
        buf = np.zeros(max(shots * wc, 1), dtype=np.uint64)
        check(self._lib.zxh_get_results(self._h, buf.ctypes.data_as(_lib._u64p), shots))
        return buf[: shots * wc].reshape(shots, wc)

    def results_bits(self) -> np.ndarray:
        """Last measure() as a (shots, N) uint8 array."""
        shots = self.measured_count()
        packed = self._packed(shots)
        out = np.zeros((shots, self._n), dtype=np.uint8)
        for q in range(self._n):
            out[:, q] = ((packed[:, q >> 6] >> np.uint64(q & 63)) & np.uint64(1)).astype(np.uint8)
        return out

    def get_results(self) -> List[List[bool]]:
        return [[bool(b) for b in row] for row in self.results_bits()]

    def Sampling(self, shots: int) -> List[List[bool]]:
        self.measure(shots)
        return self.get_results()

    def sample_counts(self, shots: int) -> Dict[str, int]:
        """{bitstring (qubit N-1 first): count}, sorted by key (bindings.cpp:36-54)."""
        self.measure(shots)
        bits = self.results_bits()
        counts: Dict[str, int] = {}
        for row in bits:
            s = "".join("1" if row[q] else "0" for q in range(self._n - 1, -1, -1))
            counts[s] = counts.get(s, 0) + 1
        return dict(sorted(counts.items()))

    def get_state_array(self) -> np.ndarray:
        """get_state() as a complex128 numpy array (bit-reversed order, global phase)."""
        out = np.zeros(2 << self._n, dtype=np.float64)
        check(self._lib.zxh_get_state(self._h, out.ctypes.data_as(_lib._dp), out.size))
        return out.view(np.complex128)

    def get_state(self) -> List[complex]:
        return [complex(v) for v in self.get_state_array()]

    def required_M(self) -> int:
        out = C.c_size_t()
        check(self._lib.zxh_required_M(self._h, C.byref(out)))
        return out.value

    def num_qubits(self) -> int:
        return self._n

    # ---- extensions ----
    def set_seg_bits(self, bits: int) -> None:
        check(self._lib.zxh_set_seg_bits(self._h, bits))

    def sample_real(self, shots: int) -> np.ndarray:
        """measure(shots) and return the physical addresses (ascending)."""
        self.measure(shots)
        out = np.zeros(max(shots, 1), dtype=np.uint64)
        check(self._lib.zxh_get_real_results(self._h, out.ctypes.data_as(_lib._u64p), shots))
        return out[:shots]

    def phys_state(self) -> np.ndarray:
        esz = 16 if self._prec == _lib.ZXH_C128 else 8
        buf = np.zeros((2 << self._n) if self._prec == _lib.ZXH_C128 else (1 << self._n) * 2,
                       dtype=np.float64 if self._prec == _lib.ZXH_C128 else np.float32)
        m = C.c_size_t()
        check(self._lib.zxh_get_phys_state(self._h, buf.ctypes.data, buf.nbytes, C.byref(m)))
        n_el = 1 << m.value
        arr = buf[: 2 * n_el]
        return (arr.view(np.complex128) if esz == 16 else arr.view(np.complex64)).copy()

    def set_phys_state(self, psi: np.ndarray) -> None:
        dt = np.complex128 if self._prec == _lib.ZXH_C128 else np.complex64
        a = np.ascontiguousarray(psi, dtype=dt)
        check(self._lib.zxh_set_phys_state(self._h, a.ctypes.data, a.nbytes))

    def device_state(self):
        p = C.c_void_p()
        m = C.c_size_t()
        check(self._lib.zxh_state_device_ptr(self._h, C.byref(p), C.byref(m)))
        return p.value, m.value

    def set_timing(self, on: bool) -> None:
        check(self._lib.zxh_set_timing(self._h, int(bool(on))))

    def reset_timing(self) -> None:
        check(self._lib.zxh_reset_timing(self._h))

    def timing(self, family: str):
        launches = C.c_uint64()
        ms = C.c_double()
        nbytes = C.c_uint64()
        check(self._lib.zxh_get_timing(self._h, TIMING_FAMILIES[family], C.byref(launches), C.byref(ms),
                                       C.byref(nbytes)))
        return {"launches": launches.value, "ms": ms.value, "bytes": nbytes.value}
