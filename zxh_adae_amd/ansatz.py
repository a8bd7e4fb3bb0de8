"""Parameterized circuits on the MI355X engine: expectation values of
diagonal Z / ZZ observables and their adjoint gradients (C-ABI
zxh_ansatz_*; no reference counterpart).

    a = Ansatz(n, ops, nparams, precision="c64")
    a.set_observable([(0.5, 3), (-1.2, 4, 7)])     # 0.5 Z3 - 1.2 Z4 Z7
    e, norm = a.forward(params)
    grads, e = a.backward(params)

ops are (kind, q0, q1, param, value) with kind one of RX, RY, RZ, H, CZ, CP;
param >= 0 takes params[param] as the angle, otherwise value is used.
"""
from __future__ import annotations

import ctypes as C
from typing import Dict, Sequence, Tuple

import numpy as np

from . import _lib
from ._lib import check

RX, RY, RZ, H, CZ, CP = 0, 1, 2, 3, 4, 5


class AnsatzOp(C.Structure):
    _fields_ = [("kind", C.c_int32), ("q0", C.c_int32), ("q1", C.c_int32), ("param", C.c_int32),
                ("value", C.c_double)]


FAMILIES = {"diag": 0, "u3_tile": 1, "reset": 6, "fused": 9, "obs": 10, "adjoint": 11}


class Ansatz:
    def __init__(self, n: int, ops: Sequence[tuple], nparams: int, precision: str = "c64"):
        self.lib = _lib.load()
        self.n = int(n)
        self.nparams = int(nparams)
        self.precision = precision
        arr = (AnsatzOp * max(len(ops), 1))()
        for i, o in enumerate(ops):
            kind, q0, q1, param, value = o
            arr[i].kind, arr[i].q0, arr[i].q1, arr[i].param, arr[i].value = kind, q0, q1, param, value
        h = C.c_void_p()
        prec = _lib.ZXH_C64 if precision in ("c64", "complex64") else _lib.ZXH_C128
        check(self.lib.zxh_ansatz_create(self.n, prec, C.cast(arr, C.c_void_p), len(ops), self.nparams,
                                         C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self.lib.zxh_ansatz_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def set_observable(self, terms: Sequence[tuple]):
        """terms: (coef, i) for coef * Z_i, (coef, i, j) for coef * Z_i Z_j."""
        qi = np.array([int(t[1]) for t in terms], dtype=np.int32)
        qj = np.array([int(t[2]) if len(t) > 2 else -1 for t in terms], dtype=np.int32)
        cf = np.array([float(t[0]) for t in terms], dtype=np.float64)
        self._obs = (qi, qj, cf)
        check(self.lib.zxh_ansatz_set_observable(self._h, qi.ctypes.data_as(_lib._i32p),
                                                 qj.ctypes.data_as(_lib._i32p), cf.ctypes.data_as(_lib._dp),
                                                 len(cf)))

    def _params(self, params) -> np.ndarray:
        p = np.ascontiguousarray(params, dtype=np.float64).reshape(-1)
        if p.size != self.nparams:
            raise ValueError(f"expected {self.nparams} parameters, got {p.size}")
        return p

    def forward(self, params) -> Tuple[float, float]:
        """(<psi|H|psi>, <psi|psi>); the state stays on the device."""
        p = self._params(params)
        out = np.zeros(2)
        check(self.lib.zxh_ansatz_forward(self._h, p.ctypes.data_as(_lib._dp), out.ctypes.data_as(_lib._dp)))
        return float(out[0]), float(out[1])

    def backward(self, params) -> Tuple[np.ndarray, float]:
        """(dE/dparams, E) by the adjoint method."""
        p = self._params(params)
        g = np.zeros(max(self.nparams, 1))
        out = np.zeros(2)
        check(self.lib.zxh_ansatz_backward(self._h, p.ctypes.data_as(_lib._dp), g.ctypes.data_as(_lib._dp),
                                           out.ctypes.data_as(_lib._dp)))
        return g[:self.nparams], float(out[0])

    def state(self) -> np.ndarray:
        """Host copy of the device state (after forward)."""
        ptr = C.c_void_p()
        nb = C.c_size_t()
        check(self.lib.zxh_ansatz_state(self._h, C.byref(ptr), C.byref(nb)))
        dt = np.complex64 if self.precision in ("c64", "complex64") else np.complex128
        out = np.empty(1 << nb.value, dtype=dt)
        check(self.lib.zxh_dev_download(out.ctypes.data, ptr, out.nbytes))
        return out

    def tile_passes(self) -> int:
        v = C.c_size_t()
        check(self.lib.zxh_ansatz_passes(self._h, C.byref(v)))
        return v.value

    def set_timing(self, on: bool):
        check(self.lib.zxh_ansatz_set_timing(self._h, 1 if on else 0))

    def reset_timing(self):
        check(self.lib.zxh_ansatz_reset_timing(self._h))

    def timing(self, family: str) -> Dict[str, float]:
        launches = C.c_uint64()
        ms = C.c_double()
        nbytes = C.c_uint64()
        check(self.lib.zxh_ansatz_get_timing(self._h, FAMILIES[family], C.byref(launches), C.byref(ms),
                                             C.byref(nbytes)))
        return {"launches": launches.value, "ms": ms.value, "bytes": nbytes.value}


def two_local(n: int, layers: int):
    """The two-local ansatz: per layer RY(theta) then RZ(phi) on every qubit,
    then CZ(q, q + 1) for q = 0 .. n - 2.  Returns (ops, nparams)."""
    ops = []
    k = 0
    for _ in range(layers):
        for q in range(n):
            ops.append((RY, q, 0, k, 0.0))
            ops.append((RZ, q, 0, k + 1, 0.0))
            k += 2
        for q in range(n - 1):
            ops.append((CZ, q, q + 1, -1, 0.0))
    return ops, k


class AnsatzBatch:
    """Many parameterized circuits (different widths and depths), evaluated
    together: forward pass p of every set in one persistent launch, batched
    observables, one launch per backward step, replayed from a hipGraph."""

    def __init__(self, precision: str = "c64"):
        self.lib = _lib.load()
        self.precision = precision
        prec = _lib.ZXH_C64 if precision in ("c64", "complex64") else _lib.ZXH_C128
        h = C.c_void_p()
        check(self.lib.zxh_batch_create(prec, C.byref(h)))
        self._h = h
        self.nparams = []

    def close(self):
        if getattr(self, "_h", None):
            self.lib.zxh_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def add(self, n: int, ops: Sequence[tuple], nparams: int, terms: Sequence[tuple]):
        arr = (AnsatzOp * max(len(ops), 1))()
        for i, o in enumerate(ops):
            kind, q0, q1, param, value = o
            arr[i].kind, arr[i].q0, arr[i].q1, arr[i].param, arr[i].value = kind, q0, q1, param, value
        qi = np.array([int(t[1]) for t in terms], dtype=np.int32)
        qj = np.array([int(t[2]) if len(t) > 2 else -1 for t in terms], dtype=np.int32)
        cf = np.array([float(t[0]) for t in terms], dtype=np.float64)
        check(self.lib.zxh_batch_add(self._h, int(n), C.cast(arr, C.c_void_p), len(ops), int(nparams),
                                     qi.ctypes.data_as(_lib._i32p), qj.ctypes.data_as(_lib._i32p),
                                     cf.ctypes.data_as(_lib._dp), len(cf)))
        self.nparams.append(int(nparams))

    def set_geometry(self, n: int, seg: int):
        """Tile geometry (3..6; 0 = the fewest-passes rule) for the sets of n
        qubits added after this call."""
        check(self.lib.zxh_batch_set_geometry(self._h, int(n), int(seg)))

    def finalize(self):
        check(self.lib.zxh_batch_finalize(self._h))

    def run(self, params, gradients: bool = True):
        """params: list of per-set arrays (or one concatenated array).
        Returns (energies[sets], norms[sets], grads list or None)."""
        if isinstance(params, (list, tuple)):
            p = np.concatenate([np.asarray(x, dtype=np.float64).reshape(-1) for x in params])
        else:
            p = np.ascontiguousarray(params, dtype=np.float64).reshape(-1)
        total = sum(self.nparams)
        if p.size != total:
            raise ValueError(f"expected {total} parameters, got {p.size}")
        sets = len(self.nparams)
        e = np.zeros(2 * sets)
        g = np.zeros(max(total, 1)) if gradients else None
        check(self.lib.zxh_batch_run(self._h, p.ctypes.data_as(_lib._dp), e.ctypes.data_as(_lib._dp),
                                     g.ctypes.data_as(_lib._dp) if gradients else None))
        grads = None
        if gradients:
            grads = np.split(g[:total], np.cumsum(self.nparams)[:-1])
        return e[0::2], e[1::2], grads

    def info(self):
        s, t, l = C.c_size_t(), C.c_size_t(), C.c_size_t()
        check(self.lib.zxh_batch_info(self._h, C.byref(s), C.byref(t), C.byref(l)))
        return {"sets": s.value, "params": t.value, "launches": l.value}

    def set_info(self, i: int):
        """(qubits, tile passes per forward) of set i, after finalize."""
        nb, ps = C.c_uint(), C.c_size_t()
        check(self.lib.zxh_batch_set_info(self._h, int(i), C.byref(nb), C.byref(ps)))
        return nb.value, ps.value
