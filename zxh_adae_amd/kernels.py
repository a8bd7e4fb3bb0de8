"""Kernel-level access to the HIP passes over a device state buffer.

Mirrors the reference's internal backend API (src/zxh/include/zxhsim/kernels.h
and the sampler of measure.h) through the C ABI's zxh_k_* entry points.  Used
by the per-kernel parity tests and the benchmark.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import numpy as np

from . import _lib
from ._lib import check

NO_SEL = 0xFFFFFFFFFFFFFFFF


class U3Desc(C.Structure):
    _fields_ = [("packed_sel", C.c_uint64), ("delta", C.c_uint64), ("u", C.c_double * 8)]


def pack_sel(alpha: int, beta: bool) -> int:
    return (alpha & ((1 << 63) - 1)) | ((1 << 63) if beta else 0)


class DeviceState:
    """A device buffer of 2^nbits amplitudes (complex128 or complex64)."""

    def __init__(self, nbits: int, precision: str = "c128"):
        self.lib = _lib.load()
        self.nbits = int(nbits)
        self.prec = _lib.ZXH_C128 if precision in ("c128", "complex128") else _lib.ZXH_C64
        self.dtype = np.complex128 if self.prec == _lib.ZXH_C128 else np.complex64
        self.nbytes = (1 << self.nbits) * np.dtype(self.dtype).itemsize
        p = C.c_void_p()
        check(self.lib.zxh_dev_alloc(self.nbytes, C.byref(p)))
        self.ptr = p

    @classmethod
    def wrap(cls, ptr: int, nbits: int, precision: str = "c128") -> "DeviceState":
        """A non-owning view of an existing device state (e.g. a simulator's)."""
        self = cls.__new__(cls)
        self.lib = _lib.load()
        self.nbits = int(nbits)
        self.prec = _lib.ZXH_C128 if precision in ("c128", "complex128") else _lib.ZXH_C64
        self.dtype = np.complex128 if self.prec == _lib.ZXH_C128 else np.complex64
        self.nbytes = (1 << self.nbits) * np.dtype(self.dtype).itemsize
        self.ptr = C.c_void_p(ptr)
        self._owned = False
        return self

    def free(self):
        if getattr(self, "ptr", None) and getattr(self, "_owned", True):
            self.lib.zxh_dev_free(self.ptr)
        self.ptr = None

    def __del__(self):
        self.free()

    def upload(self, psi: np.ndarray):
        a = np.ascontiguousarray(psi, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        check(self.lib.zxh_dev_upload(self.ptr, a.ctypes.data, a.nbytes))

    def gather(self, indices) -> np.ndarray:
        """The amplitudes at the given physical addresses (no full download)."""
        idx = np.ascontiguousarray(indices, dtype=np.uint64).reshape(-1)
        out = np.empty(idx.size, dtype=self.dtype)
        check(self.lib.zxh_k_gather(self.ptr, self.prec, self.nbits, idx.ctypes.data_as(_lib._u64p), idx.size,
                                    out.ctypes.data))
        return out

    def download(self) -> np.ndarray:
        out = np.empty(1 << self.nbits, dtype=self.dtype)
        check(self.lib.zxh_dev_download(out.ctypes.data, self.ptr, self.nbytes))
        return out

    # ---- passes ----
    def u3_batch(self, descs: Sequence[tuple]):
        """descs: (packed_sel, delta, 2x2 complex matrix) per gate, distinct targets."""
        arr = (U3Desc * max(len(descs), 1))()
        for i, (sel, delta, u) in enumerate(descs):
            arr[i].packed_sel = sel
            arr[i].delta = delta
            flat = np.asarray(u, dtype=np.complex128).reshape(4)
            for k in range(4):
                arr[i].u[2 * k] = flat[k].real
                arr[i].u[2 * k + 1] = flat[k].imag
        check(self.lib.zxh_k_u3_batch(self.ptr, self.prec, self.nbits, C.cast(arr, C.c_void_p), len(descs)))

    @staticmethod
    def _descs(descs: Sequence[tuple]):
        arr = (U3Desc * max(len(descs), 1))()
        for i, (sel, delta, u) in enumerate(descs):
            arr[i].packed_sel = sel
            arr[i].delta = delta
            flat = np.asarray(u, dtype=np.complex128).reshape(4)
            for k in range(4):
                arr[i].u[2 * k] = flat[k].real
                arr[i].u[2 * k + 1] = flat[k].imag
        return arr

    def fused_pass(self, descs: Sequence[tuple], sel0: Sequence[int], sel1: Sequence[int], theta: Sequence[float]):
        """U3 descriptors then a diagonal window, in one tile pass."""
        arr = self._descs(descs)
        s0 = np.ascontiguousarray(sel0, dtype=np.uint64)
        s1 = np.ascontiguousarray(sel1, dtype=np.uint64)
        th = np.ascontiguousarray(theta, dtype=np.float64)
        check(self.lib.zxh_k_fused_pass(self.ptr, self.prec, self.nbits, C.cast(arr, C.c_void_p), len(descs),
                                        s0.ctypes.data_as(_lib._u64p), s1.ctypes.data_as(_lib._u64p),
                                        th.ctypes.data_as(_lib._dp), th.size))

    @staticmethod
    def _terms(terms: Sequence[tuple]):
        """terms: (coef, i) for Z_i or (coef, i, j) for Z_i Z_j."""
        qi = np.array([int(t[1]) for t in terms], dtype=np.int32)
        qj = np.array([int(t[2]) if len(t) > 2 else -1 for t in terms], dtype=np.int32)
        cf = np.array([float(t[0]) for t in terms], dtype=np.float64)
        return qi, qj, cf

    def expect_z(self, terms: Sequence[tuple]):
        """(<psi|H|psi>, <psi|psi>) for H = sum of coef * Z_i [Z_j], FP64 sums."""
        qi, qj, cf = self._terms(terms)
        out = np.zeros(2)
        check(self.lib.zxh_k_expect_z(self.ptr, self.prec, self.nbits, qi.ctypes.data_as(_lib._i32p),
                                      qj.ctypes.data_as(_lib._i32p), cf.ctypes.data_as(_lib._dp), len(cf),
                                      out.ctypes.data_as(_lib._dp)))
        return float(out[0]), float(out[1])

    def apply_z(self, dst: "DeviceState", terms: Sequence[tuple]):
        """dst = H psi; returns the same sums as expect_z."""
        assert dst.nbits == self.nbits and dst.prec == self.prec
        qi, qj, cf = self._terms(terms)
        out = np.zeros(2)
        check(self.lib.zxh_k_apply_z(self.ptr, dst.ptr, self.prec, self.nbits, qi.ctypes.data_as(_lib._i32p),
                                     qj.ctypes.data_as(_lib._i32p), cf.ctypes.data_as(_lib._dp), len(cf),
                                     out.ctypes.data_as(_lib._dp)))
        return float(out[0]), float(out[1])

    def block5(self, targets: Sequence[int], u: np.ndarray):
        """Dense 5-qubit unitary u (32x32, bit b of its index = targets[b])."""
        t = np.ascontiguousarray(targets, dtype=np.uint32)
        assert t.size == 5
        m = np.ascontiguousarray(np.asarray(u, dtype=np.complex128).reshape(32, 32)).view(np.float64)
        check(self.lib.zxh_k_block5(self.ptr, self.prec, self.nbits, t.ctypes.data_as(C.POINTER(C.c_uint32)),
                                    m.ctypes.data_as(_lib._dp)))

    def diag(self, sel0: Sequence[int], sel1: Sequence[int], theta: Sequence[float]):
        s0 = np.ascontiguousarray(sel0, dtype=np.uint64)
        s1 = np.ascontiguousarray(sel1, dtype=np.uint64)
        th = np.ascontiguousarray(theta, dtype=np.float64)
        check(self.lib.zxh_k_diag(self.ptr, self.prec, self.nbits, s0.ctypes.data_as(_lib._u64p),
                                  s1.ctypes.data_as(_lib._u64p), th.ctypes.data_as(_lib._dp), len(th)))

    def x(self, delta: int, sel: int):
        check(self.lib.zxh_k_x(self.ptr, self.prec, self.nbits, delta, sel))

    def cx(self, delta: int, sel_cq: int, sel_q: int):
        check(self.lib.zxh_k_cx(self.ptr, self.prec, self.nbits, delta, sel_cq, sel_q))

    def cp(self, sel_cq: int, sel_q: int, theta: float):
        check(self.lib.zxh_k_cp(self.ptr, self.prec, self.nbits, sel_cq, sel_q, theta))

    def mcx(self, ctrl_mask: int, target: int):
        check(self.lib.zxh_k_mcx(self.ptr, self.prec, self.nbits, ctrl_mask, target))

    def sample(self, seg_bits: int, stream_seed: int, shots: int) -> np.ndarray:
        out = np.zeros(max(shots, 1), dtype=np.uint64)
        check(self.lib.zxh_k_sample(self.ptr, self.prec, self.nbits, seg_bits, stream_seed, shots,
                                    out.ctypes.data_as(_lib._u64p)))
        return out[:shots]

    def measure(self, seg_bits: int, rows, b, shots: int, seed=None):
        """measure() + get_results() of the reference on this state with the
        affine map (rows[q]: bit c = A[q][c]; b[q]): returns (packed words,
        ceil(nq/64) per shot, and the real addresses).  seed None draws the
        stream seed from random_device, as the reference does without one."""
        r = np.ascontiguousarray(rows, dtype=np.uint64)
        bb = np.ascontiguousarray(b, dtype=np.uint8)
        nq = r.size
        wc = (nq + 63) // 64
        packed = np.zeros(max(shots * wc, 1), dtype=np.uint64)
        real = np.zeros(max(shots, 1), dtype=np.uint64)
        check(self.lib.zxh_k_measure(self.ptr, self.prec, self.nbits, seg_bits, nq, r.ctypes.data_as(_lib._u64p),
                                     bb.ctypes.data_as(C.POINTER(C.c_uint8)), int(seed is not None),
                                     0 if seed is None else int(seed), shots, packed.ctypes.data_as(_lib._u64p),
                                     real.ctypes.data_as(_lib._u64p)))
        return packed[: shots * wc].reshape(shots, wc), real[:shots]

    def sample_thresholds(self, seg_bits: int, thresholds: np.ndarray) -> np.ndarray:
        thr = np.ascontiguousarray(thresholds, dtype=np.float64)
        out = np.zeros(max(thr.size, 1), dtype=np.uint64)
        check(self.lib.zxh_k_sample_thresholds(self.ptr, self.prec, self.nbits, seg_bits,
                                               thr.ctypes.data_as(_lib._dp), thr.size,
                                               out.ctypes.data_as(_lib._u64p)))
        return out[: thr.size]

    def scale(self, re: float, im: float = 0.0):
        check(self.lib.zxh_k_scale(self.ptr, self.prec, self.nbits, re, im))

    def copy_from(self, other: "DeviceState"):
        assert other.nbytes == self.nbytes
        check(self.lib.zxh_k_copy(self.ptr, other.ptr, self.nbytes))

    def prob_total(self) -> float:
        out = C.c_double()
        check(self.lib.zxh_k_prob_total(self.ptr, self.prec, self.nbits, C.byref(out)))
        return out.value

    def compare(self, other: "DeviceState") -> dict:
        """This state against another of the same width (either precision),
        one read of each: overlap <self|other>, both squared norms, the
        infidelity 1 - |<self|other>| and max_i |self_i - other_i|."""
        assert other.nbits == self.nbits
        out = np.zeros(5)
        check(self.lib.zxh_k_compare(self.ptr, self.prec, other.ptr, other.prec, self.nbits,
                                     out.ctypes.data_as(_lib._dp)))
        ov = complex(out[0], out[1])
        return {"overlap": ov, "norm2_self": float(out[2]), "norm2_other": float(out[3]),
                "infidelity": 1.0 - abs(ov), "max_amp_error": float(out[4])}


def sync():
    check(_lib.load().zxh_dev_sync())


FAMILIES = {"diag": 0, "u3_tile": 1, "u3_pair": 2, "perm": 3, "cp": 4, "sample": 5, "reset": 6, "copy": 7,
            "scale": 8, "fused": 9, "obs": 10, "adjoint": 11, "block": 12, "u3_upper": 13, "tree": 14}


def set_timing(on: bool):
    check(_lib.load().zxh_k_set_timing(int(bool(on))))


def reset_timing():
    check(_lib.load().zxh_k_reset_timing())


def timing(family: str) -> dict:
    """Launch count, total milliseconds (HIP events on the launch stream) and
    algorithmic bytes of the kernel-level calls of one family."""
    launches, ms, nbytes = C.c_uint64(), C.c_double(), C.c_uint64()
    check(_lib.load().zxh_k_get_timing(FAMILIES[family], C.byref(launches), C.byref(ms), C.byref(nbytes)))
    return {"launches": launches.value, "ms": ms.value, "bytes": nbytes.value}


def apply_layer(st: DeviceState, layer: dict):
    """One layer of workloads.c5_layers on a device state: the dense blocks,
    the diagonal batch as one sweep, then the multi-controlled X cascade."""
    for targets, u in layer["blocks"]:
        st.block5(targets, u)
    s0, s1, th = [], [], []
    for code, cq, q, theta in layer["diag"]:
        if code == 4:  # P
            s0.append(1 << q)
            s1.append(NO_SEL)
        else:  # CP
            s0.append(1 << cq)
            s1.append(1 << q)
        th.append(theta)
    if th:
        st.diag(s0, s1, th)
    for mask, t in layer["mcx"]:
        st.mcx(mask, t)


def set_basis_zero(st: DeviceState):
    """|0...0> in place (zero fill, then amplitude 0 = 1)."""
    check(st.lib.zxh_k_reset(st.ptr, st.prec, st.nbits))


def mix_seed(base: int, salt: int = 0) -> int:
    """SplitMix64 finaliser of base + golden gamma + salt: the sampler's
    threshold-stream seed (measure_general.cpp:16-22)."""
    m = 0xFFFFFFFFFFFFFFFF
    z = (base + 0x9E3779B97F4A7C15 + salt) & m
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
    return z ^ (z >> 31)
