"""The no-allocation diagonal sweep (zxh_k_diag_noalloc) against zxh_k_diag.

Both are called through the C ABI on the same input.  Where both run the same
kernel the results are bit-identical; where the workspace makes the
no-allocation call take the polynomial sweep instead of the flat one
(complex128 above 27 qubits), they agree within the rounding bound of the
oracle parity tests (1e-12 complex128).  Two calls in a row, the second on a
smaller state, check that the shared workspace carries nothing over."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NO_SEL = 0xFFFFFFFFFFFFFFFF


def window(rng, n, count):
    s0, s1, th = [], [], []
    for _ in range(count):
        a = 1 << int(rng.integers(n))
        if rng.random() < 0.5:
            s0.append(a)
            s1.append(NO_SEL)
        else:
            b = 1 << int(rng.integers(n))
            while b == a:
                b = 1 << int(rng.integers(n))
            s0.append(a)
            s1.append(b)
        th.append(float(rng.uniform(0, 2 * math.pi)))
    return (np.array(s0, dtype=np.uint64), np.array(s1, dtype=np.uint64), np.array(th))


def last_error():
    from zxh_adae_amd import _lib
    return _lib.load().zxh_last_error().decode()


def call(lib, st, s0, s1, th, noalloc):
    from zxh_adae_amd import _lib
    fn = lib.zxh_k_diag_noalloc if noalloc else lib.zxh_k_diag
    return fn(st.ptr, st.prec, st.nbits, s0.ctypes.data_as(_lib._u64p), s1.ctypes.data_as(_lib._u64p),
              th.ctypes.data_as(_lib._dp), th.size)


def run_pair(gpu_engine, n, prec, count, seed):
    from zxh_adae_amd.kernels import DeviceState
    rng = np.random.default_rng(seed)
    dt = np.complex64 if prec == "c64" else np.complex128
    psi = (rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)).astype(dt)
    s0, s1, th = window(rng, n, count)
    a, b = DeviceState(n, prec), DeviceState(n, prec)
    a.upload(psi)
    b.upload(psi)
    assert call(a.lib, a, s0, s1, th, False) == 0
    assert call(b.lib, b, s0, s1, th, True) == 0, last_error()
    return a.download(), b.download()


@pytest.mark.parametrize("prec", ["c128", "c64"])
def test_noalloc_equals_diag_two_calls_second_smaller(gpu_engine, prec):
    for n, count, seed in ((22, 64, 1), (15, 17, 2)):
        want, got = run_pair(gpu_engine, n, prec, count, seed)
        assert np.array_equal(want, got), (prec, n)


def test_noalloc_general_window(gpu_engine):
    """A window with a three-bit selector (not a polynomial): the per-gate
    sweep, no workspace; bit-identical to zxh_k_diag."""
    from zxh_adae_amd.kernels import DeviceState
    n = 16
    rng = np.random.default_rng(5)
    psi = (rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)).astype(np.complex128)
    s0 = np.array([0b111 << 3, 1 << 2], dtype=np.uint64)
    s1 = np.array([NO_SEL, 1 << 9], dtype=np.uint64)
    th = np.array([0.7, 1.9])
    a, b = DeviceState(n, "c128"), DeviceState(n, "c128")
    a.upload(psi)
    b.upload(psi)
    assert call(a.lib, a, s0, s1, th, False) == 0
    assert call(b.lib, b, s0, s1, th, True) == 0
    assert np.array_equal(a.download(), b.download())


@pytest.mark.parametrize("n", [30])
def test_noalloc_at_the_widest_state_then_narrow(gpu_engine, n):
    """30 qubits complex128: the polynomial sweep in the workspace against the
    flat sweep of zxh_k_diag (checked on gathered amplitudes), then a narrow
    call that must equal zxh_k_diag exactly."""
    from zxh_adae_amd.kernels import DeviceState
    rng = np.random.default_rng(30)
    s0, s1, th = window(rng, n, 48)
    from zxh_adae_amd import workloads as W
    from zxh_adae_amd.kernels import set_basis_zero
    hall = [((1 << q), (1 << q), W.u3_matrix(math.pi / 2, math.pi, 0.0)) for q in range(n)]
    a, b = DeviceState(n, "c128"), DeviceState(n, "c128")
    for st in (a, b):  # uniform superposition: H on every qubit of |0...0>
        set_basis_zero(st)
        st.u3_batch(hall)
    idx = np.unique(rng.integers(0, 1 << n, size=4096, dtype=np.uint64))
    assert call(a.lib, a, s0, s1, th, False) == 0
    assert call(b.lib, b, s0, s1, th, True) == 0, last_error()
    ga, gb = a.gather(idx), b.gather(idx)
    assert np.max(np.abs(ga - gb)) < 1e-12 * 2.0 ** (-n / 2) * 64
    a.free()
    b.free()
    want, got = run_pair(gpu_engine, 14, "c128", 9, 31)
    assert np.array_equal(want, got)


def test_noalloc_rejects_wider_state(gpu_engine):
    """31 qubits complex128 (32 GiB): the coefficients do not fit the
    workspace; the call fails with ZXH_EINVAL before any launch."""
    from zxh_adae_amd.kernels import DeviceState
    n = 31
    st = DeviceState(n, "c128")
    s0 = np.array([1 << 30], dtype=np.uint64)
    s1 = np.array([1], dtype=np.uint64)
    th = np.array([0.5])
    rc = call(st.lib, st, s0, s1, th, True)
    st.free()
    assert rc == 1
    assert "workspace" in last_error()
