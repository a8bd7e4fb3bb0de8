"""Diagonal Z / ZZ observables on the GPU: expectation values, compensated
norms and H|psi>.  There is no reference counterpart (the reference only
samples); the checks are exact FP64 sums (math.fsum over the same amplitudes),
closed forms of product states, and the observable applied elementwise."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rand_state(rng, n, dtype=np.complex128):
    v = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    v /= np.linalg.norm(v)
    return v.astype(dtype)


def rand_terms(rng, n, count):
    terms = []
    for _ in range(count):
        c = float(rng.normal())
        if rng.random() < 0.5:
            terms.append((c, int(rng.integers(n))))
        else:
            i, j = rng.choice(n, 2, replace=False)
            terms.append((c, int(i), int(j)))
    return terms


def h_values(n, terms):
    x = np.arange(1 << n, dtype=np.int64)
    h = np.zeros(1 << n)
    for t in terms:
        zi = 1.0 - 2.0 * ((x >> t[1]) & 1)
        if len(t) > 2:
            zi = zi * (1.0 - 2.0 * ((x >> t[2]) & 1))
        h += t[0] * zi
    return h


def exact_sums(psi, h):
    p = psi.real.astype(np.float64) ** 2 + psi.imag.astype(np.float64) ** 2
    return math.fsum(p * h), math.fsum(p)


@pytest.mark.parametrize("n", [3, 8, 13, 14, 15, 16, 19])
@pytest.mark.parametrize("prec", ["c128", "c64"])
def test_expectation_matches_exact_sum(gpu_engine, n, prec):
    """Small-state path (n below one block: 14 bits c128, 15 bits c64) and the
    block path; complex64 amplitudes are squared and summed in FP64."""
    from zxh_adae_amd.kernels import DeviceState
    rng = np.random.default_rng(100 + n)
    psi = rand_state(rng, n, np.complex64 if prec == "c64" else np.complex128)
    terms = rand_terms(rng, n, 40)
    d = DeviceState(n, prec)
    d.upload(psi)
    e, nrm = d.expect_z(terms)
    e_ref, n_ref = exact_sums(psi, h_values(n, terms))
    scale = sum(abs(t[0]) for t in terms)
    assert abs(e - e_ref) <= 1e-13 * scale
    assert abs(nrm - n_ref) <= 1e-14


@pytest.mark.parametrize("n", [10, 16, 20])
@pytest.mark.parametrize("prec", ["c128", "c64"])
def test_apply_observable_elementwise(gpu_engine, n, prec):
    from zxh_adae_amd.kernels import DeviceState
    rng = np.random.default_rng(7 * n)
    dt = np.complex64 if prec == "c64" else np.complex128
    psi = rand_state(rng, n, dt)
    terms = rand_terms(rng, n, 25)
    a = DeviceState(n, prec)
    b = DeviceState(n, prec)
    a.upload(psi)
    e, _ = a.apply_z(b, terms)
    h = h_values(n, terms)
    want = psi.astype(np.complex128) * h
    got = b.download().astype(np.complex128)
    tol = 1e-13 if prec == "c128" else 2e-6 * np.max(np.abs(h))
    assert np.max(np.abs(got - want)) <= tol
    assert abs(e - exact_sums(psi, h)[0]) <= 1e-13 * sum(abs(t[0]) for t in terms)


def test_product_state_closed_form(gpu_engine):
    """RY(theta_q) on every qubit of |0..0>: <Z_q> = cos theta_q and
    <Z_i Z_j> = cos theta_i cos theta_j."""
    from zxh_adae_amd.kernels import DeviceState
    n = 16
    rng = np.random.default_rng(3)
    th = rng.uniform(-math.pi, math.pi, n)
    psi = np.array([1.0 + 0j])
    for q in range(n):
        v = np.array([math.cos(th[q] / 2), math.sin(th[q] / 2)], dtype=np.complex128)
        psi = np.kron(v, psi)  # qubit q is address bit q
    terms = rand_terms(rng, n, 60)
    want = 0.0
    for t in terms:
        want += t[0] * math.cos(th[t[1]]) * (math.cos(th[t[2]]) if len(t) > 2 else 1.0)
    d = DeviceState(n, "c128")
    d.upload(psi)
    e, nrm = d.expect_z(terms)
    assert abs(e - want) < 1e-12
    assert abs(nrm - 1.0) < 1e-13


def test_compensated_norm_of_wide_dynamic_range(gpu_engine):
    """One amplitude of 1 and 2^20 amplitudes of 1e-5 in complex64: a plain
    FP32 sum would lose the small terms entirely; the FP64 compensated sum
    matches math.fsum."""
    from zxh_adae_amd.kernels import DeviceState
    n = 20
    psi = np.full(1 << n, 1e-5, dtype=np.complex64)
    psi[12345] = 1.0
    d = DeviceState(n, "c64")
    d.upload(psi)
    _, nrm = d.expect_z([(1.0, 0)])
    want = math.fsum((psi.real.astype(np.float64)) ** 2)
    assert abs(nrm - want) <= 1e-15 * want
    assert nrm - 1.0 > 1.0e-4  # the 2^20 small terms (about 1.05e-4) are all there


def test_observable_rejects_bad_qubits(gpu_engine):
    from zxh_adae_amd import ZXHError
    from zxh_adae_amd.kernels import DeviceState
    d = DeviceState(8, "c128")
    d.upload(rand_state(np.random.default_rng(1), 8))
    with pytest.raises(ZXHError):
        d.expect_z([(1.0, 8)])
    with pytest.raises(ZXHError):
        d.expect_z([(1.0, 0, 9)])


# ---------------- state comparison (complex64 result vs complex128) ----------------

@pytest.mark.parametrize("n", [1, 5, 12, 17])
@pytest.mark.parametrize("pa,pb", [("c64", "c128"), ("c128", "c128"), ("c64", "c64"), ("c128", "c64")])
def test_compare_matches_exact_sums(gpu_engine, n, pa, pb):
    """<a|b>, |a|^2, |b|^2 against math.fsum over the same (widened) values;
    the max amplitude error is exact (an order-free max)."""
    from zxh_adae_amd.kernels import DeviceState
    rng = np.random.default_rng(7 * n + len(pa) + 3 * len(pb))
    a = rand_state(rng, n, np.complex64 if pa == "c64" else np.complex128)
    b = (a.astype(np.complex128) + 1e-4 * rand_state(rng, n)).astype(np.complex64 if pb == "c64" else np.complex128)
    da, db = DeviceState(n, pa), DeviceState(n, pb)
    da.upload(a)
    db.upload(b)
    got = da.compare(db)
    a2, b2 = a.astype(np.complex128), b.astype(np.complex128)
    ov = np.conj(a2) * b2
    want_re, want_im = math.fsum(ov.real), math.fsum(ov.imag)
    # compensated per-thread sums, then fixed-order trees: a few ulps of the
    # unit-norm totals
    assert abs(got["overlap"].real - want_re) <= 1e-14
    assert abs(got["overlap"].imag - want_im) <= 1e-14
    assert abs(got["norm2_self"] - math.fsum(np.abs(a2) ** 2)) <= 1e-14
    assert abs(got["norm2_other"] - math.fsum(np.abs(b2) ** 2)) <= 1e-14
    assert got["max_amp_error"] == pytest.approx(float(np.max(np.abs(a2 - b2))), rel=1e-14, abs=0)
    assert got["infidelity"] == pytest.approx(1.0 - abs(complex(want_re, want_im)), abs=1e-14)


def test_compare_identical_and_phase_rotated(gpu_engine):
    """A state against itself: zero error, overlap = norm; against i*state:
    overlap i*|a|^2 (so the infidelity ignores a global phase) and
    max error sqrt(2) max|a_i|."""
    from zxh_adae_amd.kernels import DeviceState
    rng = np.random.default_rng(11)
    n = 18
    a = rand_state(rng, n)
    da, db = DeviceState(n, "c128"), DeviceState(n, "c128")
    da.upload(a)
    db.upload(a)
    same = da.compare(db)
    assert same["max_amp_error"] == 0.0
    assert abs(same["infidelity"]) < 1e-14
    db.upload(1j * a)
    rot = da.compare(db)
    assert abs(rot["overlap"] - 1j * same["norm2_self"]) < 1e-14
    assert rot["max_amp_error"] == pytest.approx(math.sqrt(2) * float(np.max(np.abs(a))), rel=1e-14)
