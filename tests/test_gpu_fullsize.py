"""Properties at the benchmark sizes, where the oracle is too slow to run:
closed forms, round trips, norms and sampling statistics of the 30-qubit QFT
step (the benchmark's own circuit and policy), and the 28-qubit complex64
adjoint gradient against the FP64 path.  Amplitudes are read back with the
gather kernel (a few thousand addresses, not the 16 GiB state)."""
import math

import numpy as np
import pytest

import oracle_py as O
from zxh_adae_amd import workloads as W

pytestmark = pytest.mark.gpu

NQ = 30


def qft_sim(Z, prec, gates=None):
    s = Z.ZXH(NQ, False, True, precision=prec, policy="fused")
    s.gates(W.qft(NQ) if gates is None else gates)
    s.execute()
    return s


def view(s, prec):
    from zxh_adae_amd.kernels import DeviceState
    ptr, m = s.device_state()
    assert m == NQ
    return DeviceState.wrap(ptr, m, prec)


def qft_closed_form(idx):
    """H(t) meets qubit t still in |0>, then CP(j, t, pi / 2^(t - j)) for
    j < t: amplitude(x) = prod_t H[x_t, 0] * exp(i theta(x)) with
    theta(x) = sum_{j<t} pi / 2^(t-j) x_j x_t (H = U3(pi/2, pi, 0) exactly as
    the reference builds it)."""
    h = O.u3_matrix(math.pi / 2, math.pi, 0.0)
    amp = np.ones(idx.size, dtype=np.complex128)
    theta = np.zeros(idx.size)
    for t in range(NQ):
        xt = (idx >> np.uint64(t)) & np.uint64(1)
        amp *= np.where(xt == 1, h[1, 0], h[0, 0])
        for j in range(t):
            xj = (idx >> np.uint64(j)) & np.uint64(1)
            theta += np.where((xj & xt) == 1, math.pi / 2 ** (t - j), 0.0)
    return amp * np.exp(1j * theta)


@pytest.mark.parametrize("prec", ["c128", "c64"])
def test_qft30_amplitudes_match_closed_form(gpu_engine, prec):
    s = qft_sim(gpu_engine, prec)
    rng = np.random.default_rng(30)
    idx = np.concatenate([np.array([0, (1 << NQ) - 1], dtype=np.uint64),
                          rng.integers(0, 1 << NQ, 4094, dtype=np.uint64)])
    got = view(s, prec).gather(idx).astype(np.complex128)
    want = qft_closed_form(idx)
    scale = 2.0 ** (-NQ / 2)
    tol = 1e-12 if prec == "c128" else 2e-5 * scale
    assert np.max(np.abs(got - want)) < tol


@pytest.mark.parametrize("prec", ["c128", "c64"])
def test_qft30_norm_and_inverse_round_trip(gpu_engine, prec):
    Z = gpu_engine
    from zxh_adae_amd.kernels import DeviceState  # noqa: F401
    s = qft_sim(Z, prec)
    total = view(s, prec).prob_total()
    assert abs(total - 1.0) < (1e-12 if prec == "c128" else 1e-5)
    fwd = W.qft(NQ)
    inv = [(c, cq, q, -th if c == W.CP else th, la, ph) for (c, cq, q, th, la, ph) in reversed(fwd)]
    r = qft_sim(Z, prec, fwd + [(W.BARRIER, 0, 0, 0, 0, 0)] + inv)
    rng = np.random.default_rng(31)
    idx = np.concatenate([np.array([0], dtype=np.uint64), rng.integers(1, 1 << NQ, 2047, dtype=np.uint64)])
    amp = view(r, prec).gather(idx).astype(np.complex128)
    tol = 1e-12 if prec == "c128" else 2e-5
    assert abs(amp[0] - 1.0) < tol
    assert np.max(np.abs(amp[1:])) < tol


def test_qft30_samples_are_uniform(gpu_engine):
    """|amplitude|^2 = 2^-30 everywhere, so the top four address bits of
    10^5 shots fall into 16 equal bins (chi-square, 15 degrees of freedom,
    bound at p ~ 1e-6)."""
    s = qft_sim(gpu_engine, "c128")
    s.set_seed(2718)
    real = s.sample_real(100000)
    assert np.all(np.diff(real.astype(np.int64)) >= 0)  # ascending physical addresses
    bins = np.bincount((real >> np.uint64(NQ - 4)).astype(np.int64), minlength=16)
    expect = real.size / 16
    chi2 = float(np.sum((bins - expect) ** 2 / expect))
    assert chi2 < 55.0, bins


def test_adjoint_28q_complex64_within_1e5_of_fp64(gpu_engine):
    """The benchmark's adjoint configuration: two-local, 28 qubits, 8 layers
    (448 parameters), 200 Z/ZZ terms; the complex64 gradient within 1e-5 of
    the FP64 (complex128) gradient, state infidelity below 1e-6."""
    from zxh_adae_amd.ansatz import Ansatz, two_local
    n, layers = 28, 8
    ops, npar = two_local(n, layers)
    params = np.array(W.ansatz_params(npar, 3))
    terms = W.zz_observable(n, 200, 3)
    a = Ansatz(n, ops, npar, "c64")
    a.set_observable(terms)
    g32, e32 = a.backward(params)
    a.forward(params)
    psi32 = a.state()
    a.close()
    b = Ansatz(n, ops, npar, "c128")
    b.set_observable(terms)
    g64, e64 = b.backward(params)
    b.forward(params)
    psi64 = b.state()
    b.close()
    assert np.max(np.abs(g32 - g64)) < 1e-5
    assert abs(e32 - e64) < 1e-5
    ov = abs(np.vdot(psi64, psi32.astype(np.complex128)))
    assert 1.0 - ov < 1e-6
