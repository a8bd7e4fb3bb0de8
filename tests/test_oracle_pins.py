"""Pins of the oracle that do not need the reference: known answers, closed
forms and brute-force dense products.  Each check is chosen so that a
plausible mistake (a dropped term, a transposed operand, a wrong bit) fails."""
import math

import numpy as np
import pytest

import oracle_py as O
from zxh_adae_amd import workloads as W


def test_mt19937_64_known_answer():
    # C++11 [rand.predef]: the 10000th invocation of a default-constructed
    # mt19937_64 (seed 5489) produces 9981545732273789042.
    assert O.oracle_lib().zo_mt64_nth(5489, 10000) == 9981545732273789042


def test_mix_seed_is_splitmix64_finaliser():
    # SplitMix64 reference outputs for state 0: next() = 0xe220a8397b1dcdaf
    assert O.mix_seed(0, 0) == 0xE220A8397B1DCDAF
    assert O.mix_seed(0x9E3779B97F4A7C15, 0) == 0x6E789E6AA1B965F4
    # salt is added before the finaliser
    assert O.mix_seed(5, 7) == O.mix_seed(12, 0)


def test_threshold_stream_sorted_and_inside_interval():
    t = O.thresholds(O.mix_seed(123, 0), 5000, 0.0, 2.5)
    assert np.all(np.diff(t) >= 0)
    assert t[0] >= 0.0 and t[-1] < 2.5
    # order statistics of uniforms: the mean of the k-th of n is k/(n+1)
    k = np.arange(1, 5001)
    assert abs(np.mean(t / 2.5 - k / 5001.0)) < 0.01


def test_u3_matrix_closed_form():
    th, la, ph = 0.7, 1.1, -0.4
    u = O.u3_matrix(th, la, ph)
    ref = np.array([[math.cos(th / 2), -np.exp(1j * la) * math.sin(th / 2)],
                    [np.exp(1j * ph) * math.sin(th / 2), np.exp(1j * (ph + la)) * math.cos(th / 2)]])
    assert np.max(np.abs(u - ref)) < 1e-15
    # H is U3(pi/2, pi, 0): not exactly real (imaginary residue of e^{i pi})
    h = O.u3_matrix(math.pi / 2, math.pi, 0.0)
    assert np.max(np.abs(h - np.array([[1, 1], [1, -1]]) / math.sqrt(2))) < 1e-15
    assert h[0, 1].imag != 0.0


def _dense_apply(n, gates):
    """Brute force: dense 2^n x 2^n matrices built by Kronecker products (qubit q = bit q)."""
    psi = np.zeros(1 << n, dtype=np.complex128)
    psi[0] = 1
    idx = np.arange(1 << n)
    gphase = 1.0 + 0j
    for (code, cq, q, th, la, ph) in gates:
        if code in (W.H, W.U3):
            u = O.u3_matrix(*(math.pi / 2, math.pi, 0.0) if code == W.H else (th, la, ph))
            mats = [np.eye(2)] * n
            mats[q] = u
            full = np.array([[1.0]])
            for k in range(n - 1, -1, -1):
                full = np.kron(full, mats[k])
            psi = full @ psi
        elif code == W.P:
            psi = psi * np.where((idx >> q) & 1, np.exp(1j * th), 1.0)
        elif code == W.CP:
            psi = psi * np.where(((idx >> q) & 1) & ((idx >> cq) & 1), np.exp(1j * th), 1.0)
        elif code == W.RZ:
            psi = psi * np.where((idx >> q) & 1, np.exp(1j * th), 1.0)
            gphase *= np.exp(-0.5j * th)
        elif code == W.X:
            psi = psi[idx ^ (1 << q)]
        elif code == W.CX:
            psi = psi[np.where((idx >> cq) & 1, idx ^ (1 << q), idx)]
    return psi * gphase


def _bitrev(n):
    idx = np.arange(1 << n)
    rev = np.zeros_like(idx)
    for b in range(n):
        rev |= ((idx >> b) & 1) << (n - 1 - b)
    return rev


@pytest.mark.parametrize("eager,disable_x", [(True, False), (False, False), (False, True)])
def test_oracle_matches_dense_brute_force(eager, disable_x):
    n = 6
    gs = W.random_mixed(n, 70, 5, kinds=(W.H, W.U3, W.P, W.CP, W.RZ, W.X, W.CX))
    o = O.OracleSim(n, disable_x, eager)
    o.gates(gs)
    o.execute()
    got = o.get_state()
    ref = _dense_apply(n, gs)
    # get_state is in bit-reversed ("cudaq") order
    assert np.max(np.abs(got[_bitrev(n)] - ref)) < 1e-12


def test_diag_on_uniform_superposition_closed_form():
    n = 9
    gs = W.hadamard_all(n) + W.diag_batch(n, 40, 3)
    o = O.OracleSim(n, False, True)
    o.gates(gs)
    o.execute()
    got = o.phys_state()
    idx = np.arange(1 << n)
    theta = np.zeros(1 << n)
    for (code, cq, q, th, _, _) in gs[n:]:
        if code == W.P:
            theta += np.where((idx >> q) & 1, th, 0.0)
        else:
            theta += np.where(((idx >> q) & 1) & ((idx >> cq) & 1), th, 0.0)
    ref = np.exp(1j * theta) / math.sqrt(1 << n)
    assert np.max(np.abs(got - ref)) < 1e-13


def test_z_is_rz_pi_with_global_phase():
    o = O.OracleSim(1, False, False)
    o.gates([(W.H, 0, 0, 0, 0, 0), (W.Z, 0, 0, 0, 0, 0)])
    o.execute()
    assert abs(o.global_phase() - np.exp(-0.5j * math.pi)) < 1e-15


def test_u3_identity_products_are_dropped():
    # a product within 1e-5 of I is dropped, a larger deviation is applied
    o = O.OracleSim(2, False, True)
    o.gates([(W.U3, 0, 0, 0.4, 0.3, 0.2), (W.U3, 0, 0, -0.4 + 1e-7, -0.2, -0.3),
             (W.U3, 0, 1, 0.4, 0.3, 0.2), (W.U3, 0, 1, -0.4 + 1e-3, -0.2, -0.3)])
    o.execute()
    st = o.phys_state()
    assert st[0] != 1.0  # qubit 1 deviates
    # qubit 0 exactly untouched: amplitudes with bit0 = 1 are exactly zero
    assert st[1] == 0 and st[3] == 0


def test_sampler_single_amplitude_and_delta():
    psi = np.zeros(16, dtype=np.complex128)
    psi[11] = 1.0
    for seg in (1, 2, 4, 18):
        idx = O.oracle_sample_state(psi, seg, O.mix_seed(9, 0), 100)
        assert np.all(idx == 11)


def test_sampler_frequencies():
    rng = np.random.default_rng(3)
    psi = rng.normal(size=64) + 1j * rng.normal(size=64)
    psi /= np.linalg.norm(psi)
    idx = O.oracle_sample_state(psi, 3, O.mix_seed(11, 0), 200000)
    freq = np.bincount(idx.astype(np.int64), minlength=64) / 200000
    assert np.max(np.abs(freq - np.abs(psi) ** 2)) < 0.006
    assert np.all(np.diff(idx.astype(np.int64)) >= 0)  # ascending real addresses
