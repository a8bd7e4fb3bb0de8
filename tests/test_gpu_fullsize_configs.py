"""Correctness at the benchmark configurations' full sizes, where the oracle
cannot run the whole circuit: properties the mathematics fixes (unitarity,
exact permutations, block-then-inverse round trips), the oracle on the GPU's
own state where it finishes in seconds (sampling at 2^26..2^28), and the
single-circuit path as the check of the batched one.

  * the large random circuit (5-qubit Haar blocks on the matrix cores, 64-gate
    diagonal batches, multi-controlled X cascades), 40 layers at 33 qubits
    complex64 and 32 qubits complex128 (2^33 / 2^32 amplitudes, 64 GiB):
    the total probability, a block followed by its inverse on gathered
    addresses below and above 2^32, and the MCX cascade alone on a basis
    state against the classical permutation;
  * sampling bit-exact with the oracle at m = 26, 27 (complex64) and 28,
    segment sizes 2^18 and 2^9 (the probability tree's third pass and wide
    segment tables);
  * the batched uneven ansatz at its full configuration (512 sets, 16..24
    qubits, 2..10 layers, 50 terms) against the single-circuit path on every
    8th set, over two replays, complex64 and complex128;
  * the 28-qubit adjoint gradient against exact parameter shifts on 16 of
    its 448 parameters (complex128).
"""
import math

import numpy as np
import pytest

import oracle_py as O
from zxh_adae_amd import workloads as W

pytestmark = pytest.mark.gpu


def c5_probes(rng, n, count=4096):
    """Gather addresses: 0, the last, both sides of 2^32 when the state is
    that wide, and random ones (half of them at or above 2^32)."""
    fixed = [0, (1 << n) - 1]
    if n > 32:
        fixed += [(1 << 32) - 1, 1 << 32, (1 << 32) + 1]
    lo = rng.integers(0, 1 << min(n, 32), (count - len(fixed)) // 2, dtype=np.uint64)
    hi = rng.integers(1 << (n - 1), 1 << n, count - len(fixed) - lo.size, dtype=np.uint64)
    return np.unique(np.concatenate([np.array(fixed, dtype=np.uint64), lo, hi]))


@pytest.mark.parametrize("n,prec", [(33, "c64"), (32, "c128")])
def test_c5_full_size_norm_and_block_round_trip(gpu_engine, n, prec):
    from zxh_adae_amd.kernels import DeviceState, apply_layer, set_basis_zero
    spec = W.c5_layers(n, 40, 5)
    st = DeviceState(n, prec)
    try:
        set_basis_zero(st)
        for L in spec:
            apply_layer(st, L)
        total = st.prob_total()
        assert abs(total - 1.0) < (1e-5 if prec == "c64" else 1e-12), total
        rng = np.random.default_rng(n)
        idx = c5_probes(rng, n)
        before = st.gather(idx).astype(np.complex128)
        targets = [int(t) for t in rng.choice(n, 5, replace=False)]
        u = W.haar_unitary(W.SplitMix64(77 + n))
        st.block5(targets, u)
        moved = st.gather(idx).astype(np.complex128)
        st.block5(targets, np.conj(u.T))
        after = st.gather(idx).astype(np.complex128)
        scale = float(np.max(np.abs(before)))
        tol = 1e-12 if prec == "c128" else 2e-5 * scale
        assert np.max(np.abs(after - before)) < tol
        # the block did act on these addresses
        assert np.max(np.abs(moved - before)) > 100 * tol
    finally:
        st.free()


@pytest.mark.parametrize("n,prec", [(33, "c64"), (32, "c128")])
def test_c5_mcx_cascade_is_the_classical_permutation(gpu_engine, n, prec):
    from zxh_adae_amd.kernels import DeviceState, set_basis_zero
    spec = W.c5_layers(n, 40, 5)
    rng = np.random.default_rng(100 + n)
    x = int(rng.integers(0, 1 << n, dtype=np.uint64)) | (1 << (n - 1))  # above 2^32
    st = DeviceState(n, prec)
    try:
        set_basis_zero(st)
        for q in range(n):  # |0> -> |x> with X gates (exact permutations)
            if (x >> q) & 1:
                st.x(1 << q, 1 << q)
        want = x
        for L in spec:
            for mask, t in L["mcx"]:
                st.mcx(mask, t)
                if want & mask == mask:
                    want ^= 1 << t
        probe = np.array([want, x, want ^ 1, 0], dtype=np.uint64)
        amp = st.gather(probe)
        assert amp[0] == 1.0
        if want != x:
            assert amp[1] == 0.0
        assert amp[2] == 0.0
        assert st.prob_total() == 1.0
    finally:
        st.free()


@pytest.mark.parametrize("m,prec,seg", [(26, "c128", 18), (26, "c128", 9), (27, "c64", 18), (28, "c128", 18)])
def test_sampling_bit_exact_at_large_widths(gpu_engine, m, prec, seg):
    """A random-layer state made on the GPU, read back, and sampled by the
    oracle on the CPU: the same indices for 10^5 shots."""
    s = gpu_engine.ZXH(m, False, True, precision=prec)
    s.gates(W.random_layers(m, 2, 40 + m))
    s.execute()
    s.set_seg_bits(seg)
    s.set_seed(9000 + m)
    got = s.sample_real(100000)
    psi = s.phys_state()
    del s
    want = O.oracle_sample_state(psi.astype(np.complex128) if prec == "c64" else psi, seg,
                                 O.mix_seed(9000 + m, 0), 100000)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("prec,tol", [("c64", 2e-5), ("c128", 1e-10)])
def test_c4_full_batch_matches_single_circuits(gpu_engine, prec, tol):
    from zxh_adae_amd.ansatz import Ansatz, AnsatzBatch, two_local
    sets = W.c4_sets()
    b = AnsatzBatch(prec)
    nps = []
    for n, layers, terms in sets:
        ops, npar = two_local(n, layers)
        b.add(n, ops, npar, terms)
        nps.append(npar)
    b.finalize()
    offs = np.concatenate([[0], np.cumsum(nps)])
    try:
        for replay in range(2):
            params = np.array(W.ansatz_params(int(offs[-1]), 4 + replay))
            e, nrm, g = b.run(params, gradients=True)
            for i in range(0, len(sets), 8):
                n, layers, terms = sets[i]
                ops, npar = two_local(n, layers)
                a = Ansatz(n, ops, npar, prec)
                a.set_observable(terms)
                gi, ei = a.backward(params[offs[i]:offs[i + 1]])
                a.close()
                scale = max(1.0, sum(abs(t[0]) for t in terms))
                assert abs(e[i] - ei) < tol * scale, (replay, i)
                assert np.max(np.abs(g[i] - gi)) < tol * scale, (replay, i)
                assert abs(nrm[i] - 1.0) < (1e-5 if prec == "c64" else 1e-12)
    finally:
        b.close()


def test_c3_adjoint_28q_against_parameter_shift(gpu_engine):
    """RY and RZ are exp(-i theta P / 2) up to a global phase, so
    dE/dtheta = (E(theta + pi/2) - E(theta - pi/2)) / 2 exactly; 16 of the 448
    parameters, complex128."""
    from zxh_adae_amd.ansatz import Ansatz, two_local
    n, layers = 28, 8
    ops, npar = two_local(n, layers)
    params = np.array(W.ansatz_params(npar, 3))
    terms = W.zz_observable(n, 200, 3)
    a = Ansatz(n, ops, npar, "c128")
    try:
        a.set_observable(terms)
        g, _ = a.backward(params)
        rng = np.random.default_rng(28)
        for k in rng.choice(npar, 16, replace=False):
            p = params.copy()
            p[k] += math.pi / 2
            ep, _ = a.forward(p)
            p[k] -= math.pi
            em, _ = a.forward(p)
            assert abs((ep - em) / 2 - g[k]) < 1e-10, k
    finally:
        a.close()
