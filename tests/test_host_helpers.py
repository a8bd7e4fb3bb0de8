"""Host-side helpers of the package (no device needed)."""
import numpy as np

import oracle_py as O
from zxh_adae_amd import workloads as W
from zxh_adae_amd.kernels import mix_seed


def test_mix_seed_matches_oracle():
    for b in (0, 1, 5, 123, 31415926, (1 << 63) + 7, (1 << 64) - 1):
        for salt in (0, 1, 17):
            assert mix_seed(b, salt) == O.mix_seed(b, salt)


def test_haar_unitary_is_unitary_and_seeded():
    r = W.SplitMix64(9)
    u = W.haar_unitary(r)
    assert np.max(np.abs(u.conj().T @ u - np.eye(32))) < 1e-12
    v = W.haar_unitary(W.SplitMix64(9))
    assert np.array_equal(u, v)


def test_normal_moments():
    r = W.SplitMix64(1)
    x = np.array([W.normal(r) for _ in range(20000)])
    assert abs(x.mean()) < 0.03 and abs(x.std() - 1.0) < 0.03


def test_large_circuit_layers_shape():
    layers = W.c5_layers(33, 2, 5)
    for L in layers:
        qs = [q for t, _ in L["blocks"] for q in t]
        assert len(L["blocks"]) == 6 and len(set(qs)) == 30
        assert len(L["diag"]) == 64
        for mask, t in L["mcx"]:
            k = bin(mask).count("1")
            assert 2 <= k <= 6 and not (mask >> t) & 1 and mask < (1 << 33)


def test_zz_observable_terms():
    terms = W.zz_observable(28, 200, 3)
    assert len(terms) == 200
    for t in terms:
        assert 0 <= t[1] < 28
        if len(t) > 2:
            assert t[1] < t[2] < 28
    assert terms == W.zz_observable(28, 200, 3)
