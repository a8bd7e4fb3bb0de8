"""Batched parameterized circuits (uneven widths and depths, one persistent
launch per pass index, replayed from a hipGraph): every set's energy and
gradient against the single-circuit path.  No reference counterpart."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


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


@pytest.mark.parametrize("prec,widths", [("c64", [15, 16, 18, 17, 15, 19]), ("c128", [14, 16, 15, 17])])
def test_batch_matches_single_circuits_over_replays(gpu_engine, prec, widths):
    from zxh_adae_amd.ansatz import Ansatz, AnsatzBatch, two_local
    rng = np.random.default_rng(len(widths))
    sets = []
    b = AnsatzBatch(prec)
    for k, n in enumerate(widths):
        layers = 1 + (k % 3)
        ops, npar = two_local(n, layers)
        terms = rand_terms(rng, n, 12)
        b.add(n, ops, npar, terms)
        sets.append((n, ops, npar, terms))
    b.finalize()
    tol = 1e-11 if prec == "c128" else 2e-5
    for replay in range(3):
        params = [rng.uniform(-math.pi, math.pi, s[2]) for s in sets]
        e, nrm, g = b.run(params)
        e_fwd, _, none = b.run(params, gradients=False)
        assert none is None
        for k, (n, ops, npar, terms) in enumerate(sets):
            a = Ansatz(n, ops, npar, prec)
            a.set_observable(terms)
            assert b.set_info(k) == (n, a.tile_passes())
            gk, ek = a.backward(params[k])
            scale = max(1.0, sum(abs(t[0]) for t in terms))
            assert abs(e[k] - ek) < tol * scale, (replay, k)
            assert abs(e_fwd[k] - ek) < tol * scale, (replay, k)
            assert abs(nrm[k] - 1.0) < tol
            assert np.max(np.abs(g[k] - gk)) < tol * scale, (replay, k)
            a.close()


def test_batch_rejects_narrow_and_unfinalized(gpu_engine):
    from zxh_adae_amd import ZXHError
    from zxh_adae_amd.ansatz import AnsatzBatch, two_local
    b = AnsatzBatch("c64")
    ops, npar = two_local(15, 1)
    b.add(15, ops, npar, [(1.0, 0)])
    with pytest.raises(ZXHError):
        b.run(np.zeros(npar))
    with pytest.raises(ZXHError):
        b.add(8, *two_local(8, 1), [(1.0, 0)])
    b.finalize()
    with pytest.raises(ZXHError):
        b.set_info(1)


@pytest.mark.parametrize("seg", [3, 4, 6])
def test_batch_with_forced_geometry_matches_single_circuits(gpu_engine, seg):
    """The batch autotuner's per-width geometry (zxh_batch_set_geometry): a
    width class forced to geometry seg, another left to the rule, both equal
    to the single-circuit path (complex128)."""
    from zxh_adae_amd.ansatz import Ansatz, AnsatzBatch, two_local
    rng = np.random.default_rng(seg)
    widths = [17, 15, 17, 16]
    b = AnsatzBatch("c128")
    b.set_geometry(17, seg)
    sets = []
    for k, n in enumerate(widths):
        ops, npar = two_local(n, 2 + k % 2)
        terms = rand_terms(rng, n, 10)
        b.add(n, ops, npar, terms)
        sets.append((n, ops, npar, terms))
    b.finalize()
    params = [rng.uniform(-math.pi, math.pi, s[2]) for s in sets]
    e, nrm, g = b.run(params)
    for k, (n, ops, npar, terms) in enumerate(sets):
        a = Ansatz(n, ops, npar, "c128")
        a.set_observable(terms)
        gk, ek = a.backward(params[k])
        a.close()
        scale = max(1.0, sum(abs(t[0]) for t in terms))
        assert abs(e[k] - ek) < 1e-11 * scale
        assert np.max(np.abs(g[k] - gk)) < 1e-11 * scale
    from zxh_adae_amd import ZXHError
    with pytest.raises(ZXHError):
        b.set_geometry(16, 4)  # after finalize
    b2 = AnsatzBatch("c64")
    with pytest.raises(ZXHError):
        b2.set_geometry(16, 7)
