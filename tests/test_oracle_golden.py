"""The oracle against the reference's own outputs (tests/golden/, written by
tests/golden/make_golden.py from the compiled reference), and the numpy model
of the GPU sampler algorithm against the same index lists."""
import glob
import json
import os

import numpy as np
import pytest

import oracle_py as O
import sampler_model
from conftest import GOLDEN

MANIFEST = json.load(open(os.path.join(GOLDEN, "manifest.json")))
CIRCUITS = MANIFEST["circuits"]


def load_circ(name):
    return np.load(os.path.join(GOLDEN, f"circ_{name}.npz"))


def replay_oracle(d):
    n, dx, eager, shots, seed, M = (int(v) for v in d["meta"])
    o = O.OracleSim(n, bool(dx), bool(eager))
    for row in d["gates"]:
        o.gate(int(row[0]), int(row[1]), int(row[2]), row[3], row[4], row[5])
    o.execute()
    return o, n, shots, seed, M


@pytest.mark.parametrize("name", CIRCUITS)
def test_oracle_state_and_samples_match_reference(name):
    d = load_circ(name)
    o, n, shots, seed, M = replay_oracle(d)
    assert o.required_M() == M
    st = o.get_state()
    if "state" in d:
        # bit-exact: same operations in the same order, no FMA on either side
        assert np.array_equal(st.view(np.float64), d["state"].view(np.float64))
    else:
        idx = d["probe_idx"].astype(np.int64)
        assert np.array_equal(st[idx].view(np.float64), d["probe_val"].view(np.float64))
    o.set_seed(seed)
    assert np.array_equal(o.sampling_bits(shots), d["samples"])


SAMPLER_STATES = np.load(os.path.join(GOLDEN, "sampler_states.npz"))
SAMPLER_CASES = [(seg, key) for seg in MANIFEST["sampler_segments"]
                 for key in np.load(os.path.join(GOLDEN, f"samp_seg{seg}.npz")).files]


@pytest.mark.parametrize("seg,key", SAMPLER_CASES)
def test_oracle_sampler_matches_reference(seg, key):
    name, seed = key.split("__")
    psi = SAMPLER_STATES[name]
    want = np.load(os.path.join(GOLDEN, f"samp_seg{seg}.npz"))[key]
    got = O.oracle_sample_state(psi, seg, O.mix_seed(int(seed), 0), want.size)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("seg,key", [c for c in SAMPLER_CASES if not c[1].startswith("dense_m12")])
def test_gpu_sampler_algorithm_model_matches_reference(seg, key):
    """The tree/Fenwick formulation used by the GPU selects the same indices."""
    name, seed = key.split("__")
    psi = SAMPLER_STATES[name]
    want = np.load(os.path.join(GOLDEN, f"samp_seg{seg}.npz"))[key][:400]
    m = psi.size.bit_length() - 1
    total = O.oracle_lib().zo_prob_total(psi.astype(np.complex128).view(np.float64).ctypes.data_as(O._dp), m)
    thr = O.thresholds(O.mix_seed(int(seed), 0), MANIFEST["sampler_shots"], 0.0, total)[:400]
    got = sampler_model.sample(psi, seg, thr)
    assert np.array_equal(got, want)


def adversarial_thresholds(psi, seg):
    """Thresholds placed exactly on the reference's cdf values (and one ulp
    either side): any change of summation order moves some index."""
    cdf = sampler_model.reference_cdf(psi, seg)
    bound = min(O.prob_total(psi), float(np.max(sampler_model.segment_tails(psi, seg))))
    t = np.concatenate([cdf, np.nextafter(cdf, 0.0), np.nextafter(cdf, np.inf)])
    return np.sort(t[(t >= 0.0) & (t < bound)])


@pytest.mark.parametrize("seg", [1, 3, 4, 6])
def test_sampler_model_on_adversarial_thresholds(seg):
    rng = np.random.default_rng(seg)
    psi = (rng.normal(size=1 << 8) * np.exp(rng.normal(size=1 << 8) * 3)).astype(np.complex128)
    thr = adversarial_thresholds(psi, seg)
    want = O.oracle_sample_thresholds(psi, seg, thr)
    assert np.array_equal(sampler_model.sample(psi, seg, thr), want)


def test_planted_summation_order_change_is_caught():
    """Summing the cdf in plain sequential order (a plausible mistake) moves
    indices on adversarial thresholds, so the bit-exact tests bite."""
    rng = np.random.default_rng(0)
    psi = (rng.normal(size=1 << 8) * np.exp(rng.normal(size=1 << 8) * 3)).astype(np.complex128)
    thr = adversarial_thresholds(psi, 4)
    want = O.oracle_sample_thresholds(psi, 4, thr)
    p = psi.real * psi.real + psi.imag * psi.imag
    naive = np.minimum(np.searchsorted(np.cumsum(p), thr, side="right"), p.size - 1).astype(np.uint64)
    assert not np.array_equal(naive, want)
