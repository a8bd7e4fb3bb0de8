"""Per-kernel parity of the HIP passes against the oracle (same inputs)."""
import json
import math
import os

import numpy as np
import pytest

import oracle_py as O
import sampler_model
from conftest import GOLDEN

pytestmark = pytest.mark.gpu

MANIFEST = json.load(open(os.path.join(GOLDEN, "manifest.json")))
SEL_NONE = 0xFFFFFFFFFFFFFFFF
TOL128 = 1e-12  # north-star bound for f64 amplitudes
TOL64 = 2e-5    # complex64 state with f32 arithmetic


def rand_state(rng, n, dtype=np.complex128):
    psi = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    return (psi / np.linalg.norm(psi)).astype(dtype)


def dev(gpu_engine, psi, prec="c128"):
    from zxh_adae_amd.kernels import DeviceState
    n = psi.size.bit_length() - 1
    d = DeviceState(n, prec)
    d.upload(psi)
    return d


def pack(alpha, beta):
    return (alpha & ((1 << 63) - 1)) | ((1 << 63) if beta else 0)


# ---------------- sampler ----------------

SAMPLER_STATES = np.load(os.path.join(GOLDEN, "sampler_states.npz"))


@pytest.mark.parametrize("seg", MANIFEST["sampler_segments"])
def test_sampler_matches_reference_golden(gpu_engine, seg):
    gold = np.load(os.path.join(GOLDEN, f"samp_seg{seg}.npz"))
    for key in gold.files:
        name, seed = key.split("__")
        psi = SAMPLER_STATES[name]
        d = dev(gpu_engine, psi)
        got = d.sample(seg, O.mix_seed(int(seed), 0), gold[key].size)
        assert np.array_equal(got, gold[key]), (name, seed, seg)


@pytest.mark.parametrize("seg", [1, 2, 4, 5, 7, 18])
def test_sampler_adversarial_thresholds(gpu_engine, seg):
    rng = np.random.default_rng(100 + seg)
    psi = (rng.normal(size=1 << 9) * np.exp(rng.normal(size=1 << 9) * 3)).astype(np.complex128)
    cdf = sampler_model.reference_cdf(psi, seg)
    bound = min(O.prob_total(psi), float(np.max(sampler_model.segment_tails(psi, seg))))
    thr = np.concatenate([cdf, np.nextafter(cdf, 0.0), np.nextafter(cdf, np.inf)])
    thr = np.sort(thr[(thr >= 0) & (thr < bound)])
    want = O.oracle_sample_thresholds(psi, seg, thr)
    d = dev(gpu_engine, psi)
    assert np.array_equal(d.sample_thresholds(seg, thr), want)


@pytest.mark.parametrize("m,seg", [(14, 18), (16, 5), (20, 18), (21, 9)])
def test_sampler_random_states(gpu_engine, m, seg):
    rng = np.random.default_rng(m * 31 + seg)
    psi = rand_state(rng, m)
    psi[rng.random(psi.size) < 0.3] = 0
    d = dev(gpu_engine, psi)
    for seed in (3, 99):
        want = O.oracle_sample_state(psi, seg, O.mix_seed(seed, 0), 20000)
        assert np.array_equal(d.sample(seg, O.mix_seed(seed, 0), 20000), want)
    assert d.prob_total() == O.prob_total(psi)


def test_sampler_c64_state_bit_exact_on_widened_probabilities(gpu_engine):
    rng = np.random.default_rng(5)
    psi = rand_state(rng, 13, np.complex64)
    d = dev(gpu_engine, psi, "c64")
    want = O.oracle_sample_state(psi.astype(np.complex128), 6, O.mix_seed(77, 0), 5000)
    assert np.array_equal(d.sample(6, O.mix_seed(77, 0), 5000), want)


def test_sampler_zero_state_raises(gpu_engine):
    from zxh_adae_amd import ZXHError
    d = dev(gpu_engine, np.zeros(1 << 6, dtype=np.complex128))
    with pytest.raises(ZXHError):
        d.sample(3, 1, 10)


# ---------------- diagonal window ----------------

def diag_window(rng, n, count, multi_bit=False):
    s0, s1, th = [], [], []
    for _ in range(count):
        def sel():
            if multi_bit and rng.random() < 0.5:
                alpha = int(rng.integers(1, 1 << n))
            else:
                alpha = 1 << int(rng.integers(n)) if rng.random() > 0.05 else 0
            return pack(alpha, bool(rng.random() < 0.3))
        a = sel()
        if rng.random() < 0.5:
            s0.append(a)
            s1.append(SEL_NONE)
        else:
            b = sel()
            s0.append(a)
            s1.append(b)
        th.append(float(rng.uniform(0, 2 * math.pi)))
    return np.array(s0, dtype=np.uint64), np.array(s1, dtype=np.uint64), np.array(th)


def oracle_diag(psi, n, s0, s1, th):
    out = psi.astype(np.complex128).copy()
    lib = O.oracle_lib()
    # apply in bucketable groups (the oracle rejects cross CP like the reference)
    rc = lib.zo_k_diag(out.view(np.float64).ctypes.data_as(O._dp), n, s0.ctypes.data_as(O._u64p),
                       s1.ctypes.data_as(O._u64p), th.ctypes.data_as(O._dp), th.size)
    assert rc == 0
    return out


def bucketable(s0, s1):
    def cls(p):
        a = int(p) & ((1 << 63) - 1)
        return 0 if a >> 8 == 0 else (1 if a & 0xFF == 0 else 2)
    keep = [i for i in range(s0.size) if s1[i] == SEL_NONE or (cls(s0[i]) != 2 and cls(s1[i]) != 2)]
    return np.array(keep, dtype=np.int64)


@pytest.mark.parametrize("n,count", [(5, 20), (11, 40), (12, 1), (12, 64), (13, 256), (16, 100), (18, 16)])
@pytest.mark.parametrize("prec", ["c128", "c64"])
def test_diag_single_bit_selectors(gpu_engine, n, count, prec):
    rng = np.random.default_rng(n * 1000 + count)
    psi = rand_state(rng, n)
    s0, s1, th = diag_window(rng, n, count)
    want = oracle_diag(psi, n, s0, s1, th)
    d = dev(gpu_engine, psi.astype(np.complex64) if prec == "c64" else psi, prec)
    d.diag(s0, s1, th)
    got = d.download().astype(np.complex128)
    assert np.max(np.abs(got - want)) < (TOL128 if prec == "c128" else TOL64)


@pytest.mark.parametrize("n,count", [(12, 40), (14, 300), (17, 120), (20, 64)])
@pytest.mark.parametrize("prec", ["c128", "c64"])
def test_diag_two_bit_single_selectors_polynomial(gpu_engine, n, count, prec):
    """P / Rz-type gates on two-bit selectors (an Rzz in lazy mode) mixed with
    single-bit CP: still a degree-2 phase polynomial, swept by the fast path."""
    rng = np.random.default_rng(n * 31 + count)
    psi = rand_state(rng, n)
    s0, s1, th = diag_window(rng, n, count)
    for i in range(count):
        if s1[i] == SEL_NONE and rng.random() < 0.7:
            u, v = (int(b) for b in rng.choice(n, 2, replace=False))
            s0[i] = pack((1 << u) | (1 << v), bool(rng.random() < 0.4))
    want = oracle_diag(psi, n, s0, s1, th)
    d = dev(gpu_engine, psi.astype(np.complex64) if prec == "c64" else psi, prec)
    d.diag(s0, s1, th)
    got = d.download().astype(np.complex128)
    assert np.max(np.abs(got - want)) < (TOL128 if prec == "c128" else TOL64)


@pytest.mark.parametrize("n", [6, 12, 14])
def test_diag_multi_bit_selectors(gpu_engine, n):
    rng = np.random.default_rng(7 + n)
    psi = rand_state(rng, n)
    s0, s1, th = diag_window(rng, n, 50, multi_bit=True)
    k = bucketable(s0, s1)
    s0, s1, th = s0[k], s1[k], th[k]
    want = oracle_diag(psi, n, s0, s1, th)
    d = dev(gpu_engine, psi)
    d.diag(s0, s1, th)
    assert np.max(np.abs(d.download() - want)) < TOL128


# ---------------- U3 batches ----------------

def rand_u(rng):
    return O.u3_matrix(*rng.uniform(0, 2 * math.pi, 3))


def oracle_u3(psi, n, descs):
    out = psi.astype(np.complex128).copy()
    arr = (O.U3Desc * len(descs))()
    for i, (sel, delta, u) in enumerate(descs):
        arr[i].packed_sel = sel
        arr[i].delta = delta
        f = np.asarray(u).reshape(4)
        for k in range(4):
            arr[i].u[2 * k] = f[k].real
            arr[i].u[2 * k + 1] = f[k].imag
    O.oracle_lib().zo_k_u3_batch(out.view(np.float64).ctypes.data_as(O._dp), n, arr, len(descs))
    return out


@pytest.mark.parametrize("n,targets", [
    (4, [0, 1, 2, 3]), (12, list(range(12))), (13, [0, 5, 12]), (16, list(range(16))),
    (16, [12, 13, 14, 15]), (17, [2, 9, 16]), (20, list(range(0, 20, 2))), (22, [21]), (12, [3]), (14, [4, 8, 13])])
@pytest.mark.parametrize("prec", ["c128", "c64"])
def test_u3_batch_plain_selectors(gpu_engine, n, targets, prec):
    rng = np.random.default_rng(n * 7 + len(targets))
    psi = rand_state(rng, n)
    descs = [(pack(1 << q, False), 1 << q, rand_u(rng)) for q in targets]
    want = oracle_u3(psi, n, descs)
    d = dev(gpu_engine, psi.astype(np.complex64) if prec == "c64" else psi, prec)
    d.u3_batch(descs)
    got = d.download().astype(np.complex128)
    assert np.max(np.abs(got - want)) < (TOL128 if prec == "c128" else TOL64)


def random_affine(rng, n, cx_count):
    """Rows alpha_q of A and pair offsets delta_q = A^-1 e_q of an affine map
    built by CX row transport from the identity (alpha_i . delta_j = [i == j])."""
    rows = [1 << q for q in range(n)]
    for _ in range(cx_count):
        c, t = rng.choice(n, 2, replace=False)
        rows[int(t)] ^= rows[int(c)]
    # invert over GF(2): augment and eliminate
    aug = [(rows[q], 1 << q) for q in range(n)]
    for col in range(n):
        piv = next(r for r in range(col, n) if (aug[r][0] >> col) & 1)
        aug[col], aug[piv] = aug[piv], aug[col]
        for r in range(n):
            if r != col and (aug[r][0] >> col) & 1:
                aug[r] = (aug[r][0] ^ aug[col][0], aug[r][1] ^ aug[col][1])
    inv_rows = [aug[q][1] for q in range(n)]  # rows of A^-1
    deltas = [sum(((inv_rows[r] >> q) & 1) << r for r in range(n)) for q in range(n)]
    return rows, deltas


@pytest.mark.parametrize("n,cx,prec", [(8, 6, "c128"), (13, 3, "c128"), (15, 10, "c128"), (16, 40, "c128"),
                                       (20, 120, "c128"), (22, 200, "c128"), (14, 30, "c64"), (19, 90, "c64")])
def test_u3_batch_affine_selectors(gpu_engine, n, cx, prec):
    """Selectors and pair offsets of a CX-transported map: extra alpha bits
    (per-pair flip), beta = 1 (swapped roles), multi-bit deltas.  From 12
    qubits (13 for complex64) every gate runs in tile passes whose slots
    carry the whole offsets, except complex64 offsets that use address bit 0
    with others (one pair sweep each)."""
    from zxh_adae_amd import kernels as K
    rng = np.random.default_rng(n)
    rows, deltas = random_affine(rng, n, cx)
    for q in range(n):
        assert bin(rows[q] & deltas[q]).count("1") % 2 == 1
    psi = rand_state(rng, n)
    descs = [(pack(rows[q], bool(rng.random() < 0.5)), deltas[q], rand_u(rng)) for q in range(n) if rng.random() < 0.7]
    assert sum(bin(dl).count("1") > 1 for _, dl, _ in descs) >= 2
    want = oracle_u3(psi, n, descs)
    d = dev(gpu_engine, psi.astype(np.complex64) if prec == "c64" else psi, prec)
    K.set_timing(True)
    K.reset_timing()
    try:
        d.u3_batch(descs)
        pair = K.timing("u3_pair")["launches"]
        tile = K.timing("u3_tile")["launches"]
    finally:
        K.set_timing(False)
    got = d.download().astype(np.complex128)
    assert np.max(np.abs(got - want)) < (TOL128 if prec == "c128" else TOL64)
    if n >= (13 if prec == "c64" else 12):
        split = [dl for _, dl, _ in descs if prec == "c64" and dl & 1 and dl != 1]
        assert pair == len(split)
        assert 1 <= tile <= len(descs)


def test_u3_batch_of_commuting_wide_gates_in_few_passes(gpu_engine):
    """A full layer of 20 gates after a CX chain (every offset but one has
    several bits): 16 high pivots over 8 upper slots, so three tile passes
    instead of 20 pair sweeps."""
    from zxh_adae_amd import kernels as K
    n = 20
    rows = [1 << q for q in range(n)]
    for q in range(n - 1):
        rows[q + 1] ^= rows[q]  # CX(q, q+1) transports row q into row q+1
    # rows of the map: alpha_q = bits 0..q; offsets: delta_q = e_q + e_(q+1)
    deltas = [(1 << q) | ((1 << (q + 1)) if q + 1 < n else 0) for q in range(n)]
    for i in range(n):
        for j in range(n):
            assert bin(rows[i] & deltas[j]).count("1") % 2 == int(i == j)
    rng = np.random.default_rng(11)
    psi = rand_state(rng, n)
    descs = [(pack(rows[q], False), deltas[q], rand_u(rng)) for q in range(n)]
    want = oracle_u3(psi, n, descs)
    d = dev(gpu_engine, psi)
    K.set_timing(True)
    K.reset_timing()
    try:
        d.u3_batch(descs)
        pair = K.timing("u3_pair")["launches"]
        tile = K.timing("u3_tile")["launches"]
    finally:
        K.set_timing(False)
    assert np.max(np.abs(d.download() - want)) < TOL128
    assert pair == 0
    assert tile <= 3


def test_u3_batch_that_does_not_commute_runs_in_order(gpu_engine):
    """Two gates on one pair offset (not one batch of a map): applied one
    after the other, as the oracle's batch loop does."""
    rng = np.random.default_rng(12)
    n = 14
    psi = rand_state(rng, n)
    descs = [(pack(1 << 9, False), 1 << 9, rand_u(rng)), (pack((1 << 9) | 2, True), 1 << 9, rand_u(rng)),
             (pack(1 << 3, False), (1 << 3) | (1 << 12), rand_u(rng))]
    want = oracle_u3(psi, n, descs)
    d = dev(gpu_engine, psi)
    d.u3_batch(descs)
    assert np.max(np.abs(d.download() - want)) < TOL128


# ---------------- permutations and CP ----------------

@pytest.mark.parametrize("n", [5, 13])
def test_x_cx_cp_against_oracle(gpu_engine, n):
    rng = np.random.default_rng(n + 40)
    psi = rand_state(rng, n)
    lib = O.oracle_lib()
    want = psi.copy()
    d = dev(gpu_engine, psi)
    for step in range(6):
        q, c = rng.choice(n, 2, replace=False)
        sel_q = pack(1 << int(q) | (int(rng.integers(0, 1 << n)) & ~(1 << int(q)) if step % 2 else 0), step % 3 == 0)
        sel_c = pack(1 << int(c), step % 2 == 1)
        p = want.view(np.float64).ctypes.data_as(O._dp)
        if step % 3 == 0:
            lib.zo_k_x(p, n, 1 << int(q), sel_q)
            d.x(1 << int(q), sel_q)
        elif step % 3 == 1:
            lib.zo_k_cx(p, n, 1 << int(q), sel_c, sel_q)
            d.cx(1 << int(q), sel_c, sel_q)
        else:
            lib.zo_k_cp(p, n, sel_c, sel_q, 0.9 + step)
            d.cp(sel_c, sel_q, 0.9 + step)
    assert np.max(np.abs(d.download() - want)) < TOL128


@pytest.mark.parametrize("n,ctrls,target", [(6, [0, 2], 5), (14, [1, 3, 7, 9, 11, 13], 0), (16, [15, 4], 8)])
def test_mcx_is_exact_permutation(gpu_engine, n, ctrls, target):
    rng = np.random.default_rng(n)
    psi = rand_state(rng, n)
    idx = np.arange(1 << n)
    mask = sum(1 << c for c in ctrls)
    perm = np.where((idx & mask) == mask, idx ^ (1 << target), idx)
    d = dev(gpu_engine, psi)
    d.mcx(mask, target)
    assert np.array_equal(d.download(), psi[perm])


# ---------------- fused tile pass (U3 batch then diagonal window) ----------------

@pytest.mark.parametrize("n,low,high,ndiag", [(12, [0, 2], [5, 11], 30), (13, [0, 1, 2, 3], list(range(4, 12)), 100),
                                             (13, [0, 2, 4], [5, 12], 30), (14, [], list(range(6, 14)), 50),
                                             (16, [], [15, 14, 4, 9], 64), (20, [1], [19, 12, 7], 256),
                                             (14, [3], [], 10), (15, [], [], 40), (18, [4], [17], 0)])
@pytest.mark.parametrize("prec", ["c128", "c64"])
def test_fused_pass_matches_u3_then_diag(gpu_engine, tile_seg, n, low, high, ndiag, prec):
    """Targets below the tile's low bits (seg for complex128, seg + 1 for
    complex64) ride free; at most 12 - seg others; the tile needs 12 / 13
    address bits."""
    rng = np.random.default_rng(n * 13 + ndiag)
    psi = rand_state(rng, n)
    descs = [(pack(1 << q, False), 1 << q, rand_u(rng)) for q in low + high]
    s0, s1, th = diag_window(rng, n, ndiag) if ndiag else ([], [], [])
    want = oracle_diag(oracle_u3(psi, n, descs) if descs else psi, n, s0, s1, th) if ndiag else oracle_u3(psi, n, descs)
    d = dev(gpu_engine, psi.astype(np.complex64) if prec == "c64" else psi, prec)
    lowbits = tile_seg + (1 if prec == "c64" else 0)
    shift = 1 if prec == "c64" else 0
    from zxh_adae_amd import autotune
    # the wide (8192-unit) tile takes up to nine targets above a 16-unit run
    wide = (autotune.tune_get(f"tile.wide.{prec}") == 1 and n >= 13 + shift and len(low + high) <= 9
            and all(q >= 4 + shift for q in low + high))
    if (n < lowbits + 12 - tile_seg or sum(q >= lowbits for q in low + high) > 12 - tile_seg) and not wide:
        from zxh_adae_amd import ZXHError
        with pytest.raises(ZXHError):
            d.fused_pass(descs, s0, s1, th)
        return
    d.fused_pass(descs, s0, s1, th)
    got = d.download().astype(np.complex128)
    assert np.max(np.abs(got - want)) < (TOL128 if prec == "c128" else TOL64)


@pytest.mark.parametrize("n,targets", [(13, list(range(13))), (20, list(range(20))), (17, [0, 5, 6, 9, 16])])
@pytest.mark.parametrize("prec", ["c128", "c64"])
def test_u3_batch_under_each_tile_geometry(gpu_engine, tile_seg, n, targets, prec):
    rng = np.random.default_rng(n * 5 + tile_seg)
    psi = rand_state(rng, n)
    descs = [(pack(1 << q, False), 1 << q, rand_u(rng)) for q in targets]
    want = oracle_u3(psi, n, descs)
    d = dev(gpu_engine, psi.astype(np.complex64) if prec == "c64" else psi, prec)
    d.u3_batch(descs)
    got = d.download().astype(np.complex128)
    assert np.max(np.abs(got - want)) < (TOL128 if prec == "c128" else TOL64)


@pytest.fixture
def upper_knobs(gpu_engine):
    """Restores the tile geometry and upper-kernel knobs after a test."""
    from zxh_adae_amd import autotune
    knobs = ("tile.upper.c128", "tile.upper.c64", "tile.seg.c128", "tile.seg.c64")
    old = {k: autotune.tune_get(k) for k in knobs}
    yield autotune
    for k, v in old.items():
        autotune.tune_set(k, v)


@pytest.mark.parametrize("prec,seg,n,targets,ndiag,upper", [
    ("c128", 6, 18, [12, 13, 14, 15, 16, 17], 0, 1), ("c128", 6, 20, [19, 18, 14, 15, 16, 17], 0, 1),
    ("c128", 6, 20, [8, 9, 10, 11, 18, 19], 40, 1), ("c128", 6, 22, [16, 21], 0, 1),
    ("c128", 6, 20, [3, 15, 19], 0, 0), ("c128", 4, 20, [10, 12, 14, 16, 18, 19], 64, 1),
    ("c64", 6, 20, [0, 14, 15, 16, 17, 18, 19], 0, 1), ("c64", 4, 19, [0, 9, 11, 13, 18], 30, 1),
    ("c64", 3, 19, [1, 5, 9, 13, 18], 0, 0)])
def test_upper_kernel_matches_the_tile_kernel(upper_knobs, prec, seg, n, targets, ndiag, upper):
    """Passes whose targets all sit on the tile's unit bits 6..11 (and the
    complex64 in-unit bit) may run on the LDS-free upper kernel: forced on
    (knob 0) it gives the tile kernel's amplitudes (forced off, knob 64) with
    and without a diagonal epilogue, exactly in complex128 (the same
    arithmetic) and to float rounding in complex64 (the tile kernel applies
    complex64 gates on the matrix cores, the upper kernel on the VALU), and
    matches the oracle; a pass with a target inside the contiguous run stays
    on the tile kernel."""
    from zxh_adae_amd import kernels as K
    A = upper_knobs
    A.tune_set(f"tile.seg.{prec}", seg)
    rng = np.random.default_rng(n * 31 + seg + ndiag)
    psi = rand_state(rng, n)
    descs = [(pack(1 << q, False), 1 << q, rand_u(rng)) for q in targets]
    s0, s1, th = diag_window(rng, n, ndiag) if ndiag else ([], [], [])
    want = oracle_u3(psi, n, descs)
    if ndiag:
        want = oracle_diag(want, n, s0, s1, th)
    got, count = {}, {}
    for knob in (64, 0):
        A.tune_set(f"tile.upper.{prec}", knob)
        d = dev(upper_knobs, psi.astype(np.complex64) if prec == "c64" else psi, prec)
        K.set_timing(True)
        K.reset_timing()
        try:
            if ndiag:
                d.fused_pass(descs, s0, s1, th)
            else:
                d.u3_batch(descs)
            count[knob] = K.timing("u3_upper")["launches"]
            passes = K.timing("u3_tile")["launches"] + K.timing("fused")["launches"]
        finally:
            K.set_timing(False)
        got[knob] = d.download()
        assert passes == 1
    assert count == {64: 0, 0: upper}
    if prec == "c128" or not upper:
        assert np.array_equal(got[0], got[64])
    else:
        scale = float(np.max(np.abs(got[64])))
        assert np.max(np.abs(got[0].astype(np.complex128) - got[64])) < 1e-5 * scale
    assert np.max(np.abs(got[0].astype(np.complex128) - want)) < (TOL128 if prec == "c128" else TOL64)


@pytest.mark.parametrize("prec,seg,n,targets,upper", [
    ("c128", 6, 20, [14, 15, 16, 17, 18, 19], 1), ("c128", 6, 22, [21, 16, 12], 1),
    ("c128", 4, 21, [10, 12, 14, 16, 18, 20], 1), ("c64", 6, 20, [0, 14, 15, 16, 17, 18, 19], 1),
    ("c64", 3, 20, [13, 15, 17, 19], 1), ("c128", 6, 20, [2, 15, 19], 0)])
def test_upper_kernel_with_per_pair_selectors(upper_knobs, prec, seg, n, targets, upper):
    """Selectors with bits beyond the target (the side of a pair follows the
    selector's parity over its address, beta = 1 swaps it), including the
    complex64 in-unit bit: the upper kernel (knob 0) gives the tile kernel's
    amplitudes exactly (knob 64) and matches the oracle."""
    from zxh_adae_amd import kernels as K
    A = upper_knobs
    A.tune_set(f"tile.seg.{prec}", seg)
    rng = np.random.default_rng(n * 17 + seg + len(targets))
    psi = rand_state(rng, n)
    others = [q for q in range(n) if q not in targets]
    descs = []
    for q in targets:
        extra = sum(1 << int(b) for b in rng.choice(others, 3, replace=False))
        descs.append((pack((1 << q) | extra, bool(rng.random() < 0.5)), 1 << q, rand_u(rng)))
    want = oracle_u3(psi, n, descs)
    got, count = {}, {}
    for knob in (64, 0):
        A.tune_set(f"tile.upper.{prec}", knob)
        d = dev(upper_knobs, psi.astype(np.complex64) if prec == "c64" else psi, prec)
        K.set_timing(True)
        K.reset_timing()
        try:
            d.u3_batch(descs)
            count[knob] = K.timing("u3_upper")["launches"]
            assert K.timing("u3_tile")["launches"] == 1
        finally:
            K.set_timing(False)
        got[knob] = d.download()
    assert count == {64: 0, 0: upper}
    assert np.array_equal(got[0], got[64])
    assert np.max(np.abs(got[0].astype(np.complex128) - want)) < (TOL128 if prec == "c128" else TOL64)


@pytest.mark.parametrize("n,cx,prec", [(20, 60, "c128"), (22, 200, "c128"), (21, 90, "c64")])
def test_upper_kernel_on_affine_maps(upper_knobs, n, cx, prec):
    """A full layer on a CX-transported map (wide pair offsets, per-pair
    flips): every pass that qualifies runs on the upper kernel (knob 0) with
    the tile kernel's amplitudes (knob 64), exactly in complex128 and to float
    rounding in complex64 (matrix cores against VALU on passes without
    flips), matching the oracle."""
    from zxh_adae_amd import kernels as K
    A = upper_knobs
    rng = np.random.default_rng(n + cx)
    rows, deltas = random_affine(rng, n, cx)
    psi = rand_state(rng, n)
    descs = [(pack(rows[q], bool(rng.random() < 0.5)), deltas[q], rand_u(rng)) for q in range(n)]
    want = oracle_u3(psi, n, descs)
    got, count = {}, {}
    for knob in (64, 0):
        A.tune_set(f"tile.upper.{prec}", knob)
        d = dev(upper_knobs, psi.astype(np.complex64) if prec == "c64" else psi, prec)
        K.set_timing(True)
        K.reset_timing()
        try:
            d.u3_batch(descs)
            count[knob] = K.timing("u3_upper")["launches"]
        finally:
            K.set_timing(False)
        got[knob] = d.download()
    print(n, cx, prec, count)
    assert count[64] == 0
    if prec == "c128":
        assert np.array_equal(got[0], got[64])
    else:
        scale = float(np.max(np.abs(got[64])))
        assert np.max(np.abs(got[0].astype(np.complex128) - got[64])) < 1e-5 * scale
    assert np.max(np.abs(got[0].astype(np.complex128) - want)) < (TOL128 if prec == "c128" else TOL64)


@pytest.mark.parametrize("prec", ["c128", "c64"])
def test_autotuners_follow_their_rules(upper_knobs, prec):
    """tune_diag keeps the fastest diagonal variant; tune_tile keeps the
    fastest layer among the geometries with and without the wide tile, with
    the pass count of layer_passes.  Both set their knobs, and a U3 batch
    under the tuned geometry matches the oracle."""
    from zxh_adae_amd import kernels as K
    A = upper_knobs
    n = 20
    diag_knobs = (f"diag.ept.{prec}", f"diag.nt.{prec}", f"tile.wide.{prec}")
    old = {k: A.tune_get(k) for k in diag_knobs}
    try:
        st = K.DeviceState(n, prec)
        st.scale(0.0)
        d = A.tune_diag(st)
        key = f"ept{d['best']['ept']}_nt{d['best']['nt']}"
        assert d["ms"][key] == min(d["ms"].values())
        assert (A.tune_get(diag_knobs[0]), A.tune_get(diag_knobs[1])) == (d["best"]["ept"], d["best"]["nt"])
        t = A.tune_tile(st)

        def key(k):
            return (int(k[3:].split("_")[0]), 1 if k.endswith("_wide") else 0)
        ms = {key(k): v for k, v in t["layer_ms"].items()}
        assert sorted(ms) == sorted((s, w) for s in A.TILE_SEGS for w in (0, 1))
        best = (t["best"], t["wide"])
        assert ms[best] <= min(ms.values()) + 1e-4  # layer_ms is rounded to 0.1 us
        assert A.tune_get(f"tile.seg.{prec}") == t["best"]
        assert A.tune_get(f"tile.wide.{prec}") == t["wide"]
        for (s, w) in ms:
            name = f"seg{s}" + ("_wide" if w else "")
            assert t["passes"][name] == len(A.layer_passes(n, s, prec == "c64", bool(w)))
        st.free()
        rng = np.random.default_rng(5)
        psi = rand_state(rng, n)
        descs = [(pack(1 << q, False), 1 << q, rand_u(rng)) for q in range(n)]
        want = oracle_u3(psi, n, descs)
        d2 = dev(upper_knobs, psi.astype(np.complex64) if prec == "c64" else psi, prec)
        d2.u3_batch(descs)
        assert np.max(np.abs(d2.download().astype(np.complex128) - want)) < (TOL128 if prec == "c128" else TOL64)
    finally:
        for k, v in old.items():
            A.tune_set(k, v)


def spanned_window(rng, n, count, rank):
    """count P/CP gates whose selectors are random XOR combinations (with
    random beta) of `rank` random address forms: a window of that rank."""
    basis = [int(rng.integers(1, 1 << n)) for _ in range(rank)]

    def sel():
        a = 0
        for b in basis:
            if rng.random() < 0.5:
                a ^= b
        return pack(a or basis[0], bool(rng.random() < 0.3))
    s0, s1, th = [], [], []
    for _ in range(count):
        s0.append(sel())
        s1.append(sel() if rng.random() < 0.5 else SEL_NONE)
        th.append(float(rng.uniform(-math.pi, math.pi)))
    return np.array(s0, dtype=np.uint64), np.array(s1, dtype=np.uint64), np.array(th)


def brute_window(psi, n, s0, s1, th):
    """The window by its definition: amplitude x times exp(i sum of theta_g
    over the gates whose selectors hold at x), a selector holding where the
    parity of alpha & x differs from beta.  (The oracle's kernel-level
    diagonal rejects two-selector gates that mix the low eight and higher
    address bits, as the reference's does; such gates are the point here.)"""
    x = np.arange(1 << n, dtype=np.uint64)

    def holds(sel):
        a = np.uint64(int(sel) & ((1 << 63) - 1))
        par = np.zeros(x.size, dtype=np.uint64)
        v = x & a
        while np.any(v):
            par ^= v & np.uint64(1)
            v >>= np.uint64(1)
        return par.astype(bool) ^ bool(int(sel) >> 63)
    ph = np.zeros(x.size)
    for a, b, t in zip(s0, s1, th):
        on = holds(a)
        if int(b) != int(SEL_NONE):
            on &= holds(b)
        ph += np.where(on, t, 0.0)
    return psi.astype(np.complex128) * np.exp(1j * ph)


@pytest.mark.parametrize("prec,seg,n,targets,rank,upper", [
    ("c128", 6, 14, [0, 2, 5, 9], 6, 64), ("c128", 6, 18, [12, 13, 14, 15, 16, 17], 10, 0),
    ("c128", 4, 16, [1, 7, 15], 8, 64), ("c64", 3, 15, [0, 3, 8, 14], 9, 64),
    ("c64", 6, 19, [0, 14, 15, 16, 17, 18], 10, 0), ("c128", 6, 15, [3, 14], 11, 64),
    ("c64", 4, 16, [2, 11], 12, 64)])
def test_fused_pass_with_a_tabulated_window(upper_knobs, prec, seg, n, targets, rank, upper):
    """A window whose selectors are parities of many address bits (beyond the
    phase-polynomial form) rides the tile pass as a table over its selector
    forms (tile kernel and, forced on, the upper kernel): the oracle's U3
    batch then window, complex128 1e-12, complex64 float rounding; windows
    spanning more than ten dimensions are rejected."""
    from zxh_adae_amd import ZXHError
    A = upper_knobs
    A.tune_set(f"tile.seg.{prec}", seg)
    A.tune_set(f"tile.upper.{prec}", upper)
    rng = np.random.default_rng(n * 7 + rank)
    psi = rand_state(rng, n)
    descs = [(pack(1 << q, False), 1 << q, rand_u(rng)) for q in targets]
    s0, s1, th = spanned_window(rng, n, 40, rank)
    d = dev(upper_knobs, psi.astype(np.complex64) if prec == "c64" else psi, prec)
    if rank > 10:
        with pytest.raises(ZXHError):
            d.fused_pass(descs, s0, s1, th)
        return
    want = brute_window(oracle_u3(psi, n, descs), n, s0, s1, th)
    d.fused_pass(descs, s0, s1, th)
    got = d.download().astype(np.complex128)
    assert np.max(np.abs(got - want)) < (TOL128 if prec == "c128" else TOL64)


@pytest.fixture
def wide_knobs(gpu_engine):
    """Turns the wide (8192-unit) tile on for a test and restores the knobs."""
    from zxh_adae_amd import autotune
    knobs = ("tile.wide.c128", "tile.wide.c64", "tile.seg.c128", "tile.seg.c64")
    old = {k: autotune.tune_get(k) for k in knobs}
    autotune.tune_set("tile.wide.c128", 1)
    autotune.tune_set("tile.wide.c64", 1)
    yield autotune
    for k, v in old.items():
        autotune.tune_set(k, v)


@pytest.mark.parametrize("prec,n,targets,ndiag", [
    ("c128", 13, [4, 5, 6, 7, 8, 9, 10, 11, 12], 0), ("c128", 16, [15, 4, 9, 12, 13, 7, 10, 11, 14], 40),
    ("c128", 18, [17, 16, 5], 64), ("c128", 20, [12, 13, 14, 15, 16, 17, 18, 19, 11], 100),
    ("c64", 14, [5, 6, 7, 8, 9, 10, 11, 12, 13], 0), ("c64", 17, [16, 5, 9, 12, 13, 7, 10, 11, 14], 40),
    ("c64", 21, [20, 19, 18, 17, 16, 15, 14, 13, 12], 64)])
def test_wide_tile_pass_matches_the_oracle(wide_knobs, prec, n, targets, ndiag):
    """Nine targets above the 16-unit run exceed the 4096-unit tile's slots
    at any geometry with a run of 16 or more units; with the wide tile on they
    run as one 8192-unit tile pass, then the polynomial window, and match the
    oracle's U3 batch and diagonal sweep (complex128 1e-12, complex64 float
    rounding)."""
    wide_knobs.tune_set(f"tile.seg.{prec}", 4)
    rng = np.random.default_rng(n * 3 + ndiag)
    psi = rand_state(rng, n)
    descs = [(pack(1 << q, False), 1 << q, rand_u(rng)) for q in targets]
    s0, s1, th = diag_window(rng, n, ndiag) if ndiag else ([], [], [])
    want = oracle_u3(psi, n, descs)
    if ndiag:
        want = oracle_diag(want, n, s0, s1, th)
    d = dev(wide_knobs, psi.astype(np.complex64) if prec == "c64" else psi, prec)
    d.fused_pass(descs, s0, s1, th)
    got = d.download().astype(np.complex128)
    assert np.max(np.abs(got - want)) < (TOL128 if prec == "c128" else TOL64)


@pytest.mark.parametrize("prec,n", [("c128", 20), ("c128", 23), ("c64", 21)])
def test_u3_batch_on_every_qubit_with_the_wide_tile(wide_knobs, prec, n):
    """All targets of a layer: the first pass as the 4096-unit tile takes it,
    the rest in wide passes of nine; against the oracle."""
    from zxh_adae_amd import kernels as K
    rng = np.random.default_rng(n)
    psi = rand_state(rng, n)
    descs = [(pack(1 << q, False), 1 << q, rand_u(rng)) for q in range(n)]
    want = oracle_u3(psi, n, descs)
    d = dev(wide_knobs, psi.astype(np.complex64) if prec == "c64" else psi, prec)
    K.reset_timing()
    K.set_timing(True)
    try:
        d.u3_batch(descs)
        passes = K.timing("u3_tile")["launches"]
    finally:
        K.set_timing(False)
    got = d.download().astype(np.complex128)
    assert np.max(np.abs(got - want)) < (TOL128 if prec == "c128" else TOL64)
    assert passes >= 2


@pytest.mark.parametrize("prec,n,seed,count,above_run", [
    ("c128", 20, 1, 12, False), ("c128", 22, 2, 17, False), ("c128", 22, 3, 13, True),
    ("c64", 21, 4, 15, False), ("c64", 23, 5, 11, True),
    ("c128", 23, 3, 18, True), ("c64", 24, 5, 18, True)])
def test_u3_batch_on_random_targets_follows_the_plan(wide_knobs, prec, n, seed, count, above_run):
    """Random target subsets (mixed, or all above the wide tile's run, where
    the split may be wide only): as many passes as the engine's published
    split, and the oracle's result."""
    from zxh_adae_amd import kernels as K
    rng = np.random.default_rng(seed)
    c64 = prec == "c64"
    lo = (5 if c64 else 4) if above_run else 0
    targets = sorted(int(q) for q in rng.choice(np.arange(lo, n), size=count, replace=False))
    plan = wide_knobs.tile_layer_plan(n, targets, c64)
    assert sorted(q for p in plan for q in p) == (targets if plan else [])
    psi = rand_state(rng, n)
    descs = [(pack(1 << q, False), 1 << q, rand_u(rng)) for q in targets]
    want = oracle_u3(psi, n, descs)
    d = dev(wide_knobs, psi.astype(np.complex64) if c64 else psi, prec)
    K.reset_timing()
    K.set_timing(True)
    try:
        d.u3_batch(descs)
        passes = K.timing("u3_tile")["launches"]
    finally:
        K.set_timing(False)
    got = d.download().astype(np.complex128)
    assert np.max(np.abs(got - want)) < (TOL64 if c64 else TOL128)
    if plan:
        assert passes == len(plan)


def test_fused_pass_rejects_multibit_and_overfull(gpu_engine):
    from zxh_adae_amd import ZXHError
    rng = np.random.default_rng(1)
    d = dev(gpu_engine, rand_state(rng, 13))
    with pytest.raises(ZXHError):
        d.fused_pass([(pack(3, False), 1, rand_u(rng))], [], [], [])
    # more targets above the low bits than either tile takes (the 4096-unit
    # tile: 12 - seg; the wide tile: nine, none below unit bit 4)
    with pytest.raises(ZXHError):
        d.fused_pass([(pack(1 << q, False), 1 << q, rand_u(rng)) for q in range(3, 13)], [], [], [])


@pytest.mark.parametrize("prec", ["c128", "c64"])
def test_gather_reads_the_addressed_amplitudes(gpu_engine, prec):
    from zxh_adae_amd import ZXHError
    rng = np.random.default_rng(44)
    dt = np.complex64 if prec == "c64" else np.complex128
    psi = rand_state(rng, 12, dt)
    d = dev(gpu_engine, psi, prec)
    idx = np.concatenate([[0, 4095], rng.integers(0, 4096, 300)]).astype(np.uint64)
    assert np.array_equal(d.gather(idx), psi[idx.astype(np.int64)])
    assert d.gather(np.zeros(0, dtype=np.uint64)).size == 0
    with pytest.raises(ZXHError):
        d.gather(np.array([4096], dtype=np.uint64))


# ZXHSIM_SEG_MIN_I read by the drop-in at construction, with the reference's
# parse rules (parse_env_size_t, sv_common.cpp:49-66): decimal strtoull of the
# whole string (leading blanks and a sign accepted); 0, empty or trailing text
# give the default 18.  The indices must equal the reference's for that
# segment size on the same state and seed.
@pytest.mark.parametrize("env,seg", [("1", 1), ("2", 2), ("4", 4), ("18", 18), (" 2", 2), ("+4", 4),
                                     ("0", 18), ("4x", 18), ("", 18), (None, 18)])
def test_seg_min_i_environment_through_zxh(gpu_engine, monkeypatch, env, seg):
    if env is None:
        monkeypatch.delenv("ZXHSIM_SEG_MIN_I", raising=False)
    else:
        monkeypatch.setenv("ZXHSIM_SEG_MIN_I", env)
    gold = np.load(os.path.join(GOLDEN, f"samp_seg{seg}.npz"))
    for key in gold.files:
        name, seed = key.split("__")
        psi = SAMPLER_STATES[name]
        m = psi.size.bit_length() - 1
        if m == 0:
            continue  # the drop-in needs at least one qubit
        sim = gpu_engine.ZXH(m, False, True)
        sim.execute()
        sim.set_phys_state(psi)
        sim.set_seed(int(seed))
        got = sim.sample_real(gold[key].size)
        assert np.array_equal(got, gold[key]), (env, name, seed)
        del sim


# Kernel-level measure (measure.h:12-51): the reference's sampler on a caller's
# state and affine map, results packed into virtual words on the device.  The
# oracle runs lazy-mode circuits (X/CX transport the map: A is not the
# identity, b is not zero) and its packed results are the expected words.
@pytest.mark.parametrize("n,count,seed", [(10, 120, 1), (14, 300, 2), (17, 400, 3)])
def test_kernel_level_affine_measure_matches_oracle(gpu_engine, n, count, seed):
    from zxh_adae_amd import workloads as W
    from zxh_adae_amd.kernels import DeviceState
    ora = O.OracleSim(n, False, False)
    ora.gates(W.random_mixed(n, count, seed))
    ora.execute()
    psi = ora.phys_state()
    rows, b, m = ora.affine()
    assert psi.size == 1 << m
    assert any(int(r) != (1 << q) for q, r in enumerate(rows)) or any(b)
    shots = 3000
    ora.set_seed(500 + seed)
    ora.measure(shots)
    wc = (n + 63) // 64
    want = np.zeros(shots * wc, dtype=np.uint64)
    assert ora.lib.zo_get_results(ora.h, O._u64ptr(want), shots) == 0
    want_real = ora.real_results(shots)
    d = DeviceState(m, "c128")
    d.upload(psi)
    packed, real = d.measure(18, rows, b, shots, seed=500 + seed)
    assert np.array_equal(real, want_real)
    assert np.array_equal(packed.reshape(-1), want)


def test_dev_realloc_set_zero_and_mem_set(gpu_engine):
    """worker_realloc keeps the common prefix and zero-fills the rest;
    worker_set_zero clears a range; worker_mem_set writes amplitude 0."""
    import ctypes as C
    from zxh_adae_amd import _lib
    lib = _lib.load()
    n = 1000
    host = (np.arange(n) + 1j * np.arange(n)).astype(np.complex128)
    p = C.c_void_p()
    assert lib.zxh_dev_alloc(host.nbytes, C.byref(p)) == 0
    assert lib.zxh_dev_upload(p, host.ctypes.data, host.nbytes) == 0
    q = C.c_void_p()
    assert lib.zxh_dev_realloc(p, host.nbytes, 2 * host.nbytes, C.byref(q)) == 0
    out = np.zeros(2 * n, dtype=np.complex128)
    assert lib.zxh_dev_download(out.ctypes.data, q, out.nbytes) == 0
    assert np.array_equal(out[:n], host) and not np.any(out[n:])
    assert lib.zxh_dev_set_zero(q, _lib.ZXH_C128, 10, 20) == 0
    assert lib.zxh_dev_mem_set(q, _lib.ZXH_C128, 0.5, -0.25) == 0
    assert lib.zxh_dev_download(out.ctypes.data, q, out.nbytes) == 0
    want = np.concatenate([host, np.zeros(n, dtype=np.complex128)])
    want[10:20] = 0
    want[0] = 0.5 - 0.25j
    assert np.array_equal(out, want)
    r = C.c_void_p()
    assert lib.zxh_dev_realloc(q, 2 * host.nbytes, 100 * 16, C.byref(r)) == 0
    small = np.zeros(100, dtype=np.complex128)
    assert lib.zxh_dev_download(small.ctypes.data, r, small.nbytes) == 0
    assert np.array_equal(small, want[:100])
    z = C.c_void_p(1)
    assert lib.zxh_dev_realloc(r, 100 * 16, 0, C.byref(z)) == 0
    assert z.value is None
