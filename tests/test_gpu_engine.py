"""End-to-end parity of the ZXH object on the GPU: against the reference's
golden outputs, the oracle on random circuits in every execution mode, and
size-independent properties at full sizes."""
import json
import math
import os

import numpy as np
import pytest

import oracle_py as O
from conftest import GOLDEN
from zxh_adae_amd import workloads as W

pytestmark = pytest.mark.gpu

MANIFEST = json.load(open(os.path.join(GOLDEN, "manifest.json")))
TOL128 = 1e-12


POLICIES = ["fused", "reference"]


def make_sim(Z, n, dx, eager, gates, prec="c128", policy="fused"):
    s = Z.ZXH(n, dx, eager, precision=prec, policy=policy)
    s.gates(gates)
    s.execute()
    return s


@pytest.mark.parametrize("policy", POLICIES)
@pytest.mark.parametrize("name", MANIFEST["circuits"])
def test_golden_circuit(gpu_engine, name, policy):
    Z = gpu_engine
    d = np.load(os.path.join(GOLDEN, f"circ_{name}.npz"))
    n, dx, eager, shots, seed, M = (int(v) for v in d["meta"])
    gates = [(int(r[0]), int(r[1]), int(r[2]), r[3], r[4], r[5]) for r in d["gates"]]
    s = make_sim(Z, n, bool(dx), bool(eager), gates, policy=policy)
    assert s.required_M() == M
    st = s.get_state_array()
    if "state" in d:
        assert np.max(np.abs(st - d["state"])) < TOL128
    else:
        idx = d["probe_idx"].astype(np.int64)
        assert np.max(np.abs(st[idx] - d["probe_val"])) < TOL128
    # sampling: bit-exact against the oracle's sampler on the GPU's own state
    # (identical probability vector and uniform stream) ...
    s.set_seed(seed)
    real = s.sample_real(shots)
    phys = s.phys_state()
    seg = min(M, 18)
    want_real = O.oracle_sample_state(phys, seg, O.mix_seed(seed, 0), shots)
    assert np.array_equal(real, want_real)
    # ... the virtual bit rows packed on the device: exactly the oracle's
    # affine map (same recording rules) applied to these real addresses ...
    bits = s.results_bits()
    o = O.OracleSim(n, bool(dx), bool(eager))
    o.gates(gates)
    o.execute()
    rows, b, _ = o.affine()
    want_bits = np.zeros_like(bits)
    for q in range(n):
        par = np.array([bin(int(r) & int(rows[q])).count("1") & 1 for r in real], dtype=np.uint8)
        want_bits[:, q] = par ^ b[q]
    assert np.array_equal(bits, want_bits)
    # ... and against the reference's own Sampling() rows
    mismatch = np.mean(np.any(bits != d["samples"], axis=1))
    assert mismatch <= 0.01, mismatch


@pytest.mark.parametrize("policy", POLICIES)
@pytest.mark.parametrize("seed", range(12))
def test_random_circuits_vs_oracle(gpu_engine, seed, policy):
    Z = gpu_engine
    n = 4 + seed % 11
    eager = seed % 2 == 1
    dx = seed % 3 == 0
    gs = W.random_mixed(n, 40 + 8 * seed, 5000 + seed)
    s = make_sim(Z, n, dx, eager, gs, policy=policy)
    o = O.OracleSim(n, dx, eager)
    o.gates(gs)
    o.execute()
    assert s.required_M() == o.required_M()
    assert np.max(np.abs(s.get_state_array() - o.get_state())) < TOL128


@pytest.mark.parametrize("policy", POLICIES)
def test_layers_20q_plain_full_config(gpu_engine, policy):
    """20 qubits, 20 layers of U3-all / CP ladder / P-all, 1024 shots."""
    Z = gpu_engine
    d = np.load(os.path.join(GOLDEN, "circ_layers_n20_plain.npz"))
    gates = [(int(r[0]), int(r[1]), int(r[2]), r[3], r[4], r[5]) for r in d["gates"]]
    s = make_sim(Z, 20, False, True, gates, policy=policy)
    st = s.get_state_array()
    idx = d["probe_idx"].astype(np.int64)
    assert np.max(np.abs(st[idx] - d["probe_val"])) < TOL128
    assert abs(np.sum(np.abs(st) ** 2) - d["state_norm2"][0]) < 1e-12


def test_complex64_infidelity_and_max_error(gpu_engine):
    Z = gpu_engine
    gs = W.random_layers(14, 6, 21)
    a = make_sim(Z, 14, False, True, gs, "c128").get_state_array()
    b = make_sim(Z, 14, False, True, gs, "c64").get_state_array()
    infid = 1.0 - abs(np.vdot(a, b))
    assert infid < 1e-6
    assert np.max(np.abs(a - b)) < 1e-5


def test_errors_are_raised_not_aborts(gpu_engine):
    Z = gpu_engine
    s = Z.ZXH(3)
    with pytest.raises(Z.ZXHError):
        s.H(5)
    with pytest.raises(Z.ZXHError):
        s.get_state()


@pytest.mark.parametrize("n", [26])
def test_large_uniform_diag_closed_form(gpu_engine, n):
    """H on every qubit then one diagonal batch: amplitude x is
    exp(i theta(x)) / 2^(n/2) with theta summed over the batch."""
    Z = gpu_engine
    batch = W.diag_batch(n, 256, 2)
    s = make_sim(Z, n, False, True, W.hadamard_all(n) + [(W.BARRIER, 0, 0, 0, 0, 0)] + batch)
    phys = s.phys_state()
    rng = np.random.default_rng(0)
    idx = rng.integers(0, 1 << n, 4096)
    theta = np.zeros(idx.size)
    for (code, cq, q, th, _, _) in batch:
        if code == W.P:
            theta += np.where((idx >> q) & 1, th, 0.0)
        else:
            theta += np.where(((idx >> q) & 1) & ((idx >> cq) & 1), th, 0.0)
    # H's matrix is not exactly real; build its product-state amplitude exactly
    h = O.u3_matrix(math.pi / 2, math.pi, 0.0)
    amp0 = h[0, 0] ** n  # every amplitude of H^n|0> has the same value up to rounding
    ref = amp0 * np.exp(1j * theta)
    assert np.max(np.abs(phys[idx] - ref)) < 1e-12


def test_large_u3_round_trip(gpu_engine):
    """U3 layer then its inverse layer returns the state (28 qubits)."""
    Z = gpu_engine
    n = 28
    fwd = W.u3_all(n, 9)
    inv = [(W.U3, 0, q, -th, -ph, -la) for (_, _, q, th, la, ph) in fwd]
    base = W.hadamard_all(n) + [(W.BARRIER, 0, 0, 0, 0, 0)] + W.diag_batch(n, 64, 4)
    a = make_sim(Z, n, False, True, base).phys_state()
    s = make_sim(Z, n, False, True, base + [(W.BARRIER, 0, 0, 0, 0, 0)] + fwd + [(W.P, 0, 0, 0.0, 0, 0)] + inv)
    b = s.phys_state()
    assert np.max(np.abs(a - b)) < 1e-12


@pytest.mark.parametrize("n,gates_fn", [(16, lambda n: W.qft(n)), (14, lambda n: W.random_layers(n, 5, 3)),
                                       (13, lambda n: W.random_mixed(n, 150, 77, kinds=(W.H, W.U3, W.P, W.CP, W.RZ, W.Z,
                                                                                       W.BARRIER)))])
def test_fused_policy_matches_oracle_and_reference_policy(gpu_engine, tile_seg, n, gates_fn):
    """Plain-method circuits through the fused policy (U3 batches merged into
    tile passes with diagonal epilogues) agree with the oracle and with the
    reference-policy run."""
    Z = gpu_engine
    gs = gates_fn(n)
    o = O.OracleSim(n, False, True)
    o.gates(gs)
    o.execute()
    want = o.get_state()
    for prec, tol in (("c128", TOL128), ("c64", 2e-5)):
        a = make_sim(Z, n, False, True, gs, prec, "fused").get_state_array()
        b = make_sim(Z, n, False, True, gs, prec, "reference").get_state_array()
        assert np.max(np.abs(a - want)) < tol
        assert np.max(np.abs(b - want)) < tol


def test_fused_policy_composes_across_commuting_windows(gpu_engine):
    """U3(q); P(q + 6); U3(q)^-1: the P commutes with both U3 gates, so the two
    U3 matrices meet in one pass (composed exactly, no identity skip) and must
    agree with the oracle, which applies them in separate sweeps."""
    Z = gpu_engine
    n = 12
    gs = W.hadamard_all(n) + [(W.BARRIER, 0, 0, 0, 0, 0)]
    for q in range(6):
        gs += [(W.U3, 0, q, 0.4, 0.3, 0.2), (W.P, 0, q + 6, 0.7, 0, 0), (W.U3, 0, q, -0.4, -0.2, -0.3)]
    o = O.OracleSim(n, False, True)
    o.gates(gs)
    o.execute()
    s = make_sim(Z, n, False, True, gs, "c128", "fused")
    assert np.max(np.abs(s.get_state_array() - o.get_state())) < TOL128


def test_open_batch_keeps_each_runs_product(gpu_engine):
    """Lazy mode (CX as affine updates).  The fused policy keeps a U3 batch
    open across later gates that commute with it, but one qubit's products of
    two runs stay apart: ry(0.3) and, one run later, ry(-0.3 + 3e-6) on qubit 0
    are applied as two matrices, as the reference's batches do, rather than
    as one near-identity product that the 1e-5 identity skip would drop."""
    Z = gpu_engine
    n = 14
    gs = W.hadamard_all(n) + [(W.BARRIER, 0, 0, 0, 0, 0)]
    gs += [(W.U3, 0, 0, 0.3, 0.0, 0.0), (W.CX, 1, 2, 0, 0, 0), (W.U3, 0, 0, -0.3 + 3e-6, 0.0, 0.0),
           (W.U3, 0, 3, 0.5, 0.1, 0.2), (W.CX, 3, 4, 0, 0, 0), (W.U3, 0, 5, 0.7, 0.0, 0.3),
           (W.U3, 0, 3, 0.2, 0.4, 0.0), (W.P, 0, 6, 0.9, 0, 0), (W.U3, 0, 7, 1.1, 0.2, 0.1)]
    for dx, eager in ((False, False), (False, True)):
        o = O.OracleSim(n, dx, eager)
        o.gates(gs)
        o.execute()
        for policy in POLICIES:
            s = make_sim(Z, n, dx, eager, gs, "c128", policy)
            assert np.max(np.abs(s.phys_state() - o.phys_state())) < TOL128, (eager, policy)


def two_local_ladder(n, layers, seed):
    """The reference's VQE two-local workload shape: ry on qubit k, then CX
    from every lower qubit onto k, so that no two ry gates are adjacent."""
    rng = np.random.default_rng(seed)
    gs = []
    for _ in range(layers):
        for k in range(n):
            gs.append((W.U3, 0, k, float(rng.uniform(0, 2 * math.pi)), 0.0, 0.0))
            gs += [(W.CX, j, k, 0, 0, 0) for j in range(k)]
    return gs


@pytest.mark.parametrize("n", [16, 20])
def test_open_batch_spans_a_cx_ladder_layer(gpu_engine, n):
    """Each layer's ry gates, separated by CX gates, reach the device as one
    batch of wide pair offsets: a few tile passes per layer instead of one
    batch per gate.  Parity against the oracle and the reference policy
    (which batches runs only)."""
    Z = gpu_engine
    gs = W.hadamard_all(n) + [(W.BARRIER, 0, 0, 0, 0, 0)] + two_local_ladder(n, 3, n)
    o = O.OracleSim(n, False, False)
    o.gates(gs)
    o.execute()
    launches = {}
    for policy in POLICIES:
        s = Z.ZXH(n, False, False, precision="c128", policy=policy)
        s.gates(gs)
        s.set_timing(True)
        s.execute()
        s.set_timing(False)
        assert np.max(np.abs(s.phys_state() - o.phys_state())) < TOL128, policy
        launches[policy] = sum(s.timing(f)["launches"] for f in ("u3_tile", "u3_pair", "fused"))
    print(launches)
    # the fused policy: n expansions plus at most 3 passes per layer
    assert launches["fused"] <= n + 3 * 3
    assert launches["reference"] >= n + 3 * (n - 1)


def _swept_bytes(s, fams=("u3_tile", "u3_pair", "fused", "diag")):
    return sum(s.timing(f)["bytes"] for f in fams)


@pytest.mark.parametrize("name", ["qaoa_n20", "vqe_two_local_n20"])
def test_open_batch_on_reference_workloads(gpu_engine, name):
    """The reference's QAOA and two-local workloads at 20 qubits in lazy mode:
    the fused policy's state matches the oracle and moves far fewer bytes
    (the sweeps weighted by the state size at the time) than batching runs
    only."""
    from zxh_adae_amd import frontend as F
    Z = gpu_engine
    path = os.path.join(os.path.dirname(__file__), "golden", "workloads", name + ".qasm")
    circ = F.load_qasm(path)
    gs = F.to_gate_tuples(F.lower_circuit(circ))
    n = circ.num_qubits
    o = O.OracleSim(n, False, False)
    o.gates(gs)
    o.execute()
    counts = {}
    for policy in POLICIES:
        s = Z.ZXH(n, False, False, precision="c128", policy=policy)
        s.gates(gs)
        s.set_timing(True)
        s.execute()
        s.set_timing(False)
        assert np.max(np.abs(s.phys_state() - o.phys_state())) < TOL128, policy
        counts[policy] = _swept_bytes(s)
    print(name, counts, counts["reference"] / counts["fused"])
    assert counts["fused"] * 3 <= counts["reference"]


SOUP_KINDS = (W.H, W.U3, W.P, W.CP, W.RZ, W.Z, W.X, W.CX, W.CRZ, W.RX)
DIAG_SOUP_KINDS = (W.H, W.U3, W.P, W.CP, W.RZ, W.Z, W.CRZ, W.RX)


@pytest.mark.parametrize("n,count,seed", [(13, 500, 81), (16, 700, 82), (19, 400, 83)])
def test_plain_soups_with_fans(gpu_engine, n, count, seed):
    """Plain method, single-qubit and diagonal gates only: every diagonal
    gate that sits between two gates on its qubits becomes a fan, in passes
    that keep circuit order (low and high targets mixed).  Both policies and
    both geometries' extremes against the oracle."""
    from zxh_adae_amd import autotune as A
    Z = gpu_engine
    gs = W.random_mixed(n, count, seed, kinds=DIAG_SOUP_KINDS)
    o = O.OracleSim(n, False, True)
    o.gates(gs)
    o.execute()
    want = o.get_state()
    old = A.tune_get("tile.seg.c128")
    try:
        for seg in (3, 6):
            A.tune_set("tile.seg.c128", seg)
            for policy in POLICIES:
                s = make_sim(Z, n, False, True, gs, policy=policy)
                assert np.max(np.abs(s.get_state_array() - want)) < TOL128, (seg, policy)
    finally:
        A.tune_set("tile.seg.c128", old)


@pytest.mark.parametrize("n,count,seed", [(13, 400, 71), (14, 600, 72), (16, 500, 73), (17, 300, 74)])
@pytest.mark.parametrize("mode", ["lazy", "nox", "exp"])
def test_random_soups_through_the_scheduler(gpu_engine, n, count, seed, mode):
    """Barrier-free random circuits over every gate kind, wide enough for tile
    passes: the open batch, wide pair offsets, per-pair flips, cleaned
    fillers and two-bit polynomial windows all meet here.  Both policies
    against the oracle (state, width and map); complex64 within 2e-5."""
    Z = gpu_engine
    dx, eager = {"lazy": (False, False), "nox": (True, False), "exp": (False, True)}[mode]
    gs = W.random_mixed(n, count, seed, kinds=SOUP_KINDS)
    o = O.OracleSim(n, dx, eager)
    o.gates(gs)
    o.execute()
    want = o.get_state()
    for policy in POLICIES:
        s = make_sim(Z, n, dx, eager, gs, policy=policy)
        assert s.required_M() == o.required_M()
        assert np.max(np.abs(s.get_state_array() - want)) < TOL128, policy
    s = make_sim(Z, n, dx, eager, gs, prec="c64")
    assert np.max(np.abs(s.get_state_array() - want)) < 2e-5


def textbook_qft(n, seed=None):
    """The QFT in the order of the reference's qft workload: for t from the
    top qubit down, the controlled phases from every higher qubit onto t,
    then H(t).  Each H then depends on the phases before it (no two commute
    past each other).  With a seed, random U3 gates replace the H gates."""
    rng = np.random.default_rng(seed) if seed is not None else None
    gs = []
    for t in range(n - 1, -1, -1):
        for k in range(n - 1, t, -1):
            gs.append((W.CP, k, t, math.pi / 2 ** (k - t), 0, 0))
        if rng is None:
            gs.append((W.H, 0, t, 0, 0, 0))
        else:
            a, b, c = (float(x) for x in rng.uniform(-math.pi, math.pi, 3))
            gs.append((W.U3, 0, t, a, b, c))
    return gs


@pytest.mark.parametrize("n,seed,seg", [(16, None, 6), (20, None, 4), (18, 5, 3), (21, 6, 6), (22, None, 5)])
def test_fans_let_dependent_gates_share_a_pass(gpu_engine, n, seed, seg):
    """Plain method, complex128: the controlled phases between two H gates
    that involve the later one's target become that gate's fan, so up to one
    pass per tile-slot group instead of one pass per gate.  Parity against
    the oracle and the reference policy; complex64 (no fans) as well."""
    from zxh_adae_amd import autotune as A
    Z = gpu_engine
    gs = W.hadamard_all(n) + [(W.BARRIER, 0, 0, 0, 0, 0)] + textbook_qft(n, seed)
    o = O.OracleSim(n, False, True)
    o.gates(gs)
    o.execute()
    want = o.get_state()
    passes = {}
    old = A.tune_get("tile.seg.c128")
    try:
        A.tune_set("tile.seg.c128", seg)
        for policy in POLICIES:
            s = Z.ZXH(n, False, True, precision="c128", policy=policy)
            s.gates(gs)
            s.set_timing(True)
            s.execute()
            s.set_timing(False)
            assert np.max(np.abs(s.get_state_array() - want)) < TOL128, policy
            passes[policy] = sum(s.timing(f)["launches"] for f in ("u3_tile", "fused", "diag"))
    finally:
        A.tune_set("tile.seg.c128", old)
    print(n, seed, seg, passes)
    # the H layer, then one pass per slot group of the QFT (the low targets
    # fill the contiguous run in circuit order)
    assert passes["fused"] <= 2 + -(-n // (12 - seg)) + 1
    assert passes["fused"] * 2 < passes["reference"]
    c64 = make_sim(Z, n, False, True, gs, prec="c64")
    assert np.max(np.abs(c64.get_state_array() - want)) < 2e-5


def transported_windows(n, layers, width, seed):
    """A CX network (every row of the affine map several bits wide), then
    layers of H on a few qubits followed by a window of P and CP gates on
    `width` random qubits: in lazy mode each H cluster has wide pair offsets
    or per-pair selectors, and each window's selectors are parities of many
    address bits, spanning up to `width` dimensions."""
    rng = np.random.default_rng(seed)
    gs = [(W.H, 0, q, 0, 0, 0) for q in range(n)]
    for _ in range(3):
        for q in range(n - 1):
            gs.append((W.CX, q, q + 1, 0, 0, 0))
        for q in range(0, n - 2, 2):
            gs.append((W.CX, q + 2, q, 0, 0, 0))
    for _ in range(layers):
        for q in rng.choice(n, 3, replace=False):
            gs.append((W.H, 0, int(q), 0, 0, 0))
        qs = [int(q) for q in rng.choice(n, width, replace=False)]
        for k, q in enumerate(qs):
            gs.append((W.P, 0, q, float(rng.uniform(-math.pi, math.pi)), 0, 0))
            if k:
                gs.append((W.CP, qs[k - 1], q, float(rng.uniform(-math.pi, math.pi)), 0, 0))
    return gs


@pytest.mark.parametrize("n,width,seed", [(16, 4, 81), (17, 9, 82), (18, 10, 83), (16, 12, 84)])
def test_transported_windows_ride_the_tile_passes(gpu_engine, n, width, seed):
    """Windows after CX transport (selectors over many address bits) as table
    epilogues of the preceding cluster's last tile pass: the state matches
    the oracle (complex128 1e-12, complex64 2e-5) and, while the windows span
    at most ten dimensions, no diagonal sweep of their own runs but for the
    first window, whose commuting part is flushed before any cluster.  The
    open batch merges the windows of neighbouring layers, and a merged window
    spanning more than ten dimensions falls back to a sweep: with wider
    windows only some ride."""
    Z = gpu_engine
    gs = transported_windows(n, 12, width, seed)
    o = O.OracleSim(n, False, False)
    o.gates(gs)
    o.execute()
    want = o.get_state()
    s = Z.ZXH(n, False, False)
    s.gates(gs)
    s.execute()
    assert s.required_M() == o.required_M()
    assert np.max(np.abs(s.get_state_array() - want)) < TOL128
    s.reset_timing()
    s.set_timing(True)
    s.execute()
    s.set_timing(False)
    assert np.max(np.abs(s.get_state_array() - want)) < TOL128
    sweeps = s.timing("diag")["launches"]
    if width <= 4:
        assert sweeps <= 1, sweeps
    elif width <= 10:
        assert sweeps < 12, sweeps
    r = make_sim(Z, n, False, False, gs, policy="reference")
    assert np.max(np.abs(r.get_state_array() - want)) < TOL128
    c = make_sim(Z, n, False, False, gs, prec="c64")
    assert np.max(np.abs(c.get_state_array() - want)) < 2e-5


@pytest.mark.parametrize("n,prec", [(16, "c128"), (21, "c128"), (22, "c64"), (24, "c128")])
def test_qft_with_the_wide_tile(gpu_engine, n, prec):
    """The fused policy with the wide (8192-unit) tile on: passes of up to
    nine targets above the contiguous run, the last with the controlled-phase
    window as its epilogue; the QFT against the oracle."""
    from zxh_adae_amd import autotune as A
    Z = gpu_engine
    old = {k: A.tune_get(k) for k in ("tile.wide.c128", "tile.wide.c64")}
    try:
        A.tune_set("tile.wide.c128", 1)
        A.tune_set("tile.wide.c64", 1)
        gs = W.qft(n)
        o = O.OracleSim(n, False, True)
        o.gates(gs)
        o.execute()
        want = o.get_state()
        s = make_sim(Z, n, False, True, gs, prec=prec)
        got = s.get_state_array()
        assert np.max(np.abs(got - want)) < (TOL128 if prec == "c128" else 2e-5)
    finally:
        for k, v in old.items():
            A.tune_set(k, v)
