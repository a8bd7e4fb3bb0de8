"""Benchmark of the MI355X state-vector engine (one JSON line on stdout).

Step: one execute() of the 30-qubit QFT-structured circuit (H on every
target followed by CP(pi/2^k) to all lower qubits, no swaps: 30 H + 435 CP)
in complex128 through the engine's drop-in ZXH object with the default fused
policy: the reference's own U3 clusters and diagonal windows, merged into
tile passes where they commute.  The state is resident in HBM; execute()
starts from the reset state |0...0> of 16 GiB, as the reference's execute()
does: the fused policy merges the reset into the first tile pass, which forms
that state in registers instead of reading it (every amplitude still goes
through the pass's gates and is stored).  value = gates / second of the timed
steps.

Reported beside it (same JSON line):
  roofline      dominant kernel of the step: algorithmic bytes per launch /
                average launch time (HIP events on the launch stream)
  reference_policy  the same step with one pass per flushed batch, as the
                reference runs it (30 U3 passes + 29 diagonal sweeps)
  autotune      the diagonal-sweep variant and the tile geometry chosen on the
                resident state before the timed steps
  diag_batches  one fused sweep per batch of B random P/CP gates, B in
                {1,4,16,64,256}, uniform-superposition input, c128 and c64
  u3_all        U3 on all 30 qubits as tile passes
  c64           the same QFT step with a complex64 state
  sampling      1e6 shots on the 30-qubit state
  c5            33q complex64 and 32q complex128 random circuits of 5-qubit
                Haar blocks (matrix cores), diagonal batches and MCX cascades,
                40 layers, then 1e6 shots
  adjoint_c3    two-local ansatz, 28 qubits complex64, 448 parameters, 200
                Z/ZZ terms: forward and adjoint backward times, grads/s, and
                the gradient error and state infidelity against complex128
  batch_c4      512 parameter sets (16..24 qubits, 2..10 layers, 50 terms,
                complex64) in one batched engine replayed from a hipGraph
  zxh_workloads the reference's own 30-45-qubit benchmark circuits (QASM),
                execute() time in lazy ZXH mode and in the plain method
  c1            the 20-qubit plain random circuit (1180 gates) and
                Sampling(1024) on the GPU, execute() and Sampling() timed apart
  cpu_baseline  the oracle's restatement of the reference's OpenMP CPU path
                (oracle/liboracle_base.so) in child processes on this host, at
                full threads and at one thread: C1 in full, the 30-qubit U3 and
                diagonal passes on bounded samples, the QFT-30 step estimated
                from them (value), with nproc / lscpu

roofline.traffic is the HBM bytes per launch of the dominant kernel as the
two counter runs summarised in profiles/pmc_traffic_qft30_c128.txt measured
them (FETCH_SIZE and WRITE_SIZE scaled by the in-place scale kernel).

--section NAME runs one extra alone (after the tile-geometry autotuning the
full run does first), for profiling it.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0
METRIC = "30q gates/s + HBM TB/s vs peak; 28q adjoint grads/s (c64 vs c128 plain)"
NQ = 30


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def family_stats(sim, families=("diag", "u3_tile", "fused", "u3_pair", "reset", "cp", "perm")):
    out = {}
    for f in families:
        t = sim.timing(f)
        if t["launches"]:
            out[f] = t
    return out


TRAFFIC_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "pmc_traffic_qft30_c128.txt")


def measured_traffic(family):
    """HBM bytes per launch of a kernel family from the committed counter
    summary (line "traffic <family> <bytes>"), or None."""
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 3 and parts[0] == "traffic" and parts[1] == family:
                    return int(parts[2])
    except OSError:
        return None
    return None


def roofline_of(stats):
    fam = max(stats, key=lambda k: stats[k]["ms"])
    t = stats[fam]
    per_launch_bytes = t["bytes"] / t["launches"]
    avg_ms = t["ms"] / t["launches"]
    achieved = per_launch_bytes / (avg_ms * 1e-3) / 1e9
    return fam, {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                 "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": measured_traffic(fam),
                 "kernel": fam, "bytes_per_launch": int(per_launch_bytes), "avg_launch_ms": round(avg_ms, 4),
                 "launches": t["launches"]}


def qft_step(Z, W, prec, steps, warmup, policy="fused", sim=None):
    gates = W.qft(NQ)
    if sim is None:
        sim = Z.ZXH(NQ, False, True, precision=prec, policy=policy)
    sim.clear_gates()
    sim.gates(gates)
    for i in range(warmup):
        sim.execute()
    from zxh_adae_amd import kernels as K
    K.sync()
    sim.reset_timing()
    sim.set_timing(True)
    t0 = time.perf_counter()
    for i in range(steps):
        sim.execute()
    K.sync()
    dt = time.perf_counter() - t0
    sim.set_timing(False)
    stats = family_stats(sim)
    return sim, len(gates), dt / steps, stats


def wrap_state(sim_state, prec):
    from zxh_adae_amd.kernels import DeviceState
    ptr, m = sim_state
    return DeviceState.wrap(ptr, m, prec)


def diag_batches(Z, W, K, sim_state, prec, repeats=5):
    from zxh_adae_amd.kernels import NO_SEL
    st = wrap_state(sim_state, prec)
    m = st.nbits
    out = {}
    for B in (1, 4, 16, 64, 256):
        gs = W.diag_batch(m, B, 100 + B)
        s0 = [(1 << g[2]) for g in gs]
        s1 = [NO_SEL if g[0] == W.P else (1 << g[1]) for g in gs]
        th = [g[3] for g in gs]
        st.diag(s0, s1, th)  # warmup
        K.reset_timing()
        K.set_timing(True)
        for _ in range(repeats):
            st.diag(s0, s1, th)
        t = K.timing("diag")
        K.set_timing(False)
        avg_ms = t["ms"] / t["launches"]
        gbps = t["bytes"] / t["launches"] / (avg_ms * 1e-3) / 1e9
        out[str(B)] = {"ms": round(avg_ms, 4), "GB/s": round(gbps, 1), "frac": round(gbps / HBM_PEAK_GBPS, 4),
                       "gates/s": round(B / (avg_ms * 1e-3), 1)}
    return out


def u3_all(Z, W, K, sim_state, prec, repeats=3):
    st = wrap_state(sim_state, prec)
    m = st.nbits
    descs = [((1 << g[2]), (1 << g[2]), W.u3_matrix(g[3], g[4], g[5])) for g in W.u3_all(m, 77)]
    st.u3_batch(descs)
    K.reset_timing()
    K.set_timing(True)
    t0 = time.perf_counter()
    for _ in range(repeats):
        st.u3_batch(descs)
    K.sync()
    wall = (time.perf_counter() - t0) / repeats
    t = K.timing("u3_tile")
    K.set_timing(False)
    passes = t["launches"] / repeats
    kern_ms = t["ms"] / repeats
    gbps = t["bytes"] / t["launches"] / (t["ms"] / t["launches"] * 1e-3) / 1e9
    return {"gates": m, "passes": passes, "ms": round(kern_ms, 3), "wall_ms": round(wall * 1e3, 3),
            "gates/s": round(m / (kern_ms * 1e-3), 1), "GB/s_per_pass": round(gbps, 1),
            "frac": round(gbps / HBM_PEAK_GBPS, 4)}


def adjoint_c3(W, repeats=3, check_fp64=True):
    """Two-local ansatz, 28 qubits, 8 layers (448 parameters), 200 Z/ZZ
    terms: complex64 forward and adjoint backward, with the FP64 (complex128)
    run of the same circuit for the gradient error and the state infidelity."""
    import numpy as np
    from zxh_adae_amd.ansatz import Ansatz, two_local
    n, layers = 28, 8
    ops, npar = two_local(n, layers)
    params = np.array(W.ansatz_params(npar, 3))
    terms = W.zz_observable(n, 200, 3)
    a = Ansatz(n, ops, npar, "c64")
    a.set_observable(terms)
    a.forward(params)
    a.backward(params)  # warm
    tf = tb = 0.0
    for _ in range(repeats):
        t0 = time.perf_counter()
        e, _ = a.forward(params)
        t1 = time.perf_counter()
        g32, _ = a.backward(params)  # reuses the forward state
        t2 = time.perf_counter()
        tf += t1 - t0
        tb += t2 - t1
    tf /= repeats
    tb /= repeats
    a.reset_timing()
    a.set_timing(True)
    a.forward(params)
    a.backward(params)
    a.set_timing(False)
    fams = {}
    for fam in ("fused", "u3_tile", "diag", "obs", "adjoint", "reset"):
        t = a.timing(fam)
        if t["launches"]:
            avg = t["ms"] / t["launches"]
            fams[fam] = {"launches": t["launches"], "avg_ms": round(avg, 4),
                         "GB/s": round(t["bytes"] / t["launches"] / (avg * 1e-3) / 1e9, 1)}
    out = {"qubits": n, "layers": layers, "params": npar, "terms": len(terms), "state": "complex64",
           "forward_ms": round(tf * 1e3, 3), "backward_ms": round(tb * 1e3, 3),
           "backward_over_forward": round(tb / tf, 3), "grads/s": round(npar / (tf + tb), 1),
           "energy": e, "tile_passes_per_forward": a.tile_passes(), "kernels": fams}
    if check_fp64:
        a.forward(params)
        psi32 = a.state()
        del a
        b = Ansatz(n, ops, npar, "c128")
        b.set_observable(terms)
        e64, _ = b.forward(params)
        psi64 = b.state()
        g64, _ = b.backward(params)
        ov = abs(np.vdot(psi64, psi32.astype(np.complex128)))
        out["fp64"] = {"max_grad_error": float(np.max(np.abs(g32 - g64))),
                       "max_abs_grad": float(np.max(np.abs(g64))), "energy_error": abs(e - e64),
                       "infidelity": float(1.0 - ov),
                       "max_amp_error": float(np.max(np.abs(psi32.astype(np.complex128) - psi64)))}
        del psi32, psi64, b
    return out


def batch_c4(W, repeats=3, sequential=64):
    """512 parameter sets of the two-local ansatz, widths 16..24 and depths
    2..10, 50-term observables, complex64: one batched engine (forward pass p
    of every set in one launch, replayed from a hipGraph).  Timed forward-only
    and forward+adjoint runs; the single-circuit path on every k-th set gives
    the sequential estimate and the per-set gradient check."""
    import numpy as np
    from zxh_adae_amd.ansatz import Ansatz, AnsatzBatch, two_local
    from zxh_adae_amd import autotune
    sets = W.c4_sets()
    full = []
    for n, layers, terms in sets:
        ops, npar = two_local(n, layers)
        full.append((n, ops, npar, terms))
    nps = [s[2] for s in full]
    offs = np.concatenate([[0], np.cumsum(nps)])
    params = np.array(W.ansatz_params(int(offs[-1]), 4))
    # the fewest-passes rule for every width, timed as a reference
    b = AnsatzBatch("c64")
    for n, ops, npar, terms in full:
        b.add(n, ops, npar, terms)
    b.finalize()
    b.run(params, gradients=True)
    t0 = time.perf_counter()
    for _ in range(repeats):
        b.run(params, gradients=True)
    t_rule = (time.perf_counter() - t0) / repeats
    b.close()
    # geometry per width class from the timing autotuner
    t0 = time.perf_counter()
    tuned = autotune.tune_batch(full, "c64")
    t_tune = time.perf_counter() - t0
    b = AnsatzBatch("c64")
    for n, seg in tuned["choice"].items():
        b.set_geometry(int(n), int(seg))
    for n, ops, npar, terms in full:
        b.add(n, ops, npar, terms)
    t0 = time.perf_counter()
    b.finalize()
    t_setup = time.perf_counter() - t0
    b.run(params, gradients=True)
    tf = tb = 0.0
    for _ in range(repeats):
        t0 = time.perf_counter()
        b.run(params, gradients=False)
        t1 = time.perf_counter()
        e, norms, grads = b.run(params, gradients=True)
        t2 = time.perf_counter()
        tf += t1 - t0
        tb += t2 - t1
    tf /= repeats
    tb /= repeats
    info = b.info()
    # algorithmic bytes (S = state bytes): forward = reset write S + passes x 2S
    # + observable read S; backward adds lam = H psi (2S), the last window on
    # both states (4S) and passes x 4S
    fwd_bytes = bwd_bytes = 0
    for i in range(len(sets)):
        nb, ps = b.set_info(i)
        S = 8 << nb
        fwd_bytes += S + ps * 2 * S + S
        bwd_bytes += S + ps * 2 * S + 2 * S + 4 * S + ps * 4 * S
    del b
    step = max(1, len(sets) // sequential)
    t_seq = 0.0
    nseq = 0
    gdiff = ediff = 0.0
    for i in range(0, len(sets), step):
        n, layers, terms = sets[i]
        ops, npar = two_local(n, layers)
        a = Ansatz(n, ops, npar, "c64")
        a.set_observable(terms)
        p = params[offs[i]:offs[i + 1]]
        a.forward(p)
        a.backward(p)
        t0 = time.perf_counter()
        e1, _ = a.forward(p)
        g1, _ = a.backward(p)
        t_seq += time.perf_counter() - t0
        nseq += 1
        gdiff = max(gdiff, float(np.max(np.abs(g1 - grads[i]))))
        ediff = max(ediff, abs(e1 - e[i]))
        a.close()
    seq_est = t_seq * len(sets) / nseq
    return {"sets": len(sets), "qubits": "16..24", "layers": "2..10", "terms": 50, "state": "complex64",
            "params": int(offs[-1]), "launches_per_run": info["launches"], "setup_s": round(t_setup, 3),
            "forward_ms": round(tf * 1e3, 3), "forward_backward_ms": round(tb * 1e3, 3),
            "sets/s": round(len(sets) / tb, 1), "grads/s": round(int(offs[-1]) / tb, 1),
            "forward_GB/s": round(fwd_bytes / tf / 1e9, 1), "forward_backward_GB/s": round(bwd_bytes / tb / 1e9, 1),
            "rule_forward_backward_ms": round(t_rule * 1e3, 3),
            "autotune": {**tuned, "seconds": round(t_tune, 2)},
            "sequential_single_circuit_ms": round(seq_est * 1e3, 3), "sequential_sample": nseq,
            "speedup_vs_sequential": round(seq_est / tb, 2),
            "max_grad_diff_vs_single": gdiff, "max_energy_diff_vs_single": ediff}


ZXH_WORKLOADS = ["qft_n30", "qaoa_n30", "vqe_two_local_n30", "vqe_su2_n30", "kernel_crz_dom_n30", "kernel_rz_dom_n30",
                 "kernel_h_inter_dom_n30", "kernel_h_intra_dom_n30", "bv_n30", "ghz_n32", "multiplier_n45"]


def zxh_workloads(Z, repeats=3, breakdown=False, names=None):
    """The reference's own benchmark circuits (its QASM workloads, kept as
    test fixtures) through the drop-in ZXH object, complex128: execute() time
    in the default lazy ZXH mode (X/CX as affine updates, qubits expanded on
    demand) and in the plain method (eager_expand_all), with the gate count
    after lowering and the physical width each mode reached."""
    from zxh_adae_amd import frontend as F
    from zxh_adae_amd import kernels as K
    root = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden", "workloads")
    out = {}
    for name in names or ZXH_WORKLOADS:
        circ = F.load_qasm(os.path.join(root, name + ".qasm"))
        gs = F.to_gate_tuples(F.lower_circuit(circ))
        row = {"qubits": circ.num_qubits, "gates": len(gs)}
        for mode, eager in (("lazy", False), ("plain", True)):
            if eager and circ.num_qubits > 32:
                continue
            s = Z.ZXH(circ.num_qubits, False, eager)
            s.gates(gs)
            s.execute()
            t0 = time.perf_counter()
            for _ in range(repeats):
                s.execute()
            t = (time.perf_counter() - t0) / repeats
            row[mode] = {"execute_ms": round(t * 1e3, 3), "gates/s": round(len(gs) / t, 1),
                         "width": int(s.device_state()[1])}
            if breakdown:
                # one more run with HIP-event timing per kernel family
                s.reset_timing()
                s.set_timing(True)
                s.execute()
                K.sync()
                s.set_timing(False)
                fams = {}
                for fam in K.FAMILIES:
                    tm = s.timing(fam)
                    if tm["launches"]:
                        fams[fam] = {"launches": tm["launches"], "ms": round(tm["ms"], 3)}
                row[mode]["families"] = fams
            del s
        out[name] = row
    return out


def c5_circuit(W, K, prec, n, layers=40, shots=1000000, seed=5):
    """The large random circuit: per layer floor(n/5) Haar 5-qubit blocks
    (matrix cores), a 64-gate diagonal batch, 8 multi-controlled X gates;
    then 1e6 shots.  Kernel times by HIP events; MFMA rate of the blocks."""
    from zxh_adae_amd.kernels import DeviceState, apply_layer, mix_seed, set_basis_zero
    spec = W.c5_layers(n, layers, seed)
    st = DeviceState(n, prec)
    set_basis_zero(st)
    K.sync()
    K.reset_timing()
    K.set_timing(True)
    t0 = time.perf_counter()
    for L in spec:
        apply_layer(st, L)
    K.sync()
    t_circ = time.perf_counter() - t0
    fams = {}
    for fam in ("block", "diag", "perm"):
        t = K.timing(fam)
        if t["launches"]:
            avg = t["ms"] / t["launches"]
            fams[fam] = {"launches": t["launches"], "avg_ms": round(avg, 4),
                         "GB/s": round(t["bytes"] / t["launches"] / (avg * 1e-3) / 1e9, 1)}
    blk = K.timing("block")
    flops = 256.0 * (1 << n)  # 32 complex MACs = 8 * 2^5 real flops per amplitude
    tflops = flops / (blk["ms"] / blk["launches"] * 1e-3) / 1e12
    peak = 157.3 if prec == "c64" else 78.6
    t1 = time.perf_counter()
    st.sample(18, mix_seed(seed, 0), shots)
    t_samp = time.perf_counter() - t1
    tree = K.timing("tree")
    K.set_timing(False)
    st.free()
    tree_ms = tree["ms"] / max(1, tree["launches"])
    tree_gbps = tree["bytes"] / max(1, tree["launches"]) / (tree_ms * 1e-3) / 1e9 if tree["launches"] else 0.0
    nblocks = sum(len(L["blocks"]) for L in spec)
    ngates = nblocks + sum(len(L["diag"]) + len(L["mcx"]) for L in spec)
    return {"qubits": n, "state": "complex64" if prec == "c64" else "complex128", "layers": layers,
            "gates": ngates, "blocks": nblocks, "circuit_s": round(t_circ, 3), "gates/s": round(ngates / t_circ, 1),
            "kernels": fams, "block_tflops": round(tflops, 2), "block_mfma_frac": round(tflops / peak, 4),
            "mfma_peak_tflops": peak,
            "sampling": {"shots": shots, "seconds": round(t_samp, 3), "tree_ms": round(tree_ms, 3),
                         "tree_GB/s": round(tree_gbps, 1), "tree_frac": round(tree_gbps / HBM_PEAK_GBPS, 4)}}


def sampling(sim, shots=1000000, repeats=3):
    """measure(shots) on the simulator's state: wall time of the whole call
    (tree build, host thresholds overlapped with it, search, results back),
    the tree pass's and the search's kernel times, and the tree pass against
    the HBM roofline (bytes: one read of the state plus the stored levels)."""
    sim.set_seed(5)
    sim.measure(shots)  # warm: pinned buffers, scratch
    sim.reset_timing()
    sim.set_timing(True)
    t0 = time.perf_counter()
    for _ in range(repeats):
        sim.measure(shots)
    dt = (time.perf_counter() - t0) / repeats
    sim.set_timing(False)
    tree, search = sim.timing("tree"), sim.timing("sample")
    tree_ms = tree["ms"] / tree["launches"]
    gbps = tree["bytes"] / tree["launches"] / (tree_ms * 1e-3) / 1e9
    return {"shots": shots, "qubits": sim.device_state()[1], "measure_ms": round(dt * 1e3, 3),
            "shots/s": round(shots / dt, 1), "tree_ms": round(tree_ms, 4),
            "search_ms": round(search["ms"] / search["launches"], 4),
            "roofline": {"bound": "hbm", "achieved": round(gbps, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                         "frac": round(gbps / HBM_PEAK_GBPS, 4), "kernel": "tree",
                         "bytes_per_launch": int(tree["bytes"] / tree["launches"])}}


def sampling_section(Z, W, K):
    """measure(1e6) on the 30-qubit uniform superposition (c128, c64), and the
    kernel-level sampler on a 33-qubit complex64 state (64 GiB)."""
    from zxh_adae_amd.kernels import DeviceState, mix_seed, set_basis_zero
    out = {}
    for prec in ("c128", "c64"):
        sim = Z.ZXH(NQ, False, True, precision=prec)
        sim.gates(W.hadamard_all(NQ))
        sim.execute()
        out[prec] = sampling(sim)
        del sim
    n = 33
    st = DeviceState(n, "c64")
    set_basis_zero(st)
    st.u3_batch([((1 << q), (1 << q), W.u3_matrix(0.4 + 0.05 * q, 0.3, 0.2)) for q in range(n)])
    st.sample(18, mix_seed(5, 0), 1000000)
    K.reset_timing()
    K.set_timing(True)
    t0 = time.perf_counter()
    for _ in range(3):
        st.sample(18, mix_seed(5, 0), 1000000)
    dt = (time.perf_counter() - t0) / 3
    tree, search = K.timing("tree"), K.timing("sample")
    K.set_timing(False)
    st.scale(1.0, 0.0)  # counter calibration: one read and one write of 64 GiB
    K.sync()
    st.free()
    tree_ms = tree["ms"] / tree["launches"]
    gbps = tree["bytes"] / tree["launches"] / (tree_ms * 1e-3) / 1e9
    out["c64_33q_kernel"] = {"shots": 1000000, "seconds": round(dt, 4), "tree_ms": round(tree_ms, 3),
                             "search_ms": round(search["ms"] / search["launches"], 3),
                             "tree_GB/s": round(gbps, 1), "tree_frac": round(gbps / HBM_PEAK_GBPS, 4),
                             "read_only_GB/s": round((8 << n) / (tree_ms * 1e-3) / 1e9, 1)}
    return out


def accuracy(sim64, sim128):
    """Infidelity 1 - |<psi64|psi128>| and max amplitude error of a complex64
    simulator's physical state against a complex128 one that ran the same
    operations (both plain method, same width)."""
    from zxh_adae_amd.kernels import DeviceState
    p64, m64 = sim64.device_state()
    p128, m128 = sim128.device_state()
    assert m64 == m128
    c = DeviceState.wrap(p64, m64, "c64").compare(DeviceState.wrap(p128, m128, "c128"))
    return {"infidelity": c["infidelity"], "max_amp_error": c["max_amp_error"],
            "norm2_c64": c["norm2_self"]}


def c5_accuracy(W, K, n, layers=40, seed=5):
    """The large random circuit's layers (blocks, diagonal batch, MCX cascade)
    at a width where both precisions fit together: the complex64 state
    against the complex128 one after all layers."""
    from zxh_adae_amd.kernels import DeviceState, apply_layer, set_basis_zero
    spec = W.c5_layers(n, layers, seed)
    a, b = DeviceState(n, "c64"), DeviceState(n, "c128")
    for st in (a, b):
        set_basis_zero(st)
        for L in spec:
            apply_layer(st, L)
    c = a.compare(b)
    a.free()
    b.free()
    return {"qubits": n, "layers": layers, "infidelity": c["infidelity"], "max_amp_error": c["max_amp_error"],
            "norm2_c64": c["norm2_self"], "norm2_c128": c["norm2_other"]}


def host_info():
    """Core counts of this host: nproc, the affinity set, and lscpu's sockets x
    cores per socket x threads per core."""
    import subprocess
    info = {"nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0))}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        keys = {"Socket(s)": "sockets", "Core(s) per socket": "cores_per_socket",
                "Thread(s) per core": "threads_per_core", "Model name": "model", "CPU(s)": "cpus"}
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in keys:
                info[keys[k.strip()]] = v.strip()
    except (OSError, subprocess.SubprocessError):
        info["lscpu"] = "unavailable"
    return info


def _cpu_run(threads, extra, timeout_s):
    import subprocess
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", OMP_PLACES="cores")
    cmd = [sys.executable, os.path.join(ROOT, "oracle", "cpu_baseline.py"), *extra]
    r = subprocess.run(cmd, env=env, stdout=subprocess.PIPE, text=True, timeout=timeout_s)
    if r.returncode != 0:
        raise RuntimeError(f"cpu baseline child failed ({r.returncode})")
    return json.loads(r.stdout.strip().splitlines()[-1])


def cpu_baseline(c1_gpu):
    """The oracle's restatement of the reference's CPU path (oracle/
    liboracle_base.so, OpenMP), in child processes on this host: once on the
    cores this process may use and once on one thread.  value = the QFT-30
    step as the reference schedules it, estimated from the measured per-pass
    times at full threads (gates/s)."""
    host = host_info()
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or host["affinity"]
    log(f"cpu baseline: {threads} threads")
    full = _cpu_run(threads, ["--c1-repeats", "5", "--pass-repeats", "5"], 400)
    log("cpu baseline: 1 thread")
    one = _cpu_run(1, ["--c1-repeats", "5", "--pass-repeats", "3", "--slice-shift", "3", "--c1-budget-s", "25"], 400)
    q = full["qft30"]
    out = {"value": q["gates/s"], "unit": "gates/s", "cores": threads, "kind": "port",
           "sample": (f"QFT-30 complex128 step as the reference runs it (reset + 30 U3 sweeps + 29 diagonal "
                      f"windows), {q['estimated_ms'] / 1e3:.1f} s estimated from per-pass times measured on "
                      f"bounded samples of the 2^30 state (2^28 amplitudes for U3, B x 2^k = 2^28 selector "
                      f"evaluations per window; 1 warm-up + 5 runs each); C1 (20q, 1180 gates, 1024 shots) "
                      f"timed in full"),
           "build": "oracle/liboracle_base.so (C restatement of the reference's OpenMP CPU backend, "
                    "gcc -O3 -fopenmp -ffp-contract=off, selector calls out of line)",
           "host": host, "threads": {"full": full, "one": one}}
    if c1_gpu:
        g = c1_gpu
        out["c1_gpu_over_cpu"] = {
            "execute": round(full["c1"]["execute"]["mean_ms"] / g["execute"]["mean_ms"], 1),
            "sampling": round(full["c1"]["sampling"]["mean_ms"] / g["sampling"]["mean_ms"], 1)}
    return out


def c1_gpu(Z, W, repeats=5):
    """C1 on the GPU through the drop-in: 20 qubits, plain method, complex128,
    1180 gates, then Sampling(1024) with set_seed(31415926); execute() and
    Sampling() timed separately (1 warm-up, then repeats runs)."""
    import statistics
    gates = W.random_layers(20, 20, 1)
    sim = Z.ZXH(20, False, True)
    sim.gates(gates)
    sim.set_seed(31415926)
    sim.execute()
    sim.Sampling(1024)
    te, ts = [], []
    for _ in range(repeats):
        t0 = time.perf_counter()
        sim.execute()
        t1 = time.perf_counter()
        sim.measure(1024)
        sim.results_bits()
        t2 = time.perf_counter()
        te.append(t1 - t0)
        ts.append(t2 - t1)
    del sim

    def st(xs):
        return {"mean_ms": round(1e3 * statistics.fmean(xs), 4), "sd_ms": round(1e3 * statistics.stdev(xs), 4),
                "runs": len(xs)}
    ex = st(te)
    return {"qubits": 20, "gates": len(gates), "shots": 1024, "execute": ex, "sampling": st(ts),
            "gates/s": round(len(gates) / (ex["mean_ms"] * 1e-3), 1)}


def kernels_only(Z, K, W, repeats, prec):
    """Each hot kernel on a 30-qubit state, repeats launches each, plus the
    copy / in-place-scale calibration kernels (for profilers and counters)."""
    import numpy as np
    from zxh_adae_amd.kernels import DeviceState, NO_SEL
    st = DeviceState(NQ, prec)
    other = DeviceState(NQ, prec)
    st.scale(0.0)  # defined contents (zeros) without a host upload
    other.scale(0.0)
    gs = W.diag_batch(NQ, 64, 164)
    s0 = [(1 << g[2]) for g in gs]
    s1 = [NO_SEL if g[0] == W.P else (1 << g[1]) for g in gs]
    th = [g[3] for g in gs]
    one = [((1 << 29), (1 << 29), W.u3_matrix(0.3, 0.2, 0.1))]
    allq = [((1 << g[2]), (1 << g[2]), W.u3_matrix(g[3], g[4], g[5])) for g in W.u3_all(NQ, 77)]
    # a fused pass as the QFT's: 12 U3 targets (the lowest free bits and the
    # highest the geometry allows) plus a 32-gate P/CP window, autotuned geometry
    from zxh_adae_amd import autotune
    tuned = autotune.tune_tile(st)
    seg = tuned["best"]
    low = seg + (1 if prec == "c64" else 0)
    ftargets = list(range(low)) + list(range(NQ - (12 - seg), NQ))
    fdescs = [((1 << q), (1 << q), W.u3_matrix(0.3 + 0.01 * q, 0.2, 0.1)) for q in ftargets]
    wg = W.diag_batch(NQ, 32, 4343)
    w0 = [(1 << g[2]) for g in wg]
    w1 = [NO_SEL if g[0] == W.P else (1 << g[1]) for g in wg]
    wt = [g[3] for g in wg]
    K.reset_timing()
    K.set_timing(True)
    out = {"tile_seg": seg}
    for name, fn in (("copy", lambda: other.copy_from(st)), ("scale", lambda: st.scale(1.0, 0.0)),
                     ("diag", lambda: st.diag(s0, s1, th)), ("u3_one", lambda: st.u3_batch(one)),
                     ("u3_all", lambda: st.u3_batch(allq)), ("fused", lambda: st.fused_pass(fdescs, w0, w1, wt))):
        log(f"kernel {name} x{repeats} ({prec})")
        for _ in range(repeats):
            fn()
        K.sync()
    for fam in ("copy", "scale", "diag", "u3_tile", "fused"):
        t = K.timing(fam)
        if t["launches"]:
            avg = t["ms"] / t["launches"]
            out[fam] = {"launches": t["launches"], "avg_ms": round(avg, 4),
                        "GB/s": round(t["bytes"] / t["launches"] / (avg * 1e-3) / 1e9, 1)}
    K.set_timing(False)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--kernels", choices=["c128", "c64"], default=None,
                    help="only launch each hot kernel --steps times on a 30-qubit state (profiling)")
    ap.add_argument("--section", choices=["adjoint_c3", "batch_c4", "c5_c64", "c5_c128", "zxh_workloads", "sampling", "c1"],
                    default=None,
                    help="run one extra alone, after the tile-geometry autotuning (profiling)")
    ap.add_argument("--workloads", default=None,
                    help="with --section zxh_workloads: comma-separated circuit names (default: all)")
    args = ap.parse_args()
    unknown = [w for w in (args.workloads.split(",") if args.workloads else []) if w not in ZXH_WORKLOADS]
    if unknown:
        print(f"unknown workloads {unknown}; known: {ZXH_WORKLOADS}", file=sys.stderr)
        sys.exit(2)
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the engine is single-GPU by design",
              file=sys.stderr)
        sys.exit(2)

    import zxh_adae_amd as Z
    from zxh_adae_amd import kernels as K
    from zxh_adae_amd import workloads as W

    Z.init()
    if args.kernels:
        print(json.dumps({"kernels": kernels_only(Z, K, W, args.steps, args.kernels), "precision": args.kernels,
                          "qubits": NQ}), flush=True)
        return
    from zxh_adae_amd import autotune
    if args.section:
        prec = "c128" if args.section in ("c5_c128", "zxh_workloads", "sampling", "c1") else "c64"
        st = K.DeviceState(NQ, prec)
        st.scale(0.0)
        tuned = autotune.tune_tile(st)
        st.free()
        log(f"section {args.section} (tile geometry {tuned['best']})")
        if args.section == "adjoint_c3":
            out = adjoint_c3(W)
        elif args.section == "batch_c4":
            out = batch_c4(W)
        elif args.section == "c5_c64":
            out = c5_circuit(W, K, "c64", 33)
        elif args.section == "sampling":
            out = sampling_section(Z, W, K)
        elif args.section == "zxh_workloads":
            out = zxh_workloads(Z, breakdown=True, names=args.workloads.split(",") if args.workloads else None)
        elif args.section == "c1":
            out = c1_gpu(Z, W, repeats=20)
        else:
            out = c5_circuit(W, K, "c128", 32)
        print(json.dumps({args.section: out, "autotune": tuned}), flush=True)
        return

    # the resident 30-qubit state; the autotuner picks the diagonal-sweep variant on it
    sim = Z.ZXH(NQ, False, True, precision="c128", policy="fused")
    sim.gates(W.hadamard_all(NQ))
    sim.execute()
    log("autotune diag and tile c128")
    tuned = {"c128": autotune.tune_diag(wrap_state(sim.device_state(), "c128")),
             "tile_c128": autotune.tune_tile(wrap_state(sim.device_state(), "c128"))}

    log(f"QFT-{NQ} c128 fused: warmup {args.warmup}, steps {args.steps}")
    sim, ngates, t_step, stats = qft_step(Z, W, "c128", args.steps, args.warmup, sim=sim)
    fam, roof = roofline_of(stats)
    log(f"step {t_step * 1e3:.2f} ms, {ngates / t_step:.1f} gates/s; dominant {fam}: "
        f"{roof['achieved']:.0f} GB/s ({roof['frac']:.1%})")
    result = {
        "metric": METRIC, "value": round(ngates / t_step, 3), "unit": "gates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(t_step * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"qft{NQ}_c128_execute", "qubits": NQ, "gates": ngates, "state": "complex128",
                   "method": "plain (eager_expand_all)", "policy": "fused"},
        "roofline": roof,
        "kernels": {k: {"launches": v["launches"], "avg_ms": round(v["ms"] / v["launches"], 4),
                        "GB/s": round(v["bytes"] / v["launches"] / (v["ms"] / v["launches"] * 1e-3) / 1e9, 1)}
                    for k, v in stats.items()},
        "autotune": tuned,
    }
    log(f"QFT-{NQ} c128 reference policy")
    simr = Z.ZXH(NQ, False, True, precision="c128", policy="reference")
    simr, _, t_ref, stats_ref = qft_step(Z, W, "c128", max(1, args.steps // 2), 1, sim=simr)
    result["reference_policy"] = {"gates/s": round(ngates / t_ref, 3), "ms_per_step": round(t_ref * 1e3, 3),
                                  "launches": {k: v["launches"] for k, v in stats_ref.items()}}
    del simr

    if not args.no_extras:
        # complex64: the same QFT step, then each c64 result compared with the
        # complex128 state that went through the same operations (infidelity
        # 1 - |<psi64|psi128>| and max |psi64_i - psi128_i|, one device pass)
        log(f"QFT-{NQ} c64")
        sim64 = Z.ZXH(NQ, False, True, precision="c64", policy="fused")
        sim64.gates(W.hadamard_all(NQ))
        sim64.execute()
        tuned["c64"] = autotune.tune_diag(wrap_state(sim64.device_state(), "c64"))
        tuned["tile_c64"] = autotune.tune_tile(wrap_state(sim64.device_state(), "c64"))
        sim64, _, t64, stats64 = qft_step(Z, W, "c64", args.steps, 1, sim=sim64)
        result["c64"] = {"gates/s": round(ngates / t64, 3), "ms_per_step": round(t64 * 1e3, 3),
                         **accuracy(sim64, sim)}
        # uniform superposition input for the diagonal batches and the U3 pass
        for s_ in (sim, sim64):
            s_.clear_gates()
            s_.gates(W.hadamard_all(NQ))
            s_.execute()
        log("diag batches c128, c64")
        result["diag_batches"] = {"c128": diag_batches(Z, W, K, sim.device_state(), "c128"),
                                  "c64": diag_batches(Z, W, K, sim64.device_state(), "c64")}
        result["diag_batches"]["c64"].update(accuracy(sim64, sim))
        log("u3 all c128, c64")
        result["u3_all"] = {"c128": u3_all(Z, W, K, sim.device_state(), "c128"),
                            "c64": u3_all(Z, W, K, sim64.device_state(), "c64")}
        result["u3_all"]["c64"].update(accuracy(sim64, sim))
        log("sampling 1e6 shots c64")
        result["sampling_c64"] = sampling(sim64)
        del sim64
        log("sampling 1e6 shots")
        result["sampling"] = sampling(sim)
        del sim

    if not args.no_extras:
        log("adjoint 28q c64 (two-local, 8 layers, 200 terms)")
        result["adjoint_c3"] = adjoint_c3(W)
        log("batched ansatz: 512 sets, 16..24 qubits, 2..10 layers, c64")
        result["batch_c4"] = batch_c4(W)
        log("the reference's benchmark circuits, lazy ZXH and plain")
        result["zxh_workloads"] = zxh_workloads(Z)
        log("large random circuit 33q c64, 40 layers, 1e6 shots")
        result["c5"] = {"c64": c5_circuit(W, K, "c64", 33)}
        log("large random circuit 32q c128, 40 layers, 1e6 shots")
        result["c5"]["c128"] = c5_circuit(W, K, "c128", 32)
        log("large random circuit at 31q in both precisions (c64 accuracy)")
        acc = c5_accuracy(W, K, 31)
        result["c5"]["c64"].update({"infidelity": acc["infidelity"], "max_amp_error": acc["max_amp_error"],
                                    "accuracy_at_qubits": 31, "accuracy_31q": acc})

    log("C1: 20q plain c128, 1180 gates, Sampling(1024)")
    result["c1"] = c1_gpu(Z, W)
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(result["c1"])
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
