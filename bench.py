"""Benchmark of the MI355X state-vector engine (one JSON line on stdout).

Step: one execute() of the 30-qubit QFT-structured circuit (H on every
target followed by CP(pi/2^k) to all lower qubits, no swaps: 30 H + 435 CP)
in complex128 through the engine's drop-in ZXH object with the default fused
policy: the reference's own U3 clusters and diagonal windows, merged into
tile passes where they commute.  The state is resident in HBM; execute()
includes the reset of the 16 GiB state, as the reference's execute() does.
value = gates / second of the timed steps.

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
  cpu_baseline  the reference's CPU path (oracle/_ref when built, else the
                oracle port) on a bounded prefix of the same circuit, timed on
                this host
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


def roofline_of(stats):
    fam = max(stats, key=lambda k: stats[k]["ms"])
    t = stats[fam]
    per_launch_bytes = t["bytes"] / t["launches"]
    avg_ms = t["ms"] / t["launches"]
    achieved = per_launch_bytes / (avg_ms * 1e-3) / 1e9
    return fam, {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                 "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": None,
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


def cpu_baseline(W, targets=4):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py as O
    gates = []
    for t in range(targets):
        gates.append((W.H, 0, t, 0.0, 0.0, 0.0))
        for j in range(t):
            gates.append((W.CP, j, t, 3.141592653589793 / (2 ** (t - j)), 0.0, 0.0))
    kind = "reference" if O.ref_available() else "port"
    cls = O.RefSim if kind == "reference" else O.OracleSim
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    sim = cls(NQ, False, True)
    sim.gates(gates)
    t0 = time.perf_counter()
    sim.execute()
    dt = time.perf_counter() - t0
    del sim
    return {"value": round(len(gates) / dt, 3), "unit": "gates/s", "cores": cores, "kind": kind,
            "sample": f"first {targets} targets of the 30-qubit QFT ({len(gates)} gates, complex128, "
                      f"one execute() incl. its 16 GiB state reset), {dt:.1f} s"}


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
    K.reset_timing()
    K.set_timing(True)
    out = {}
    for name, fn in (("copy", lambda: other.copy_from(st)), ("scale", lambda: st.scale(1.0, 0.0)),
                     ("diag", lambda: st.diag(s0, s1, th)), ("u3_one", lambda: st.u3_batch(one)),
                     ("u3_all", lambda: st.u3_batch(allq))):
        log(f"kernel {name} x{repeats} ({prec})")
        for _ in range(repeats):
            fn()
        K.sync()
    for fam in ("copy", "scale", "diag", "u3_tile"):
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
    args = ap.parse_args()
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
        # uniform superposition input for the diagonal batches and the U3 pass
        sim.clear_gates()
        sim.gates(W.hadamard_all(NQ))
        sim.execute()
        log("diag batches c128")
        result["diag_batches"] = {"c128": diag_batches(Z, W, K, sim.device_state(), "c128")}
        log("u3 all c128")
        result["u3_all"] = {"c128": u3_all(Z, W, K, sim.device_state(), "c128")}
        log("sampling 1e6 shots")
        sim.set_seed(5)
        sim.set_timing(True)
        t0 = time.perf_counter()
        sim.measure(1000000)
        result["sampling"] = {"shots": 1000000, "qubits": NQ, "seconds": round(time.perf_counter() - t0, 4),
                              "search_kernel_ms": round(sim.timing("sample")["ms"], 3)}
        sim.set_timing(False)
        del sim
        log(f"QFT-{NQ} c64")
        sim64 = Z.ZXH(NQ, False, True, precision="c64", policy="fused")
        sim64.gates(W.hadamard_all(NQ))
        sim64.execute()
        tuned["c64"] = autotune.tune_diag(wrap_state(sim64.device_state(), "c64"))
        tuned["tile_c64"] = autotune.tune_tile(wrap_state(sim64.device_state(), "c64"))
        sim64, _, t64, stats64 = qft_step(Z, W, "c64", args.steps, 1, sim=sim64)
        result["c64"] = {"gates/s": round(ngates / t64, 3), "ms_per_step": round(t64 * 1e3, 3)}
        sim64.clear_gates()
        sim64.gates(W.hadamard_all(NQ))
        sim64.execute()
        result["diag_batches"]["c64"] = diag_batches(Z, W, K, sim64.device_state(), "c64")
        result["u3_all"]["c64"] = u3_all(Z, W, K, sim64.device_state(), "c64")
        del sim64

    if not args.no_cpu_baseline:
        log("cpu baseline")
        result["cpu_baseline"] = cpu_baseline(W)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
