"""Autotuner for the engine's kernel variants.

Times every variant of a pass on a device state with HIP events and keeps the
fastest, through the C ABI's tuning knobs (zxh_tune_set / zxh_tune_get).  The
knobs are process-wide; call tune() once per process before timing work.

Knobs swept:
  diag.ept.<prec>  0 = persistent diagonal sweep; 1, 2, 4 = flat grid with that
                   many 16-byte units per thread
  diag.nt.<prec>   nontemporal loads/stores in the flat sweep
  tile.seg.<prec>  contiguous unit-address bits of a tile pass (3, 4, 5, 6:
                   every wave access moves 128 B, 256 B, 512 B or 1 KiB
                   contiguous; a pass then reaches 9, 8, 7 or 6 targets above
                   them)
"""
from __future__ import annotations

import ctypes as C
from typing import Dict, Tuple

from . import _lib
from ._lib import check
from . import kernels as K
from . import workloads as W


def tune_set(name: str, value: int) -> None:
    check(_lib.load().zxh_tune_set(name.encode(), int(value)))


def tune_get(name: str) -> int:
    v = C.c_int()
    check(_lib.load().zxh_tune_get(name.encode(), C.byref(v)))
    return v.value


DIAG_VARIANTS = [(0, 0), (1, 0), (1, 1), (2, 0), (2, 1), (4, 0), (4, 1)]


def tune_diag(state: "K.DeviceState", repeats: int = 3, batch: int = 64) -> Dict[str, object]:
    """Sweep the diagonal-sweep variants on state (contents are scaled by unit
    phases only, so any state can be used).  Sets the fastest; returns the
    per-variant average milliseconds."""
    prec = "c64" if state.prec == _lib.ZXH_C64 else "c128"
    knob = f"diag.ept.{prec}"
    ntk = f"diag.nt.{prec}"
    gs = W.diag_batch(state.nbits, batch, 4242)
    s0 = [(1 << g[2]) for g in gs]
    s1 = [K.NO_SEL if g[0] == W.P else (1 << g[1]) for g in gs]
    th = [g[3] for g in gs]
    results: Dict[Tuple[int, int], float] = {}
    old = (tune_get(knob), tune_get(ntk))
    for ept, nt in DIAG_VARIANTS:
        tune_set(knob, ept)
        tune_set(ntk, nt)
        state.diag(s0, s1, th)  # warm
        K.reset_timing()
        K.set_timing(True)
        for _ in range(repeats):
            state.diag(s0, s1, th)
        t = K.timing("diag")
        K.set_timing(False)
        results[(ept, nt)] = t["ms"] / max(t["launches"], 1)
    best = min(results, key=results.get)
    tune_set(knob, best[0])
    tune_set(ntk, best[1])
    return {"knob": knob, "best": {"ept": best[0], "nt": best[1]}, "previous": {"ept": old[0], "nt": old[1]},
            "ms": {f"ept{e}_nt{n}": round(v, 4) for (e, n), v in results.items()}}


TILE_SEGS = (3, 4, 5, 6)


def tile_layer_plan(n: int, targets, c64: bool, seg: int = -1):
    """The passes the engine splits a batch of single-qubit gates on distinct
    address bits into when it needs more than one tile pass and the wide
    tile is on (zxh_tile_layer_plan): a list of target lists, the first for a
    4096-unit tile pass, the others wide (all wide when no target lies below
    the wide tile's run and that takes fewer passes); [] when the batch fits
    one 4096-unit pass or no such split applies.  Reads the tile.wide knob as
    it stands."""
    ts = (C.c_uint32 * max(1, len(targets)))(*targets)
    out = (C.c_uint32 * max(1, len(targets)))()
    npass = C.c_size_t(0)
    prec = _lib.ZXH_C64 if c64 else _lib.ZXH_C128
    check(_lib.load().zxh_tile_layer_plan(prec, n, seg, ts, len(targets), out, C.byref(npass)))
    passes = [[] for _ in range(npass.value)]
    for i, t in enumerate(targets):
        if npass.value:
            passes[out[i]].append(int(t))
    return passes


def layer_passes(n: int, seg: int, c64: bool, wide: bool = False):
    """Targets of each tile pass of a full-width single-qubit layer under
    geometry seg: the low bits ride in the first pass, then 12 - seg upper
    targets per pass.  With the wide (8192-unit) tile, the split the engine
    makes with its knob on (tile_layer_plan)."""
    low = seg + (1 if c64 else 0)
    slots = 12 - seg
    upper = list(range(low, n))
    if wide:
        wknob = "tile.wide.c64" if c64 else "tile.wide.c128"
        old = tune_get(wknob)
        tune_set(wknob, 1)
        try:
            plan = tile_layer_plan(n, list(range(n)), c64, seg)
        finally:
            tune_set(wknob, old)
        if plan:
            return plan
    chunks = [upper[i:i + slots] for i in range(0, len(upper), slots)] or [[]]
    chunks[0] = list(range(low)) + chunks[0]
    return chunks


TILE_TOLERANCE = 0.03


def tune_tile(state: "K.DeviceState", repeats: int = 2, window: int = 32) -> Dict[str, object]:
    """Times a full-width layer per tile geometry on state, with and without
    the wide (8192-unit) tile: U3 on every qubit, split into the geometry's
    tile passes (layer_passes), each pass fused with a window of P/CP gates.
    Fewer contiguous bits per tile leave room for more targets per pass
    (fewer passes) at shorter contiguous runs, so the choice is by the
    layer's total time: the fastest.  Sets the geometry and the wide-tile
    knob; the state is changed by the passes."""
    c64 = state.prec == _lib.ZXH_C64
    knob = "tile.seg.c64" if c64 else "tile.seg.c128"
    wknob = "tile.wide.c64" if c64 else "tile.wide.c128"
    n = state.nbits
    old = tune_get(knob)
    old_wide = tune_get(wknob)
    gs = W.diag_batch(n, window, 4343)
    s0 = [(1 << g[2]) for g in gs]
    s1 = [K.NO_SEL if g[0] == W.P else (1 << g[1]) for g in gs]
    th = [g[3] for g in gs]
    results: Dict[tuple, float] = {}
    npasses: Dict[tuple, int] = {}
    for seg, wide in [(sg, w) for w in (0, 1) for sg in TILE_SEGS]:
        tune_set(knob, seg)
        tune_set(wknob, wide)
        passes = [[((1 << q), (1 << q), W.u3_matrix(0.3 + 0.01 * q, 0.2, 0.1)) for q in chunk]
                  for chunk in layer_passes(n, seg, c64, bool(wide))]
        for descs in passes:  # warm
            state.fused_pass(descs, s0, s1, th)
        K.reset_timing()
        K.set_timing(True)
        for _ in range(repeats):
            for descs in passes:
                state.fused_pass(descs, s0, s1, th)
        t = K.timing("fused")
        K.set_timing(False)
        results[(seg, wide)] = t["ms"] / repeats
        npasses[(seg, wide)] = len(passes)
    # the fastest layer (the wide tile changes the pass split, so the
    # longest-runs preference among near ties no longer picks the faster
    # steps: at 30 qubits c128 seg 6 with the wide tile is within 3% of seg 4
    # on this layer and 4.5% slower on the QFT step)
    best = min(results, key=lambda k: results[k])
    tune_set(knob, best[0])
    tune_set(wknob, best[1])

    def name(k):
        return f"seg{k[0]}" + ("_wide" if k[1] else "")
    return {"knob": knob, "best": best[0], "wide": best[1], "previous": old, "previous_wide": old_wide,
            "layer_ms": {name(k): round(v, 4) for k, v in results.items()},
            "passes": {name(k): v for k, v in npasses.items()}}


def tune_batch(sets, precision: str = "c64", sample: int = 8, repeats: int = 2, seed: int = 11) -> Dict[str, object]:
    """Tile geometry per width class of a batched-ansatz workload, by timing.

    sets: the workload's (n, ops, nparams, terms).  For every width n, up to
    `sample` of its sets form a small batch, timed (forward and adjoint, one
    graph replay each, after a warm-up) under the fewest-passes rule (0) and
    under each geometry 3..6; the fastest is kept, the rule when it is within
    TILE_TOLERANCE of the fastest.  Returns {n: seg} with the timings."""
    import time

    import numpy as np

    from .ansatz import AnsatzBatch

    by_width: Dict[int, list] = {}
    for s in sets:
        by_width.setdefault(int(s[0]), []).append(s)
    choice: Dict[int, int] = {}
    timings: Dict[str, Dict[str, float]] = {}
    rng = np.random.default_rng(seed)
    for n in sorted(by_width):
        group = by_width[n][:sample]
        total = sum(int(s[2]) for s in group)
        params = rng.uniform(-np.pi, np.pi, total)
        res: Dict[int, float] = {}
        for seg in (0,) + TILE_SEGS:
            b = AnsatzBatch(precision)
            try:
                b.set_geometry(n, seg)
                for (w, ops, npar, terms) in group:
                    b.add(w, ops, npar, terms)
                b.finalize()
                b.run(params)
                t0 = time.perf_counter()
                for _ in range(repeats):
                    b.run(params)
                res[seg] = (time.perf_counter() - t0) / repeats
            finally:
                b.close()
        fastest = min(res.values())
        best = 0 if res[0] <= fastest * (1.0 + TILE_TOLERANCE) else min(res, key=res.get)
        choice[n] = best
        timings[str(n)] = {("rule" if k == 0 else f"seg{k}"): round(v * 1e3, 3) for k, v in res.items()}
    return {"choice": {str(k): v for k, v in choice.items()}, "ms": timings, "sample": sample}
