"""Autotuner for the engine's kernel variants.

Times every variant of a pass on a device state with HIP events and keeps the
fastest, through the C ABI's tuning knobs (zxh_tune_set / zxh_tune_get).  The
knobs are process-wide; call tune() once per process before timing work.

Knobs swept:
  diag.ept.<prec>  0 = persistent diagonal sweep; 1, 2, 4 = flat grid with that
                   many 16-byte units per thread
  diag.nt.<prec>   nontemporal loads/stores in the flat sweep
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
