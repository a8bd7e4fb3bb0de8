"""How a single-qubit layer is split into tile passes with the wide tile on
(zxh_tile_layer_plan; host only, no device).

The split is this engine's own (the reference applies one sweep per gate,
zxh.cpp:390-428), so the checks are its contract: a partition of the
targets, pass 0 within the 4096-unit tile (every target below the low bits,
at most 12 - seg above), the other passes within the wide tile (at most nine,
none below the 16-unit run), the fewest passes those capacities allow, and
slots spread over address bits: no wide pass other than one contiguous with
the run holds two adjacent bits when the batch spans the whole state.
"""
import ctypes as C

import pytest

from zxh_adae_amd import _lib
from zxh_adae_amd import autotune as A


def plan(n, targets, c64=False, seg=4):
    return A.tile_layer_plan(n, list(targets), c64, seg)


@pytest.fixture
def wide_on():
    old = (A.tune_get("tile.wide.c128"), A.tune_get("tile.wide.c64"))
    A.tune_set("tile.wide.c128", 1)
    A.tune_set("tile.wide.c64", 1)
    yield
    A.tune_set("tile.wide.c128", old[0])
    A.tune_set("tile.wide.c64", old[1])


def check_plan(n, targets, c64, seg, passes):
    shift = 1 if c64 else 0
    lowb = seg + shift
    wlow = 4 + shift
    slots = 12 - seg
    flat = sorted(t for p in passes for t in p)
    assert flat == sorted(targets)
    nhigh = sum(1 for t in targets if t >= lowb)
    with_first = 1 + -(-(nhigh - slots) // 9)
    if min(targets) >= wlow and -(-len(targets) // 9) < with_first:
        # no target below the wide run: wide passes only, fewer of them
        assert len(passes) == -(-len(targets) // 9)
        wide = passes
    else:
        first = passes[0]
        assert set(t for t in targets if t < lowb) <= set(first)
        assert sum(1 for t in first if t >= lowb) <= slots
        assert len(passes) == with_first
        wide = passes[1:]
    for p in wide:
        assert 1 <= len(p) <= 9
        assert min(p) >= wlow


@pytest.mark.parametrize("c64", [False, True])
@pytest.mark.parametrize("n", [14, 20, 24, 28, 30, 32, 33, 36])
@pytest.mark.parametrize("seg", [3, 4, 5, 6])
def test_full_layer_is_a_minimal_partition(wide_on, c64, n, seg):
    targets = list(range(n))
    passes = plan(n, targets, c64, seg)
    slots = 12 - seg
    nhigh = n - (seg + (1 if c64 else 0))
    if nhigh <= slots:
        assert passes == []
        return
    check_plan(n, targets, c64, seg, passes)
    # spread slots: only a wide pass contiguous with the run has neighbours
    # (on narrower states the passes' room forces some)
    wlow = 4 + (1 if c64 else 0)
    for p in passes[1:] if n >= 24 else []:
        if p == list(range(wlow, wlow + 9)):
            continue
        assert all(b - a >= 2 for a, b in zip(p, p[1:])), p


def test_qft30_complex128_split(wide_on):
    # known answer: the run-contiguous wide pass, then bits 13..29 dealt in
    # turn between the last wide pass and the first pass
    passes = plan(30, range(30))
    assert passes == [[0, 1, 2, 3] + list(range(14, 29, 2)), list(range(4, 13)), list(range(13, 30, 2))]


def test_qft30_complex64_split(wide_on):
    passes = plan(30, range(30), c64=True)
    assert passes == [[0, 1, 2, 3, 4] + list(range(15, 30, 2)), list(range(5, 14)), list(range(14, 29, 2))]


@pytest.mark.parametrize("seed", range(6))
def test_random_batches(wide_on, seed):
    import numpy as np
    rng = np.random.default_rng(seed)
    for _ in range(40):
        n = int(rng.integers(14, 37))
        k = int(rng.integers(1, n + 1))
        targets = sorted(int(t) for t in rng.choice(n, size=k, replace=False))
        c64 = bool(rng.integers(0, 2))
        seg = int(rng.integers(3, 7))
        passes = plan(n, targets, c64, seg)
        nhigh = sum(1 for t in targets if t >= seg + (1 if c64 else 0))
        if nhigh <= 12 - seg:
            assert passes == []
            continue
        check_plan(n, targets, c64, seg, passes)


def test_wide_only_batches(wide_on):
    # nine targets above the run: one wide pass; eighteen: two
    assert plan(30, range(4, 13)) == [list(range(4, 13))]
    assert plan(30, range(12, 30)) == [list(range(12, 29, 2)), list(range(13, 30, 2))]
    # a target below the run needs the 4096-unit first pass
    assert len(plan(30, [0] + list(range(4, 13)))) == 2


def test_no_plan_cases(wide_on):
    assert plan(30, range(4, 12)) == []              # fits one pass
    assert plan(12, range(12)) == []                 # state narrower than the wide tile
    assert plan(30, [3, 4, 40]) == []                # beyond the state
    A.tune_set("tile.wide.c128", 0)
    assert plan(30, range(30)) == []                 # wide tile off


def test_bad_arguments():
    lib = _lib.load()
    ts = (C.c_uint32 * 2)(1, 2)
    out = (C.c_uint32 * 2)()
    npass = C.c_size_t(0)
    assert lib.zxh_tile_layer_plan(7, 30, 4, ts, 2, out, C.byref(npass)) == 1
    assert lib.zxh_tile_layer_plan(_lib.ZXH_C128, 41, 4, ts, 2, out, C.byref(npass)) == 1
    assert lib.zxh_tile_layer_plan(_lib.ZXH_C128, 30, 4, ts, 2, out, None) == 1
    assert lib.zxh_tile_layer_plan(_lib.ZXH_C128, 30, 4, None, 2, out, C.byref(npass)) == 1
