"""The window table of the tile passes' table epilogue, on the CPU (host
code of the library, no device): the basis spans exactly the selector forms
(GF(2) rank against an independent elimination), and for every address x
the table entry at y(x) equals the window's phase computed from its
definition, exp(i sum of theta over the gates whose selectors hold at x)."""
import ctypes as C
import math
import os

import numpy as np
import pytest

from zxh_adae_amd import _lib

pytestmark = pytest.mark.skipif(not os.path.exists(_lib.LIB_PATH), reason="native library not built")

NONE = (1 << 64) - 1
AMASK = (1 << 63) - 1


def gf2_rank(vs):
    rows = {}
    for v in vs:
        while v:
            top = v.bit_length() - 1
            if top not in rows:
                rows[top] = v
                break
            v ^= rows[top]
    return len(rows)


def window(rng, n, count, rank, beta=True):
    basis = [int(rng.integers(1, 1 << n)) for _ in range(rank)]

    def sel():
        a = 0
        for b in basis:
            if rng.random() < 0.5:
                a ^= b
        a = a or basis[0]
        return a | ((1 << 63) if beta and rng.random() < 0.3 else 0)
    s0 = [sel() for _ in range(count)]
    s1 = [sel() if rng.random() < 0.5 else NONE for _ in range(count)]
    th = [float(rng.uniform(-math.pi, math.pi)) for _ in range(count)]
    return s0, s1, th


def table(n, s0, s1, th):
    lib = _lib.load()
    a0 = np.array(s0, dtype=np.uint64)
    a1 = np.array(s1, dtype=np.uint64)
    t = np.array(th, dtype=np.float64)
    basis = np.zeros(10, dtype=np.uint64)
    rank = C.c_int(-1)
    tab = np.zeros(2048)
    rc = lib.zxh_window_table(n, a0.ctypes.data_as(_lib._u64p), a1.ctypes.data_as(_lib._u64p),
                              t.ctypes.data_as(_lib._dp), t.size, basis.ctypes.data_as(_lib._u64p),
                              C.byref(rank), tab.ctypes.data_as(_lib._dp))
    return rc, [int(b) for b in basis], rank.value, tab


def holds(sel, x):
    return (bin(sel & AMASK & x).count("1") & 1) ^ (sel >> 63)


@pytest.mark.parametrize("n,count,rank,seed", [(8, 20, 3, 1), (10, 40, 7, 2), (12, 60, 10, 3), (9, 30, 9, 4),
                                               (20, 50, 10, 5)])
def test_table_is_the_window_phase(n, count, rank, seed):
    rng = np.random.default_rng(seed)
    s0, s1, th = window(rng, n, count, rank)
    rc, basis, r, tab = table(n, s0, s1, th)
    assert rc == 0
    forms = [s & AMASK for s in s0] + [s & AMASK for s in s1 if s != NONE]
    assert r == gf2_rank(forms) <= rank
    assert gf2_rank(basis[:r]) == r and all(b == 0 for b in basis[r:])
    assert gf2_rank(forms + basis[:r]) == r  # the basis spans the forms
    xs = range(1 << n) if n <= 12 else [int(v) for v in rng.integers(0, 1 << n, 4096)]
    for x in xs:
        y = sum((bin(basis[k] & x).count("1") & 1) << k for k in range(r))
        phase = sum(t for a, b, t in zip(s0, s1, th) if holds(a, x) and (b == NONE or holds(b, x)))
        assert abs(complex(tab[2 * y], tab[2 * y + 1]) - complex(math.cos(phase), math.sin(phase))) < 1e-13


def test_constant_and_empty_windows():
    # alpha = 0, beta = 1: a selector that always holds; rank 0, one entry
    rc, basis, r, tab = table(6, [1 << 63, 1 << 63], [NONE, 1 << 63], [0.25, 0.5])
    assert rc == 0 and r == 0
    assert abs(complex(tab[0], tab[1]) - complex(math.cos(0.75), math.sin(0.75))) < 1e-15
    rc, basis, r, tab = table(6, [], [], [])
    assert rc == 0 and r == 0 and tab[0] == 1.0 and tab[1] == 0.0


def test_rejects_wide_or_out_of_state_windows():
    rng = np.random.default_rng(7)
    s0, s1, th = window(rng, 16, 60, 11)
    assert table(16, s0, s1, th)[0] == 1
    assert table(8, [1 << 9], [NONE], [0.3])[0] == 1  # selector above the state
