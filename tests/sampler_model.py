"""Pure-numpy model of the GPU sampler's algorithm (kernels_sample.hip).

Used only by tests, to check on the CPU that the tree/Fenwick reformulation
selects exactly the indices of the reference's streaming sampler (restated in
the oracle): balanced-tree node sums, Fenwick prefixes, running-max segment
tails and libstdc++'s upper_bound probe sequence.
"""
from __future__ import annotations

import numpy as np


def tree_levels(p: np.ndarray):
    lv = [p]
    while lv[-1].size > 1:
        a = lv[-1]
        lv.append(a[0::2] + a[1::2])
    return lv


def segment_tails(psi: np.ndarray, seg_bits: int) -> np.ndarray:
    """prefix_s + N_I[s] per segment (the bound each segment consumes below)."""
    psi = np.asarray(psi, dtype=np.complex128)
    m = psi.size.bit_length() - 1
    p = psi.real * psi.real + psi.imag * psi.imag
    lv = tree_levels(p)
    I = min(m, seg_bits)
    out = []
    for s in range(1 << (m - I)):
        acc = 0.0
        for l in range(m - I - 1, -1, -1):
            if (s >> l) & 1:
                acc = acc + float(lv[I + l][(s >> l) - 1])
        out.append(acc + float(lv[I][s]))
    return np.array(out)


def sample(psi: np.ndarray, seg_bits: int, thresholds: np.ndarray) -> np.ndarray:
    psi = np.asarray(psi, dtype=np.complex128)
    m = psi.size.bit_length() - 1
    p = psi.real * psi.real + psi.imag * psi.imag
    lv = tree_levels(p)
    I = min(m, seg_bits)
    nseg = 1 << (m - I)
    prefix = np.zeros(nseg)
    for s in range(nseg):
        acc = 0.0
        for l in range(m - I - 1, -1, -1):
            if (s >> l) & 1:
                acc = acc + float(lv[I + l][(s >> l) - 1])
        prefix[s] = acc
    tail = prefix + lv[I][:nseg]
    cmax = np.maximum.accumulate(tail)
    out = np.zeros(len(thresholds), dtype=np.uint64)
    valid = 1 << I
    for k, tau in enumerate(thresholds):
        seg = int(np.searchsorted(cmax, tau, side="right"))
        if seg >= nseg:
            raise RuntimeError("unresolved threshold")
        sbase = seg << I
        first, ln = 0, valid
        while ln > 0:
            half = ln >> 1
            mid = first + half
            g = sbase + mid
            acc = 0.0
            for l in range(I - 1, -1, -1):
                if (mid >> l) & 1:
                    acc = acc + float(lv[l][(g >> l) - 1])
            cdf = (prefix[seg] + acc) + float(p[g])
            if tau < cdf:
                ln = half
            else:
                first = mid + 1
                ln = ln - half - 1
        out[k] = sbase + (valid - 1 if first == valid else first)
    return out


def reference_cdf(psi: np.ndarray, seg_bits: int) -> np.ndarray:
    """Every cdf value the reference's segment walk compares against
    (prefix_s + excl_j) + p_j, in address order."""
    psi = np.asarray(psi, dtype=np.complex128)
    m = psi.size.bit_length() - 1
    p = psi.real * psi.real + psi.imag * psi.imag
    lv = tree_levels(p)
    I = min(m, seg_bits)
    out = np.zeros(psi.size)
    for g in range(psi.size):
        s, j = g >> I, g & ((1 << I) - 1)
        pre = 0.0
        for l in range(m - I - 1, -1, -1):
            if (s >> l) & 1:
                pre = pre + float(lv[I + l][(s >> l) - 1])
        acc = 0.0
        for l in range(I - 1, -1, -1):
            if (j >> l) & 1:
                acc = acc + float(lv[l][(g >> l) - 1])
        out[g] = (pre + acc) + float(p[g])
    return out
