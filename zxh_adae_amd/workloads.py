"""Seeded synthetic circuits for the engine's benchmark and parity cases.

Gates are tuples ``(code, cq, q, theta, lambda, phi)`` with the reference's
gate codes (H=9, U3=10, P=4, CP=6, ...), replayable through ``ZXH.gates`` or
the oracle.  Angles come from a SplitMix64 stream, u = (x >> 11) * 2^-53, so
the same seed gives the same circuit on every machine.
"""
from __future__ import annotations

import math
from typing import List, Tuple

BARRIER, Z, RZ, P, CRZ, CP, X, CX, H, U3, RX = 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11
TWO_PI = 2.0 * math.pi

Gate = Tuple[int, int, int, float, float, float]


class SplitMix64:
    def __init__(self, seed: int):
        self.s = seed & 0xFFFFFFFFFFFFFFFF

    def next_u64(self) -> int:
        self.s = (self.s + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        return z ^ (z >> 31)

    def uniform(self) -> float:
        return (self.next_u64() >> 11) * (1.0 / 9007199254740992.0)

    def below(self, n: int) -> int:
        return int(self.uniform() * n)


def random_layers(n: int, layers: int, seed: int) -> List[Gate]:
    """Each layer: U3 on every qubit (angles in [0, 2pi)), a linear CP ladder,
    then P on every qubit."""
    r = SplitMix64(seed)
    gs: List[Gate] = []
    for _ in range(layers):
        for q in range(n):
            th, ph, la = r.uniform() * TWO_PI, r.uniform() * TWO_PI, r.uniform() * TWO_PI
            gs.append((U3, 0, q, th, la, ph))
        for q in range(n - 1):
            gs.append((CP, q, q + 1, r.uniform() * TWO_PI, 0.0, 0.0))
        for q in range(n):
            gs.append((P, 0, q, r.uniform() * TWO_PI, 0.0, 0.0))
    return gs


def qft(n: int) -> List[Gate]:
    """H on each target, then CP(pi / 2^(t-j)) from every lower qubit j; no swaps."""
    gs: List[Gate] = []
    for t in range(n):
        gs.append((H, 0, t, 0.0, 0.0, 0.0))
        for j in range(t):
            gs.append((CP, j, t, math.pi / (2 ** (t - j)), 0.0, 0.0))
    return gs


def diag_batch(n: int, count: int, seed: int) -> List[Gate]:
    """count random diagonal gates: P(q) with probability 1/2, else CP on a
    distinct random pair; angles in [0, 2pi)."""
    r = SplitMix64(seed)
    gs: List[Gate] = []
    for _ in range(count):
        if r.uniform() < 0.5:
            gs.append((P, 0, r.below(n), r.uniform() * TWO_PI, 0.0, 0.0))
        else:
            a = r.below(n)
            b = r.below(n - 1)
            if b >= a:
                b += 1
            gs.append((CP, a, b, r.uniform() * TWO_PI, 0.0, 0.0))
    return gs


def u3_matrix(theta: float, lam: float, phi: float):
    """The 2x2 U3 matrix with the reference's convention (zxh.cpp:101-113):
    [[c, -e^{i lam} s], [e^{i phi} s, e^{i (phi + lam)} c]], c, s = cos, sin(theta/2)."""
    import numpy as np
    c, s = math.cos(theta / 2.0), math.sin(theta / 2.0)
    return np.array([[c, -np.exp(1j * lam) * s], [np.exp(1j * phi) * s, np.exp(1j * (phi + lam)) * c]])


def hadamard_all(n: int) -> List[Gate]:
    return [(H, 0, q, 0.0, 0.0, 0.0) for q in range(n)]


def u3_all(n: int, seed: int) -> List[Gate]:
    r = SplitMix64(seed)
    return [(U3, 0, q, r.uniform() * TWO_PI, r.uniform() * TWO_PI, r.uniform() * TWO_PI) for q in range(n)]


def random_mixed(n: int, count: int, seed: int, kinds=(H, U3, P, CP, RZ, Z, X, CX, CRZ, RX, BARRIER)) -> List[Gate]:
    """Random gate soup over every recorded gate kind (affine-map coverage)."""
    r = SplitMix64(seed)
    gs: List[Gate] = []
    for _ in range(count):
        k = kinds[r.below(len(kinds))]
        q = r.below(n)
        if k in (CP, CX, CRZ) and n < 2:
            k = P
        cq = q
        if k in (CP, CX, CRZ):
            cq = r.below(n - 1)
            if cq >= q:
                cq += 1
        gs.append((k, cq, q, r.uniform() * TWO_PI, r.uniform() * TWO_PI, r.uniform() * TWO_PI))
    return gs
