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


def normal(r: SplitMix64) -> float:
    """N(0, 1) by the Box-Muller transform of two stream uniforms (explicit,
    so the value does not depend on a library's normal generator)."""
    u1 = r.uniform()
    u2 = r.uniform()
    return math.sqrt(-2.0 * math.log(1.0 - u1)) * math.cos(TWO_PI * u2)


def zz_observable(n: int, count: int, seed: int) -> List[tuple]:
    """count terms (coef, i) for Z_i with probability 1/2, else (coef, i, j)
    for Z_i Z_j on distinct i < j; coefficients N(0, 1)."""
    r = SplitMix64(seed)
    terms = []
    for _ in range(count):
        if r.uniform() < 0.5:
            terms.append((normal(r), r.below(n)))
        else:
            i = r.below(n)
            j = r.below(n - 1)
            if j >= i:
                j += 1
            terms.append((normal(r), min(i, j), max(i, j)))
    return terms


def ansatz_params(count: int, seed: int) -> List[float]:
    """count angles uniform in [-pi, pi)."""
    r = SplitMix64(seed)
    return [math.pi * (2.0 * r.uniform() - 1.0) for _ in range(count)]


def haar_unitary(r: SplitMix64, d: int = 32):
    """Haar-random d x d unitary: QR of a complex Ginibre matrix (entries
    N(0,1)/sqrt(2) + i N(0,1)/sqrt(2) from the stream by Box-Muller) with the
    phases of R's diagonal moved into Q (Mezzadri's construction)."""
    import numpy as np
    z = np.array([[complex(normal(r), normal(r)) for _ in range(d)] for _ in range(d)]) / math.sqrt(2.0)
    q, rr = np.linalg.qr(z)
    ph = np.diag(rr) / np.abs(np.diag(rr))
    return q * ph[None, :]


def c5_layers(n: int, layers: int, seed: int, mcx_per_layer: int = 8):
    """The large random circuit: per layer, floor(n/5) disjoint 5-qubit Haar
    blocks on a random qubit permutation, a diagonal batch of 64 P/CP gates,
    and a cascade of multi-controlled X gates (2..6 controls, distinct random
    qubits).  Returns a list of layers, each a dict with
      "blocks": [(targets[5], u 32x32)], "diag": [(code, cq, q, theta)],
      "mcx": [(control_mask, target)]."""
    r = SplitMix64(seed)
    out = []
    for _ in range(layers):
        perm = list(range(n))
        for i in range(n - 1, 0, -1):  # Fisher-Yates from the stream
            j = r.below(i + 1)
            perm[i], perm[j] = perm[j], perm[i]
        blocks = [(perm[5 * b:5 * b + 5], haar_unitary(r)) for b in range(n // 5)]
        diag = [(g[0], g[1], g[2], g[3]) for g in diag_batch(n, 64, r.next_u64() & 0xFFFFFFFF)]
        mcx = []
        for _ in range(mcx_per_layer):
            k = 2 + r.below(5)
            qs = []
            while len(qs) < k + 1:
                q = r.below(n)
                if q not in qs:
                    qs.append(q)
            mcx.append((sum(1 << q for q in qs[:k]), qs[k]))
        out.append({"blocks": blocks, "diag": diag, "mcx": mcx})
    return out


def c4_sets(count: int = 512, seed: int = 4, terms: int = 50):
    """The batched uneven ansatz workload: count parameter sets of the
    two-local ansatz with widths n ~ U{16..24} and depths ~ U{2..10}, each
    with its own terms-term Z/ZZ observable.  Returns a list of
    (n, layers, observable terms); the observables' seeds derive from seed."""
    r = SplitMix64(seed)
    out = []
    for i in range(count):
        n = 16 + r.below(9)
        layers = 2 + r.below(9)
        out.append((n, layers, zz_observable(n, terms, (seed << 32) + i)))
    return out
