"""Parameterized circuits: forward states against the oracle (the restated
reference CPU path, eager plain method), expectation values, and adjoint
gradients against closed forms, parameter shift, central finite differences
and the FP64 path.  There is no reference counterpart for gradients."""
import math

import numpy as np
import pytest

import oracle_py as O
from zxh_adae_amd import workloads as W

pytestmark = pytest.mark.gpu


def oracle_gates(ops, params):
    """The same circuit in the reference's gate set: RY = U3(t, 0, 0),
    RZ = Rz (a P gate times the global phase e^{-it/2}), RX = Rx, CZ = CP(pi)."""
    from zxh_adae_amd.ansatz import RX, RY, RZ, H, CZ, CP
    gs = []
    for kind, q0, q1, p, v in ops:
        t = params[p] if p >= 0 else v
        if kind == RY:
            gs.append((W.U3, 0, q0, t, 0.0, 0.0))
        elif kind == RZ:
            gs.append((W.RZ, 0, q0, t, 0.0, 0.0))
        elif kind == RX:
            gs.append((W.RX, 0, q0, t, 0.0, 0.0))
        elif kind == H:
            gs.append((W.H, 0, q0, 0.0, 0.0, 0.0))
        elif kind == CZ:
            gs.append((W.CP, q0, q1, math.pi, 0.0, 0.0))
        elif kind == CP:
            gs.append((W.CP, q0, q1, t, 0.0, 0.0))
    return gs


def oracle_energy(n, ops, params, terms):
    o = O.OracleSim(n, False, True)
    o.gates(oracle_gates(ops, params))
    o.execute()
    psi = o.phys_state()  # plain method: physical address bit q = qubit q
    x = np.arange(1 << n, dtype=np.int64)
    h = np.zeros(1 << n)
    for t in terms:
        z = 1.0 - 2.0 * ((x >> t[1]) & 1)
        if len(t) > 2:
            z = z * (1.0 - 2.0 * ((x >> t[2]) & 1))
        h += t[0] * z
    p = np.abs(psi) ** 2
    return math.fsum(p * h), psi


def rand_terms(rng, n, count):
    terms = []
    for _ in range(count):
        c = float(rng.normal())
        if rng.random() < 0.5:
            terms.append((c, int(rng.integers(n))))
        else:
            i, j = rng.choice(n, 2, replace=False)
            terms.append((c, int(i), int(j)))
    return terms


def make(n, ops, nparams, prec, terms):
    from zxh_adae_amd.ansatz import Ansatz
    a = Ansatz(n, ops, nparams, prec)
    a.set_observable(terms)
    return a


def test_single_rotation_closed_form(gpu_engine):
    """RY(t) on qubit 0 of 13 idle qubits, H = Z_0: E = cos t, dE/dt = -sin t;
    RZ after RY leaves E unchanged (zero gradient); RX(t) on qubit 12 with
    Z_12 adds cos t'."""
    from zxh_adae_amd.ansatz import RX, RY, RZ
    n = 13
    ops = [(RY, 0, 0, 0, 0.0), (RZ, 0, 0, 1, 0.0), (RX, 12, 0, 2, 0.0)]
    for prec, tol in (("c128", 1e-12), ("c64", 2e-6)):
        a = make(n, ops, 3, prec, [(1.0, 0), (0.5, 12)])
        for t in (0.3, -1.1, 2.9):
            p = np.array([t, 0.7, t / 3])
            e, nrm = a.forward(p)
            assert abs(e - (math.cos(t) + 0.5 * math.cos(t / 3))) < tol
            assert abs(nrm - 1.0) < tol
            g, e2 = a.backward(p)
            assert abs(e2 - e) < tol
            assert abs(g[0] + math.sin(t)) < tol
            assert abs(g[1]) < tol
            assert abs(g[2] + 0.5 * math.sin(t / 3)) < tol


@pytest.mark.parametrize("n,layers", [(13, 2), (16, 3)])
def test_forward_state_and_energy_match_oracle(gpu_engine, n, layers):
    from zxh_adae_amd.ansatz import two_local
    rng = np.random.default_rng(n * 10 + layers)
    ops, npar = two_local(n, layers)
    params = rng.uniform(-math.pi, math.pi, npar)
    terms = rand_terms(rng, n, 30)
    e_ref, psi_ref = oracle_energy(n, ops, params, terms)
    a = make(n, ops, npar, "c128", terms)
    e, nrm = a.forward(params)
    psi = a.state()
    # the oracle's RZ carries the reference's global phase e^{-it/2}: compare
    # up to one global phase
    k = int(np.argmax(np.abs(psi_ref)))
    ph = psi[k] / psi_ref[k]
    assert abs(abs(ph) - 1.0) < 1e-12
    assert np.max(np.abs(psi - ph * psi_ref)) < 1e-12
    assert abs(e - e_ref) < 1e-12
    assert abs(nrm - 1.0) < 1e-12


def test_adjoint_matches_parameter_shift(gpu_engine):
    """Every parameter of a 3-layer two-local ansatz on 14 qubits:
    dE/dt = (E(t + pi/2) - E(t - pi/2)) / 2 exactly for RY and RZ."""
    from zxh_adae_amd.ansatz import two_local
    n, layers = 14, 3
    rng = np.random.default_rng(5)
    ops, npar = two_local(n, layers)
    params = rng.uniform(-math.pi, math.pi, npar)
    terms = rand_terms(rng, n, 40)
    a = make(n, ops, npar, "c128", terms)
    g, _ = a.backward(params)
    shift = np.zeros(npar)
    for k in range(npar):
        p = params.copy()
        p[k] += math.pi / 2
        ep, _ = a.forward(p)
        p[k] -= math.pi
        em, _ = a.forward(p)
        shift[k] = (ep - em) / 2
    assert np.max(np.abs(g - shift)) < 1e-11


def test_adjoint_matches_central_differences_with_mixed_gates(gpu_engine):
    """RX, RY, RZ, H, CZ and a parameterised-free CP, shared parameters, and a
    qubit carrying three rotations (split across passes): central finite
    differences with h = 1e-6 in FP64."""
    from zxh_adae_amd.ansatz import RX, RY, RZ, H, CZ, CP
    n = 13
    rng = np.random.default_rng(11)
    ops = []
    npar = 10
    for layer in range(3):
        for q in range(n):
            k = int(rng.integers(npar))
            kind = (RX, RY, RZ)[int(rng.integers(3))]
            ops.append((kind, q, 0, k, 0.0))
            if q % 4 == 0:
                ops.append((H, q, 0, -1, 0.0))
            if q % 5 == 1:
                ops.append((RZ, q, 0, (k + 3) % npar, 0.0))
                ops.append((RY, q, 0, (k + 4) % npar, 0.0))
        for q in range(0, n - 1, 2):
            ops.append((CZ, q, q + 1, -1, 0.0))
        ops.append((CP, 0, n - 1, -1, 0.37 * (layer + 1)))
    params = rng.uniform(-math.pi, math.pi, npar)
    terms = rand_terms(rng, n, 25)
    a = make(n, ops, npar, "c128", terms)
    e_ref, _ = oracle_energy(n, ops, params, terms)
    e, _ = a.forward(params)
    assert abs(e - e_ref) < 1e-12
    g, _ = a.backward(params)
    h = 1e-6
    fd = np.zeros(npar)
    for k in range(npar):
        p = params.copy()
        p[k] += h
        ep, _ = a.forward(p)
        p[k] -= 2 * h
        em, _ = a.forward(p)
        fd[k] = (ep - em) / (2 * h)
    assert np.max(np.abs(g - fd)) < 1e-7


def test_complex64_gradient_within_1e5_of_fp64(gpu_engine):
    from zxh_adae_amd.ansatz import two_local
    n, layers = 16, 4
    rng = np.random.default_rng(21)
    ops, npar = two_local(n, layers)
    params = rng.uniform(-math.pi, math.pi, npar)
    terms = rand_terms(rng, n, 60)
    g64, e64 = make(n, ops, npar, "c128", terms).backward(params)
    g32, e32 = make(n, ops, npar, "c64", terms).backward(params)
    scale = max(1.0, sum(abs(t[0]) for t in terms))
    print(f"c64 vs c128: max grad error {np.max(np.abs(g32 - g64)):.3e}, energy error {abs(e32 - e64):.3e}")
    assert np.max(np.abs(g32 - g64)) < 1e-5 * scale
    assert abs(e32 - e64) < 1e-5 * scale


def test_torch_autograd_function(gpu_engine):
    import torch
    from zxh_adae_amd.ansatz import two_local
    from zxh_adae_amd.torch_ext import expectation
    n, layers = 13, 2
    rng = np.random.default_rng(8)
    ops, npar = two_local(n, layers)
    p0 = rng.uniform(-math.pi, math.pi, npar)
    terms = rand_terms(rng, n, 20)
    a = make(n, ops, npar, "c128", terms)
    theta = torch.tensor(p0, dtype=torch.float64, requires_grad=True)
    e = expectation(theta, a)
    (2.0 * e).backward()
    g, e_np = a.backward(p0)
    assert abs(float(e.detach()) - e_np) < 1e-13
    assert np.max(np.abs(theta.grad.numpy() - 2.0 * g)) < 1e-13


def test_ansatz_rejects_bad_input(gpu_engine):
    from zxh_adae_amd import ZXHError
    from zxh_adae_amd.ansatz import Ansatz, RY, CZ
    with pytest.raises(ZXHError):
        Ansatz(8, [(RY, 0, 0, 0, 0.0)], 1, "c128")  # narrower than a tile
    with pytest.raises(ZXHError):
        Ansatz(13, [(RY, 13, 0, 0, 0.0)], 1, "c128")
    with pytest.raises(ZXHError):
        Ansatz(13, [(RY, 0, 0, 2, 0.0)], 1, "c128")
    with pytest.raises(ZXHError):
        Ansatz(13, [(CZ, 3, 3, -1, 0.0)], 0, "c128")
    a = Ansatz(13, [(RY, 0, 0, 0, 0.0)], 1, "c128")
    with pytest.raises(ZXHError):
        a.set_observable([(1.0, 13)])
    with pytest.raises(ValueError):
        a.forward(np.zeros(2))
