"""Circuit front end and backend on the CPU: the native-gate loader against a
recording simulator (the pattern of the reference's
src/zxh/tests/unit/test_qasm_load_circuit.py:14-75), the exactness of every
frontend-gate lowering against the oracle's unitaries, the OpenQASM 2
reader on the reference's integration cases, and the backend's run / job /
result behaviour with a fake simulator (no GPU)."""
import json
import math
import os

import numpy as np
import pytest

import oracle_py as O
from zxh_adae_amd import backend as B
from zxh_adae_amd import frontend as F

CASES = os.path.join(os.path.dirname(__file__), "golden", "integration_cases")


class Recorder:
    def __init__(self, n=4):
        self.n = n
        self.gates = []
        self.seed = None
        self.cleared = 0
        self.executed = 0

    def num_qubits(self):
        return self.n

    def clear_gates(self):
        self.gates.clear()
        self.cleared += 1

    def Barrier(self):
        self.gates.append(("Barrier",))

    def X(self, q):
        self.gates.append(("X", q))

    def CX(self, cq, q):
        self.gates.append(("CX", cq, q))

    def H(self, q):
        self.gates.append(("H", q))

    def Z(self, q):
        self.gates.append(("Z", q))

    def Rz(self, q, t):
        self.gates.append(("Rz", q, t))

    def Rx(self, q, t):
        self.gates.append(("Rx", q, t))

    def P(self, q, t):
        self.gates.append(("P", q, t))

    def U3(self, q, theta, lam, phi):
        self.gates.append(("U3", q, theta, lam, phi))

    def CP(self, cq, q, t):
        self.gates.append(("CP", cq, q, t))

    def CRz(self, cq, q, t):
        self.gates.append(("CRz", cq, q, t))

    # backend surface
    def set_seed(self, s):
        self.seed = ("set", s)

    def clear_seed(self):
        self.seed = ("clear",)

    def execute(self):
        self.executed += 1

    def sample_counts(self, shots):
        return {"0" * self.n: shots}


def test_native_loader_forwards_in_order_with_u3_argument_order():
    c = F.Circuit(3)
    c.h(0).cx(0, 1).rz(0.3, 2).rx(-0.2, 1).p(0.7, 0).u(0.1, 0.2, 0.3, 2).cp(0.4, 1, 2).crz(0.5, 0, 2)
    c.z(1).x(2).barrier().append("id", [0]).measure_all()
    r = Recorder(3)
    F.load_circuit_transpiled(r, c)
    assert r.gates == [("H", 0), ("CX", 0, 1), ("Rz", 2, 0.3), ("Rx", 1, -0.2), ("P", 0, 0.7),
                       ("U3", 2, 0.1, 0.3, 0.2), ("CP", 1, 2, 0.4), ("CRz", 0, 2, 0.5), ("Z", 1), ("X", 2),
                       ("Barrier",)]


def test_native_loader_reset_clears_and_rejects_frontend_gates():
    c = F.Circuit(2).h(0).append("reset", [0]).x(1)
    r = Recorder(2)
    F.load_circuit_transpiled(r, c)
    assert r.gates == [("X", 1)]
    with pytest.raises(NotImplementedError):
        F.load_circuit_transpiled(Recorder(2), F.Circuit(2).swap(0, 1))
    with pytest.raises(ValueError):
        F.load_circuit_transpiled(Recorder(1), F.Circuit(1).append("rz", [0], ["theta"]))


def _native_to_oracle(lowered):
    gs = []
    for name, qs, ps in lowered:
        if name == "h":
            gs.append((O.H, 0, qs[0], 0.0, 0.0, 0.0))
        elif name == "x":
            gs.append((O.X, 0, qs[0], 0.0, 0.0, 0.0))
        elif name == "cx":
            gs.append((O.CX, qs[0], qs[1], 0.0, 0.0, 0.0))
        elif name == "p":
            gs.append((O.P, 0, qs[0], ps[0], 0.0, 0.0))
        elif name == "rz":
            gs.append((O.RZ, 0, qs[0], ps[0], 0.0, 0.0))
        elif name == "u":
            theta, phi, lam = ps
            gs.append((O.U3, 0, qs[0], theta, lam, phi))
        elif name == "cp":
            gs.append((O.CP, qs[0], qs[1], ps[0], 0.0, 0.0))
        else:
            raise AssertionError(name)
    return gs


def oracle_unitary(n, lowered):
    """Columns: the oracle's plain-method state of each basis input (X gates
    prepare it; disable_x makes X and CX explicit data movement), physical
    address bit q = qubit q."""
    cols = []
    for b in range(1 << n):
        prep = [(O.X, 0, q, 0.0, 0.0, 0.0) for q in range(n) if (b >> q) & 1]
        o = O.OracleSim(n, True, True)
        o.gates(prep + [(O.BARRIER, 0, 0, 0.0, 0.0, 0.0)] + _native_to_oracle(lowered))
        o.execute()
        cols.append(o.phys_state())
    return np.array(cols).T


def textbook(name, n, qs, ps):
    """Qiskit's matrix of the gate, little-endian (bit q = qubit q)."""
    dim = 1 << n
    u = np.zeros((dim, dim), dtype=complex)
    for x in range(dim):
        bits = [(x >> q) & 1 for q in qs]
        if name in ("u1", "s", "sdg", "t", "tdg"):
            lam = ps[0] if name == "u1" else F._PHASE[name]
            u[x, x] = np.exp(1j * lam) if bits[0] else 1.0
        elif name in ("u2", "ry"):
            if name == "u2":
                th, phi, lam = math.pi / 2, ps[0], ps[1]
            else:
                th, phi, lam = ps[0], 0.0, 0.0
            m = np.array([[math.cos(th / 2), -np.exp(1j * lam) * math.sin(th / 2)],
                          [np.exp(1j * phi) * math.sin(th / 2), np.exp(1j * (phi + lam)) * math.cos(th / 2)]])
            x0 = x & ~(1 << qs[0])
            for bo in (0, 1):
                u[x0 | (bo << qs[0]), x] += m[bo, bits[0]]
        elif name == "cz":
            u[x, x] = -1.0 if bits[0] and bits[1] else 1.0
        elif name == "swap":
            y = x & ~((1 << qs[0]) | (1 << qs[1]))
            y |= (bits[0] << qs[1]) | (bits[1] << qs[0])
            u[y, x] = 1.0
        elif name == "rzz":
            par = bits[0] ^ bits[1]
            u[x, x] = np.exp(-1j * ps[0] / 2) if par == 0 else np.exp(1j * ps[0] / 2)
        elif name == "ccz":
            u[x, x] = -1.0 if all(bits) else 1.0
        elif name == "ccx":
            y = x ^ (1 << qs[2]) if bits[0] and bits[1] else x
            u[y, x] = 1.0
        else:
            raise AssertionError(name)
    return u


@pytest.mark.parametrize("name,qs,ps", [("u1", [1], [0.37]), ("s", [0], []), ("sdg", [2], []), ("t", [1], []),
                                        ("tdg", [0], []), ("u2", [1], [0.23, -0.71]), ("ry", [2], [1.1]),
                                        ("cz", [0, 2], []), ("swap", [2, 0], []), ("rzz", [1, 2], [0.8]),
                                        ("ccz", [2, 0, 1], []), ("ccx", [0, 1, 2], []), ("ccx", [2, 1, 0], [])])
def test_lowering_is_exact_up_to_global_phase(name, qs, ps):
    n = 3
    got = oracle_unitary(n, F.lower_instruction(name, qs, ps))
    want = textbook(name, n, qs, ps)
    k = np.unravel_index(np.argmax(np.abs(want)), want.shape)
    ph = got[k] / want[k]
    assert abs(abs(ph) - 1.0) < 1e-12
    assert np.max(np.abs(got - ph * want)) < 1e-12


def test_qasm_expressions_and_errors():
    ev = F._eval_expr
    assert ev("pi/2") == math.pi / 2
    assert ev("-pi/4") == -math.pi / 4
    assert abs(ev("2*pi/3 + 1e-3") - (2 * math.pi / 3 + 1e-3)) < 1e-15
    assert ev("(pi+1)/2") == (math.pi + 1) / 2
    assert ev("-(0.5)") == -0.5
    assert ev("2^3") == 8.0
    for bad in ("pi/", "(1", "foo", "1 2", ""):
        with pytest.raises(ValueError):
            ev(bad)


def test_qasm_reader_on_the_reference_cases():
    bell = F.load_qasm(os.path.join(CASES, "bell_2_support.qasm"))
    assert bell.num_qubits == 2 and bell.num_clbits == 2
    assert [(i.operation.name, i.qubits) for i in bell.data] == [("h", (0,)), ("cx", (0, 1)), ("measure", (0, 1))]
    u3 = F.load_qasm(os.path.join(CASES, "u3_1_balanced.qasm"))
    assert u3.data[0].operation.params == (math.pi / 2, 0.0, 0.0)
    h11 = F.load_qasm(os.path.join(CASES, "h_11_inter_worker_delta.qasm"))
    assert h11.num_qubits == 11 and sum(i.operation.name == "h" for i in h11.data) == 22
    multi = F.parse_qasm('OPENQASM 2.0; include "qelib1.inc"; qreg a[2]; qreg b[1]; h a; cx a[1], b[0]; '
                         "barrier a, b; rz(pi) b[0];")
    assert [(i.operation.name, i.qubits) for i in multi.data] == [("h", (0,)), ("h", (1,)), ("cx", (1, 2)),
                                                                   ("barrier", (0, 1, 2)), ("rz", (2,))]
    with pytest.raises(ValueError):
        F.parse_qasm("qreg q[1]; h q[3];")
    with pytest.raises(FileNotFoundError):
        F.load_qasm(os.path.join(CASES, "missing.qasm"))


def test_fixture_expectations_cover_every_case():
    exp = json.load(open(os.path.join(CASES, "expectations.json")))
    files = sorted(f[:-5] for f in os.listdir(CASES) if f.endswith(".qasm"))
    assert sorted(exp["cases"]) == files


def fake_backend(**kw):
    made = []

    def factory(n, dx, eager):
        r = Recorder(n)
        r.flags = (dx, eager)
        made.append(r)
        return r
    return B.ZXHBackend(sim_factory=factory, **kw), made


def test_backend_run_job_result_and_options():
    be, made = fake_backend(disable_x=True)
    c = F.Circuit(3, name="ghz").h(0).cx(0, 1).cx(1, 2).swap(0, 2).measure_all()
    job = be.run(c, shots=100, seed_simulator=7)
    assert job.status() == "DONE"
    res = job.result()
    assert res.success and res.backend_name == "zxh"
    assert res.get_counts() == {"000": 100}
    assert res.results[0].shots == 100 and res.results[0].seed == 7
    assert res.results[0].metadata["sample_time"] >= 0.0
    sim = made[0]
    assert sim.flags == (True, False)
    assert sim.seed == ("set", 7)
    assert sim.executed == 1
    assert ("CX", 0, 2) in sim.gates  # swap lowered to three CX
    # unseeded runs clear the seed; options default to one shot
    res2 = be.run([c, c]).result()
    assert sim.seed == ("clear",)
    assert res2.get_counts() == [{"000": 1}, {"000": 1}]
    with pytest.raises(B.JobError):
        job.submit()


def test_backend_refuses_construction_options_and_other_widths():
    be, _ = fake_backend()
    with pytest.raises(B.JobError):
        be.run(F.Circuit(2).h(0), disable_x=True)
    with pytest.raises(B.JobError):
        be.run(F.Circuit(2).h(0), eager_expand_all=True)
    be.run(F.Circuit(2).h(0))
    with pytest.raises(B.JobError):
        be.run(F.Circuit(3).h(0))
    with pytest.raises(TypeError):
        be.run(42)


def test_registry_names():
    assert B.all_backends() == ["zxh", "zxh-nox", "zxh-exp"]
    flags = {}
    for name in B.all_backends():
        be = B.make_backend(name, sim_factory=lambda n, dx, eager: Recorder(n))
        flags[name] = (be._disable_x, be._eager_expand_all)
    assert flags == {"zxh": (False, False), "zxh-nox": (True, False), "zxh-exp": (False, True)}
    with pytest.raises(ValueError):
        B.make_backend("ddsim")
