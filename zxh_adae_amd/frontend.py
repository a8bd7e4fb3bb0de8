"""Circuit front end: the native-gate loader, an exact lowering of the wider
frontend gate set, and an OpenQASM 2 reader, for circuits given as Qiskit
``QuantumCircuit`` objects or as this module's ``Circuit``.

The loader follows the reference's ``load_circuit_transpiled``
(src/zxh/zxhsim/qasm.py:62-141): it walks ``circuit.data`` and calls the
simulator's gate methods, with Qiskit's U/U3(theta, phi, lambda) mapped to
ZXH's ``U3(q, theta, lambda, phi)`` argument order, measurements and ``id``
ignored, ``reset`` clearing the recorded gates, and anything outside the
native set rejected with ``NotImplementedError``.  The native set is the
reference's ``ZXH_NATIVE_GATES`` (src/zxh/zxhsim/gate_sets.py:35-51).

The frontend set (gate_sets.py:4-32) adds gates the reference lowers with its
Qiskit-based optimiser (circ_optimizer.py); here each is replaced by an exact
native sequence of this module's own choosing (checked against the oracle's
unitaries in the tests):
    u1(l) = s, sdg, t, tdg -> P(angle)        u2(f, l) -> U3(pi/2, l, f)
    ry(t) -> U3(t, 0, 0)                      cz -> CP(pi)
    swap -> CX CX CX                          rzz(t) -> CX Rz CX
    ccz -> 6 CX + 7 P(+-pi/4)                 ccx -> H ccz H
    cswap(c, a, b) -> CX(b, a) ccx(c, a, b) CX(b, a)
Phase-type gates become P rather than the reference's Rz, so get_state can
differ from the reference's by a global phase; samples and counts cannot.

Without Qiskit, ``Circuit`` builds circuits with the same attributes the
loader reads (``data``, ``num_qubits``, instructions with ``operation.name``,
``operation.params`` and ``qubits``), and ``parse_qasm`` reads OpenQASM 2
programs over qelib1 gates (``qreg``, ``creg``, gates with arithmetic
parameters in pi, whole-register arguments, ``measure``, ``barrier``,
``reset``).
"""
from __future__ import annotations

import math
import re
from dataclasses import dataclass, field
from pathlib import Path
from typing import List, Sequence

NATIVE_GATES = ("barrier", "measure", "reset", "id", "x", "cx", "h", "z", "rz", "rx", "p", "u", "u3", "cp", "crz")
FRONTEND_GATES = ("barrier", "measure", "reset", "id", "x", "cx", "swap", "h", "z", "cz", "rz", "rx", "ry", "p",
                  "u1", "u2", "s", "sdg", "t", "tdg", "u", "u3", "rzz", "cp", "crz", "ccz", "ccx", "cswap")


# ---------------------------------------------------------------------------
# A minimal circuit with the attributes the loader reads
# ---------------------------------------------------------------------------
@dataclass
class Operation:
    name: str
    params: tuple = ()


@dataclass
class Instruction:
    operation: Operation
    qubits: tuple
    clbits: tuple = ()


@dataclass
class Circuit:
    num_qubits: int
    num_clbits: int = 0
    name: str = "circuit"
    data: List[Instruction] = field(default_factory=list)

    def append(self, name: str, qubits: Sequence[int], params: Sequence[float] = ()) -> "Circuit":
        for q in qubits:
            if not 0 <= int(q) < self.num_qubits:
                raise ValueError(f"qubit {q} outside a {self.num_qubits}-qubit circuit")
        self.data.append(Instruction(Operation(name.lower(), tuple(params)), tuple(int(q) for q in qubits)))
        return self

    def find_bit(self, qubit):
        return _Located(int(qubit))

    # builder shorthands (Qiskit's argument order)
    def h(self, q): return self.append("h", [q])
    def x(self, q): return self.append("x", [q])
    def z(self, q): return self.append("z", [q])
    def rz(self, theta, q): return self.append("rz", [q], [theta])
    def rx(self, theta, q): return self.append("rx", [q], [theta])
    def ry(self, theta, q): return self.append("ry", [q], [theta])
    def p(self, theta, q): return self.append("p", [q], [theta])
    def u(self, theta, phi, lam, q): return self.append("u", [q], [theta, phi, lam])
    def cx(self, c, t): return self.append("cx", [c, t])
    def cz(self, a, b): return self.append("cz", [a, b])
    def cp(self, theta, c, t): return self.append("cp", [c, t], [theta])
    def crz(self, theta, c, t): return self.append("crz", [c, t], [theta])
    def swap(self, a, b): return self.append("swap", [a, b])
    def ccx(self, a, b, c): return self.append("ccx", [a, b, c])
    def barrier(self): return self.append("barrier", list(range(self.num_qubits)))
    def measure_all(self): return self.append("measure", list(range(self.num_qubits)))


@dataclass
class _Located:
    index: int


def _qubit_index(circuit, qubit) -> int:
    if isinstance(qubit, int):
        return qubit
    return circuit.find_bit(qubit).index


def _to_float(value) -> float:
    try:
        return float(value)
    except Exception as exc:
        raise ValueError(f"Cannot evaluate symbolic parameter {value!r} to float.") from exc


def _name(op) -> str:
    return op.name.lower()


# ---------------------------------------------------------------------------
# Loader (reference: src/zxh/zxhsim/qasm.py:62-141)
# ---------------------------------------------------------------------------
def load_circuit_transpiled(sim, circuit) -> None:
    """Native gates only, forwarded to the simulator in circuit order."""
    unsupported = sorted({_name(i.operation) for i in circuit.data if _name(i.operation) not in NATIVE_GATES})
    if unsupported:
        raise NotImplementedError(f"Circuit contains gates outside the allowed set: {unsupported}. "
                                  f"allowed_gates={sorted(NATIVE_GATES)}")
    for inst in circuit.data:
        op = inst.operation
        name = _name(op)
        qs = [_qubit_index(circuit, q) for q in inst.qubits]
        if name == "barrier":
            sim.Barrier()
        elif name in ("measure", "id"):
            continue
        elif name == "reset":
            sim.clear_gates()
        elif name == "x":
            sim.X(qs[0])
        elif name == "cx":
            sim.CX(qs[0], qs[1])
        elif name == "h":
            sim.H(qs[0])
        elif name == "z":
            sim.Z(qs[0])
        elif name == "rz":
            sim.Rz(qs[0], _to_float(op.params[0]))
        elif name == "rx":
            sim.Rx(qs[0], _to_float(op.params[0]))
        elif name == "p":
            sim.P(qs[0], _to_float(op.params[0]))
        elif name in ("u3", "u"):
            theta, phi, lam = (_to_float(v) for v in op.params[:3])
            sim.U3(qs[0], theta, lam, phi)  # ZXH's U3 takes (theta, lambda, phi)
        elif name == "cp":
            sim.CP(qs[0], qs[1], _to_float(op.params[0]))
        elif name == "crz":
            sim.CRz(qs[0], qs[1], _to_float(op.params[0]))
        else:  # pragma: no cover - excluded by the check above
            raise NotImplementedError(f"Unsupported gate for ZXH native loading: {name}")


# ---------------------------------------------------------------------------
# Frontend lowering (exact native sequences)
# ---------------------------------------------------------------------------
_PHASE = {"s": math.pi / 2, "sdg": -math.pi / 2, "t": math.pi / 4, "tdg": -math.pi / 4}


def _ccz(a: int, b: int, c: int):
    """CCZ as 6 CX and 7 phase gates (the textbook T-count-7 circuit with T =
    P(pi/4)); exact, no global phase."""
    t, td = math.pi / 4, -math.pi / 4
    return [("cx", (b, c), ()), ("p", (c,), (td,)), ("cx", (a, c), ()), ("p", (c,), (t,)),
            ("cx", (b, c), ()), ("p", (c,), (td,)), ("cx", (a, c), ()), ("p", (b,), (t,)), ("p", (c,), (t,)),
            ("cx", (a, b), ()), ("p", (a,), (t,)), ("p", (b,), (td,)), ("cx", (a, b), ())]


def lower_instruction(name: str, qs: Sequence[int], params: Sequence[float]):
    """The native sequence (name, qubits, params) for one frontend gate."""
    if name in NATIVE_GATES:
        return [(name, tuple(qs), tuple(params))]
    if name == "u1":
        return [("p", (qs[0],), (params[0],))]
    if name in _PHASE:
        return [("p", (qs[0],), (_PHASE[name],))]
    if name == "u2":
        phi, lam = params
        return [("u", (qs[0],), (math.pi / 2, phi, lam))]
    if name == "ry":
        return [("u", (qs[0],), (params[0], 0.0, 0.0))]
    if name == "cz":
        return [("cp", (qs[0], qs[1]), (math.pi,))]
    if name == "swap":
        a, b = qs
        return [("cx", (a, b), ()), ("cx", (b, a), ()), ("cx", (a, b), ())]
    if name == "rzz":
        a, b = qs
        return [("cx", (a, b), ()), ("rz", (b,), (params[0],)), ("cx", (a, b), ())]
    if name == "ccz":
        return _ccz(*qs)
    if name == "ccx":
        a, b, c = qs
        return [("h", (c,), ())] + _ccz(a, b, c) + [("h", (c,), ())]
    if name == "cswap":
        c, a, b = qs
        return [("cx", (b, a), ())] + lower_instruction("ccx", (c, a, b), ()) + [("cx", (b, a), ())]
    raise NotImplementedError(f"Unsupported gate: {name}. allowed_gates={sorted(FRONTEND_GATES)}")


def lower_circuit(circuit) -> Circuit:
    """A native-gate Circuit equivalent to circuit (frontend gates lowered)."""
    out = Circuit(circuit.num_qubits, getattr(circuit, "num_clbits", 0), getattr(circuit, "name", "circuit"))
    for inst in circuit.data:
        name = _name(inst.operation)
        qs = [_qubit_index(circuit, q) for q in inst.qubits]
        params = [_to_float(v) for v in inst.operation.params] if name not in ("barrier", "measure") else []
        for n2, q2, p2 in lower_instruction(name, qs, params):
            out.append(n2, q2, p2)
    return out


# reference gate codes (gate_type_t order, defs.h:65-78; RX the recording alias)
GATE_CODES = {"z": 2, "rz": 3, "p": 4, "crz": 5, "cp": 6, "x": 7, "cx": 8, "h": 9, "u3": 10, "u": 10, "rx": 11}


def to_gate_tuples(circuit) -> list:
    """A native circuit as (code, cq, q, theta, lambda, phi) tuples, the
    form ZXH.gates() and the oracle take (U/U3 in ZXH's (theta, lambda, phi)
    order; barriers kept, measurements and id dropped)."""
    out = []
    for inst in circuit.data:
        name = _name(inst.operation)
        qs = [_qubit_index(circuit, q) for q in inst.qubits]
        ps = [_to_float(v) for v in inst.operation.params]
        if name == "barrier":
            out.append((1, 0, 0, 0.0, 0.0, 0.0))
        elif name in ("measure", "id"):
            continue
        elif name in ("u", "u3"):
            out.append((10, 0, qs[0], ps[0], ps[2], ps[1]))
        elif name in ("cx", "cp", "crz"):
            out.append((GATE_CODES[name], qs[0], qs[1], ps[0] if ps else 0.0, 0.0, 0.0))
        elif name in GATE_CODES:
            out.append((GATE_CODES[name], 0, qs[0], ps[0] if ps else 0.0, 0.0, 0.0))
        else:
            raise NotImplementedError(f"no gate tuple for {name} (lower the circuit first)")
    return out


def load_circuit(sim, circuit) -> None:
    """Lower the frontend gates, then load the native circuit."""
    load_circuit_transpiled(sim, lower_circuit(circuit))


# ---------------------------------------------------------------------------
# OpenQASM 2 reader
# ---------------------------------------------------------------------------
_TOKEN = re.compile(r"\s*(?:(\d+\.\d*(?:[eE][-+]?\d+)?|\.\d+(?:[eE][-+]?\d+)?|\d+(?:[eE][-+]?\d+)?)|(pi)|([-+*/()^]))")


def _eval_expr(text: str) -> float:
    """Arithmetic over numbers and pi: + - * / ^, unary minus, parentheses."""
    toks = []
    pos = 0
    text = text.strip()
    while pos < len(text):
        m = _TOKEN.match(text, pos)
        if not m or m.end() == pos:
            raise ValueError(f"bad parameter expression {text!r}")
        num, pi, op = m.groups()
        toks.append(("n", float(num)) if num else ("n", math.pi) if pi else ("o", op))
        pos = m.end()
    i = 0

    def peek():
        return toks[i] if i < len(toks) else (None, None)

    def atom():
        nonlocal i
        kind, v = peek()
        if kind == "n":
            i += 1
            return v
        if (kind, v) == ("o", "("):
            i += 1
            r = expr()
            if peek() != ("o", ")"):
                raise ValueError(f"unbalanced parentheses in {text!r}")
            i += 1
            return r
        if (kind, v) in (("o", "-"), ("o", "+")):
            i += 1
            r = power()
            return -r if v == "-" else r
        raise ValueError(f"bad parameter expression {text!r}")

    def power():
        nonlocal i
        base = atom()
        if peek() == ("o", "^"):
            i += 1
            return base ** power()
        return base

    def term():
        nonlocal i
        r = power()
        while peek() in (("o", "*"), ("o", "/")):
            op = peek()[1]
            i += 1
            r = r * power() if op == "*" else r / power()
        return r

    def expr():
        nonlocal i
        r = term()
        while peek() in (("o", "+"), ("o", "-")):
            op = peek()[1]
            i += 1
            r = r + term() if op == "+" else r - term()
        return r

    v = expr()
    if i != len(toks):
        raise ValueError(f"bad parameter expression {text!r}")
    return v


def parse_qasm(text: str) -> Circuit:
    """An OpenQASM 2 program (qelib1 gates) as a Circuit; qubits of several
    quantum registers are numbered in declaration order."""
    src = re.sub(r"//[^\n]*", "", text)
    regs = {}
    nq = 0
    ncl = 0
    stmts = [s.strip() for s in src.split(";") if s.strip()]
    body = []
    for s in stmts:
        if s.startswith("OPENQASM") or s.startswith("include"):
            continue
        m = re.fullmatch(r"qreg\s+(\w+)\s*\[\s*(\d+)\s*\]", s)
        if m:
            regs[m.group(1)] = (nq, int(m.group(2)))
            nq += int(m.group(2))
            continue
        m = re.fullmatch(r"creg\s+(\w+)\s*\[\s*(\d+)\s*\]", s)
        if m:
            ncl += int(m.group(2))
            continue
        body.append(s)
    circ = Circuit(nq, ncl)

    def args(text):
        out = []
        for a in [x.strip() for x in text.split(",") if x.strip()]:
            m = re.fullmatch(r"(\w+)\s*(?:\[\s*(\d+)\s*\])?", a)
            if not m or m.group(1) not in regs:
                raise ValueError(f"unknown quantum argument {a!r}")
            off, size = regs[m.group(1)]
            if m.group(2) is None:
                out.append([off + k for k in range(size)])
            else:
                k = int(m.group(2))
                if k >= size:
                    raise ValueError(f"index {k} outside register {m.group(1)}[{size}]")
                out.append([off + k])
        return out

    for s in body:
        if s.startswith("measure"):
            m = re.fullmatch(r"measure\s+(.+?)\s*->\s*(.+)", s)
            if not m:
                raise ValueError(f"bad measure statement {s!r}")
            qs = [q for group in args(m.group(1)) for q in group]
            circ.append("measure", qs)
            continue
        m = re.fullmatch(r"(\w+)\s*(?:\((.*)\))?\s+(.+)", s)
        if not m:
            raise ValueError(f"bad statement {s!r}")
        name = m.group(1).lower()
        params = [_eval_expr(p) for p in m.group(2).split(",")] if m.group(2) else []
        groups = args(m.group(3))
        if name == "barrier":
            circ.append("barrier", [q for g in groups for q in g])
            continue
        width = max(len(g) for g in groups)
        for k in range(width):  # whole registers broadcast
            circ.append(name, [g[k] if len(g) > 1 else g[0] for g in groups], params)
    return circ


def load_qasm(path) -> Circuit:
    p = Path(path)
    if not p.is_file():
        raise FileNotFoundError(f"QASM file not found: {p}")
    return parse_qasm(p.read_text())
