"""The registry backends end to end on the GPU: the reference's integration
cases (QASM input fixtures, tests/golden/integration_cases) through
make_backend("zxh" | "zxh-nox" | "zxh-exp") with the manifest's oracles and
seed reproducibility, and frontend-gate circuits against the oracle's state."""
import json
import math
import os

import numpy as np
import pytest

import oracle_py as O
from zxh_adae_amd import backend as B
from zxh_adae_amd import frontend as F

pytestmark = pytest.mark.gpu

CASES = os.path.join(os.path.dirname(__file__), "golden", "integration_cases")
EXP = json.load(open(os.path.join(CASES, "expectations.json")))


@pytest.mark.parametrize("name", B.all_backends())
@pytest.mark.parametrize("case", sorted(EXP["cases"]))
def test_integration_case_oracles(gpu_engine, name, case):
    spec = EXP["cases"][case]
    circ = F.load_qasm(os.path.join(CASES, case + ".qasm"))
    shots = EXP["shots"]
    counts = B.make_backend(name).run(circ, shots=shots, seed_simulator=1234).result().get_counts()
    assert sum(counts.values()) == shots
    assert all(len(k) == circ.num_qubits for k in counts)
    if spec["oracle"] == "delta":
        assert counts == {spec["expected"]: shots}
    elif spec["oracle"] == "support":
        assert set(counts) <= set(spec["allowed"])
    else:
        for key, p in spec["expected"].items():
            assert abs(counts.get(key, 0) / shots - p) <= spec["tolerance_abs"]
    rep = spec.get("seed_repro")
    if rep:
        be = B.make_backend(name)
        runs = [be.run(circ, shots=rep["shots"], seed_simulator=rep["seed"]).result().get_counts()
                for _ in range(rep["repeats"])]
        assert all(r == runs[0] for r in runs)


def test_frontend_gates_match_the_oracle_state(gpu_engine):
    """A circuit over the frontend set, lowered and run on the engine
    (plain method), against the oracle running the same native sequence."""
    rng = np.random.default_rng(3)
    n = 6
    c = F.Circuit(n)
    for q in range(n):
        c.h(q)
    for _ in range(40):
        kind = rng.integers(9)
        a, b, d = (int(v) for v in rng.choice(n, 3, replace=False))
        t = float(rng.uniform(-math.pi, math.pi))
        [lambda: c.ry(t, a), lambda: c.cz(a, b), lambda: c.swap(a, b), lambda: c.ccx(a, b, d),
         lambda: c.append("rzz", [a, b], [t]), lambda: c.append("u2", [a], [t, -t / 2]),
         lambda: c.append("t", [a]), lambda: c.append("sdg", [b]), lambda: c.append("ccz", [a, b, d])][kind]()
    be = B.make_backend("zxh-exp")
    be.run(c, shots=16, seed_simulator=5)
    got = be._sim.get_state_array()
    lowered = F.lower_circuit(c)
    o = O.OracleSim(n, False, True)
    rec = []
    for inst in lowered.data:
        rec.append((inst.operation.name, inst.qubits, inst.operation.params))
    gs = []
    for name, qs, ps in rec:
        if name == "h":
            gs.append((O.H, 0, qs[0], 0.0, 0.0, 0.0))
        elif name == "cx":
            gs.append((O.CX, qs[0], qs[1], 0.0, 0.0, 0.0))
        elif name == "p":
            gs.append((O.P, 0, qs[0], ps[0], 0.0, 0.0))
        elif name == "rz":
            gs.append((O.RZ, 0, qs[0], ps[0], 0.0, 0.0))
        elif name == "u":
            gs.append((O.U3, 0, qs[0], ps[0], ps[2], ps[1]))
        elif name == "cp":
            gs.append((O.CP, qs[0], qs[1], ps[0], 0.0, 0.0))
        else:
            raise AssertionError(name)
    o.gates(gs)
    o.execute()
    assert np.max(np.abs(got - o.get_state())) < 1e-12
