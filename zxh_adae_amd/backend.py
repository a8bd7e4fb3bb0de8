"""The backend callers use: the registry names and a backend with the
reference's ``run`` / job / result behaviour, on this engine.

Reference: ``make_backend`` (src/scripts/backend_registry.py:16-42) maps
"zxh", "zxh-nox" and "zxh-exp" to ``ZXHQiskitBackend(disable_x=...,
eager_expand_all=...)`` (src/zxh/zxhsim/qiskit_backend.py:217-330), whose
``run(circuits, shots=, seed_simulator=)`` returns a submitted job whose
result holds, per circuit, the counts and ``metadata["sample_time"]``.  This
module keeps that behaviour:
  - one cached simulator per backend; a circuit of another width is a
    ``JobError`` (qiskit_backend.py:256-271);
  - ``disable_x`` / ``eager_expand_all`` are construction parameters and are
    refused as run options;
  - without ``seed_simulator`` the simulator's seed is cleared, otherwise
    ``set_seed(int(seed))``; gates are cleared, the circuit loaded, executed,
    and ``sample_counts(shots)`` timed as ``sample_time``;
  - options default to shots=1, seed_simulator=None.
Circuits may be Qiskit ``QuantumCircuit`` objects or ``frontend.Circuit``;
frontend gates are lowered exactly (frontend.lower_circuit) instead of
through the reference's Qiskit optimiser.  Results are plain objects with the
attribute names of Qiskit's ``Result`` (``results[i].data.counts``,
``results[i].metadata``, ``get_counts()``), so callers that read counts work
without Qiskit installed.
"""
from __future__ import annotations

import secrets
import time
import uuid
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional

from . import frontend

SUPPORTED_BACKENDS = ("zxh", "zxh-nox", "zxh-exp")


class JobError(RuntimeError):
    pass


@dataclass
class ExperimentData:
    counts: Dict[str, int]


@dataclass
class ExperimentResult:
    shots: int
    success: bool
    data: ExperimentData
    metadata: dict
    seed: Optional[int] = None
    header: dict = field(default_factory=dict)


@dataclass
class Result:
    backend_name: str
    backend_version: str
    job_id: str
    qobj_id: str
    success: bool
    results: List[ExperimentResult]

    def get_counts(self, experiment: Optional[int] = None):
        if experiment is None:
            counts = [r.data.counts for r in self.results]
            return counts[0] if len(counts) == 1 else counts
        return self.results[experiment].data.counts


class ZXHJob:
    def __init__(self, backend: "ZXHBackend", circuits: list, options: dict) -> None:
        self._backend = backend
        self._circuits = circuits
        self._options = options
        self._result: Optional[Result] = None
        self._job_id = str(uuid.uuid4())
        self._qobj_id = str(uuid.uuid4())

    def backend(self) -> "ZXHBackend":
        return self._backend

    def job_id(self) -> str:
        return self._job_id

    def submit(self) -> None:
        if self._result is not None:
            raise JobError("Job already submitted")
        self._result = self._run_job()

    def status(self) -> str:
        return "DONE" if self._result is not None else "INITIALIZING"

    def result(self) -> Result:
        if self._result is None:
            raise JobError("Job not submitted")
        return self._result

    def _run_job(self) -> Result:
        self._backend._ensure_engine()
        shots = int(self._options.get("shots", 1))
        seed = self._options.get("seed_simulator")
        results = []
        for circuit in self._circuits:
            counts, sample_time = self._backend._run_circuit(circuit, shots=shots, seed_simulator=seed)
            results.append(ExperimentResult(shots=shots, success=True, data=ExperimentData(counts=counts),
                                            metadata={"sample_time": sample_time}, seed=seed,
                                            header={"name": getattr(circuit, "name", "circuit")}))
        return Result(backend_name=self._backend.name, backend_version=self._backend.backend_version,
                      job_id=self._job_id, qobj_id=self._qobj_id, success=True, results=results)


class ZXHBackend:
    """Backend over one cached ZXH simulator (see the module docstring)."""

    def __init__(self, *, disable_x: bool = False, eager_expand_all: bool = False, precision: str = "c128",
                 policy: str = "fused", sim_factory: Optional[Callable] = None) -> None:
        from . import _core
        self.name = "zxh"
        self.description = "ZXH backend (MI355X engine)"
        self.backend_version = getattr(_core, "__version__", "0")
        self._disable_x = bool(disable_x)
        self._eager_expand_all = bool(eager_expand_all)
        self._precision = precision
        self._policy = policy
        self._factory = sim_factory
        self._sim = None
        self.options = {"shots": 1, "seed_simulator": None}

    @property
    def operation_names(self) -> List[str]:
        return sorted(frontend.FRONTEND_GATES)

    @property
    def max_circuits(self):
        return None

    def set_options(self, **options) -> None:
        self.options.update(options)

    def _ensure_engine(self) -> None:
        if self._factory is not None:
            return
        from . import _core
        if not _core.active():
            _core.init()

    def _get_or_create_sim(self, num_qubits: int):
        if self._sim is None:
            if self._factory is not None:
                self._sim = self._factory(num_qubits, self._disable_x, self._eager_expand_all)
            else:
                from . import _core
                self._sim = _core.ZXH(num_qubits, self._disable_x, self._eager_expand_all,
                                      precision=self._precision, policy=self._policy)
            return self._sim
        if self._sim.num_qubits() != num_qubits:
            raise JobError("ZXH backend caches a single simulator instance. "
                           f"Expected num_qubits={self._sim.num_qubits()}, got num_qubits={num_qubits}.")
        return self._sim

    def _run_circuit(self, circuit, *, shots: int, seed_simulator):
        sim = self._get_or_create_sim(circuit.num_qubits)
        if seed_simulator is None:
            clear = getattr(sim, "clear_seed", None)
            if clear is not None:
                clear()
            else:
                sim.set_seed(secrets.randbits(64))
        else:
            sim.set_seed(int(seed_simulator))
        sim.clear_gates()
        frontend.load_circuit(sim, circuit)
        sim.execute()
        t0 = time.perf_counter()
        counts = {str(k): int(v) for k, v in sim.sample_counts(shots).items()}
        return counts, time.perf_counter() - t0

    def run(self, run_input, **options) -> ZXHJob:
        invalid = {"disable_x", "eager_expand_all"} & set(options)
        if invalid:
            raise JobError(f"{', '.join(sorted(invalid))} is a backend construction parameter for ZXH, "
                           "not a run() option.")
        job = ZXHJob(self, _coerce_circuits(run_input), {**self.options, **options})
        job.submit()
        return job


def _is_circuit(obj) -> bool:
    return hasattr(obj, "data") and hasattr(obj, "num_qubits")


def _coerce_circuits(run_input) -> list:
    if _is_circuit(run_input):
        return [run_input]
    try:
        items = list(run_input)
    except TypeError:
        items = None
    if items is not None and items and all(_is_circuit(c) for c in items):
        return items
    raise TypeError(f"Unsupported run_input type for ZXH backend: {type(run_input)!r}")


def make_zxh_backend(*, disable_x: bool = False, eager_expand_all: bool = False, **kw) -> ZXHBackend:
    return ZXHBackend(disable_x=disable_x, eager_expand_all=eager_expand_all, **kw)


def all_backends() -> List[str]:
    return list(SUPPORTED_BACKENDS)


def make_backend(backend_name: str, **kw) -> ZXHBackend:
    """The reference registry's ZXH names: "zxh" (lazy ZXH), "zxh-nox"
    (disable_x), "zxh-exp" (eager_expand_all, the plain method).  The
    registry's other simulators (cuQuantum, ddsim, qblaze) are not part of
    this engine."""
    if backend_name == "zxh":
        return make_zxh_backend(**kw)
    if backend_name == "zxh-nox":
        return make_zxh_backend(disable_x=True, **kw)
    if backend_name == "zxh-exp":
        return make_zxh_backend(eager_expand_all=True, **kw)
    raise ValueError(f"Unsupported backend: {backend_name}. supported={SUPPORTED_BACKENDS}")
