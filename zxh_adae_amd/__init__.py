"""zxh_adae_amd -- MI355X-native state-vector engine for ZXH-Sim's hot path.

The package mirrors the reference's Python-facing surface for this path:

* ``zxh_adae_amd._core``: the ``_core`` module (``init``, ``ZXH`` ...), over the
  C ABI of ``lib/libzxh_amd.so`` (HIP kernels for gfx950);
* ``zxh_adae_amd.backend``: the registry names (``make_backend("zxh" |
  "zxh-nox" | "zxh-exp")``) and a backend with the reference's ``run`` /
  job / result behaviour; ``zxh_adae_amd.frontend``: the native-gate loader,
  frontend-gate lowering and an OpenQASM 2 reader.

There is no CPU fallback: without the native library or a GPU, calls raise.
"""
from . import _core
from ._core import ZXH, ZXHError, PrintRes, __version__, active, finalize, init, nprocs, rank

__all__ = ["_core", "ZXH", "ZXHError", "PrintRes", "init", "finalize", "active", "rank", "nprocs"]
