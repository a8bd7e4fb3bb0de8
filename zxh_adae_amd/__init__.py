"""zxh_adae_amd -- MI355X-native state-vector engine for ZXH-Sim's hot path.

The package mirrors the reference's Python-facing surface for this path:

* ``zxh_adae_amd._core``: the ``_core`` module (``init``, ``ZXH`` ...), over the
  C ABI of ``lib/libzxh_amd.so`` (HIP kernels for gfx950);
* ``zxh_adae_amd.backend_registry.make_backend``: the registry's ``run`` /
  ``sample`` entry.

There is no CPU fallback: without the native library or a GPU, calls raise.
"""
from . import _core
from ._core import ZXH, ZXHError, PrintRes, active, finalize, init, nprocs, rank

__all__ = ["_core", "ZXH", "ZXHError", "PrintRes", "init", "finalize", "active", "rank", "nprocs"]
__version__ = "0.1.0"
