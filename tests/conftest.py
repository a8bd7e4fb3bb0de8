import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built native library")
    # the oracle is C; build it if it is not there (gcc exists here and on the GPU box)
    so = os.path.join(ROOT, "oracle", "liboracle.so")
    src = os.path.join(ROOT, "oracle", "zxh_oracle.c")
    if not os.path.exists(so) or os.path.getmtime(so) < os.path.getmtime(src):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "all"], check=True, stdout=subprocess.DEVNULL)


def _gpu_available() -> bool:
    try:
        import torch  # noqa: F401  (only to count devices; not used by the engine)
        return torch.cuda.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu_engine():
    """The native engine, initialised on the GPU (fails loudly if missing)."""
    import zxh_adae_amd as Z
    Z.init()
    return Z


@pytest.fixture(params=[(3, 2), (4, 2), (5, 2), (6, 2), (3, 3), (6, 3)], ids=lambda p: f"seg{p[0]}-bufs{p[1]}")
def tile_seg(request, gpu_engine):
    """Runs a test under each tile geometry (knobs tile.seg.<prec>: 3, 4, 5
    or 6 contiguous unit-address bits per tile) and with two or three register
    buffers per tile-pass workgroup (knobs tile.bufs.<prec>); restores the
    knobs.  Yields the geometry."""
    from zxh_adae_amd import autotune
    seg, bufs = request.param
    knobs = {"tile.seg.c128": seg, "tile.seg.c64": seg, "tile.bufs.c128": bufs, "tile.bufs.c64": bufs}
    old = {k: autotune.tune_get(k) for k in knobs}
    for k, v in knobs.items():
        autotune.tune_set(k, v)
    yield seg
    for k, v in old.items():
        autotune.tune_set(k, v)
