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


@pytest.fixture(params=[4, 5, 6], ids=lambda s: f"seg{s}")
def tile_seg(request, gpu_engine):
    """Runs a test under each tile geometry (knobs tile.seg.c128 and
    tile.seg.c64: 4, 5 or 6 contiguous unit-address bits per tile) and
    restores the knobs."""
    from zxh_adae_amd import autotune
    knobs = ("tile.seg.c128", "tile.seg.c64")
    old = [autotune.tune_get(k) for k in knobs]
    for k in knobs:
        autotune.tune_set(k, request.param)
    yield request.param
    for k, v in zip(knobs, old):
        autotune.tune_set(k, v)
