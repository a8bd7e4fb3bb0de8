"""The C-ABI library loads and exports every function include/zxh_amd.h
declares (no compute call: this runs without a GPU)."""
import ctypes
import os
import re

import pytest

from conftest import ROOT
from zxh_adae_amd import _lib

HEADER = os.path.join(ROOT, "include", "zxh_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(zxh_\w+)\s*\(", text, flags=re.M)))


def test_header_parses():
    names = declared_functions()
    assert "zxh_create" in names and "zxh_k_sample" in names and len(names) > 40


@pytest.mark.skipif(not os.path.exists(_lib.LIB_PATH), reason="native library not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(_lib.LIB_PATH), reason="native library not built")
def test_python_signatures_cover_header():
    assert set(declared_functions()) == set(_lib.EXPORTED)


@pytest.mark.skipif(not os.path.exists(_lib.LIB_PATH), reason="native library not built")
def test_tuning_knobs_ranges():
    """The knobs the autotuner sweeps exist, accept their documented values,
    reject the others, and are host state only (no device needed)."""
    from zxh_adae_amd import autotune
    from zxh_adae_amd import ZXHError
    cases = {"diag.ept.c128": ([0, 1, 2, 4], [3, 5, -1]), "diag.ept.c64": ([0, 1, 2, 4], [3, 8]),
             "diag.nt.c128": ([0, 1], [2]), "diag.nt.c64": ([0, 1], [2]),
             "tile.seg.c128": ([3, 4, 5, 6], [2, 7]), "tile.seg.c64": ([3, 4, 5, 6], [0, 12]),
             "tile.bufs.c128": ([2, 3], [1, 4]), "tile.bufs.c64": ([2, 3], [0, 5]),
             "tile.wide.c128": ([0, 1], [2, -1]), "tile.wide.c64": ([0, 1], [3]),
             "tile.upper.c128": ([0, 18, 64], [-1, 65]), "tile.upper.c64": ([0, 17, 64], [-2, 100]),
             "block.variant.c128": ([0, 3, 4, 7], [-1, 8]), "block.variant.c64": ([0, 5, 7], [8, 9]),
             "block.tiles.c128": ([0, 1, 4, 4096], [-1, 4097]), "block.tiles.c64": ([0, 8], [5000])}
    for knob, (good, bad) in cases.items():
        old = autotune.tune_get(knob)
        for v in good:
            autotune.tune_set(knob, v)
            assert autotune.tune_get(knob) == v
        for v in bad:
            with pytest.raises(ZXHError):
                autotune.tune_set(knob, v)
        autotune.tune_set(knob, old)
    with pytest.raises(ZXHError):
        autotune.tune_set("no.such.knob", 1)
