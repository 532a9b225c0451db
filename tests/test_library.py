"""The C-ABI library loads and exports every function include/*.h declares."""
import glob
import os
import re

import wavetrisk_hydrostatic_amd as W
from conftest import ROOT


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:int|void|double)\s+\**\s*(wtg_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_header_declares_api():
    names = declared_functions()
    assert {"wtg_create", "wtg_trend", "wtg_rk4", "wtg_ghost_fill", "wtg_grid_create"} <= names


def test_library_exports_every_declared_symbol():
    lib = W.load_library()
    missing = [n for n in sorted(declared_functions()) if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"


def test_binding_covers_header():
    assert declared_functions() == set(W.SIGNATURES)


def test_errors_are_reported_without_gpu_calls():
    import pytest
    with pytest.raises(W.WtgError, match="level"):
        W.Grid(level=3, min_level=3, domain_level=1)  # n=8 is not a multiple of the 16-node tile
