"""The C-ABI library loads and exports every entry point of include/schemi_gpu.h."""
import os
import re

from schemi_amd import gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "schemi_gpu.h")).read()
    return sorted(set(re.findall(r"\b(schemi_gpu_\w+)\s*\(", txt)))


def test_header_symbols_exported():
    lib = gpu.library()
    names = declared_symbols()
    assert len(names) >= 14
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(gpu.exported_symbols())


def test_strerror_codes():
    lib = gpu.library()
    assert lib.schemi_gpu_strerror(0) == b"noErrors"
    assert lib.schemi_gpu_strerror(1).startswith(b"RiemannSolverError")
    assert lib.schemi_gpu_strerror(11).startswith(b"initialisationError")
