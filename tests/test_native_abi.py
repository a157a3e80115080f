"""The C-ABI library loads without a GPU and exports every symbol include/fne_hip.h declares."""

import ctypes
import re
from pathlib import Path

from fenics_nonlinear_energies_amd import native

HEADER = Path(__file__).resolve().parents[1] / "include" / "fne_hip.h"


def declared_symbols():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^(?:int|const char\*)\s+(fne_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_header_symbols():
    lib = native.lib()
    names = declared_symbols()
    assert names, "no declarations parsed from the header"
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) <= set(native.exported_symbols())


def test_version_and_error_text():
    lib = native.lib()
    assert b"gfx950" in lib.fne_version()
    # a NULL descriptor is rejected without touching the GPU
    out = ctypes.c_void_p()
    assert lib.fne_plan_create(None, ctypes.byref(out)) != 0


def test_descriptor_layout_matches_header():
    # offsets of a few fields the kernels read (x86-64 SysV layout)
    d = native.FnePlanDesc
    assert d.nnz.offset == 24 and d.edof.offset == 32
    assert d.param.offset % 8 == 0 and ctypes.sizeof(d) % 8 == 0
