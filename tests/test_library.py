"""The C-ABI library loads and exports every symbol include/csg.h declares.

CPU-only: no compute call is made (there is no GPU here), but the library is
loaded with ctypes, so a missing or mis-named export fails this test.
"""

from __future__ import annotations

import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
LIB = ROOT / "cat-surface-gpu_amd" / "cat_surface_gpu" / "libcsg_hip.so"


def _declared() -> list[str]:
    names = []
    for header in sorted((ROOT / "include").glob("*.h")):
        text = header.read_text()
        names += re.findall(r"^\s*(?:const\s+)?[a-z_]+\s*\*?\s*(csg_[a-z0-9_]+)\s*\(", text,
                            flags=re.M)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = _declared()
    assert "csg_solve_dartel_maps" in names
    assert "csg_last_error" in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol():
    if not LIB.is_file():
        pytest.fail(f"{LIB} is not built; run __graft_entry__.build()")
    handle = ctypes.CDLL(str(LIB))
    missing = [name for name in _declared() if not hasattr(handle, name)]
    assert not missing, missing


def test_python_signatures_cover_the_header():
    import sys

    sys.path.insert(0, str(ROOT / "cat-surface-gpu_amd"))
    from cat_surface_gpu import _lib

    assert set(_declared()) <= set(_lib.exported_symbols())
    handle = _lib.lib()
    assert handle.csg_version().decode().startswith("csg-hip")
    # size queries are host-only and safe without a device
    assert handle.csg_solve_workspace_bytes(512, 256, 2, 1) > 100 * 512 * 256 * 8


def test_errors_map_to_reference_exception_types():
    import sys

    sys.path.insert(0, str(ROOT / "cat-surface-gpu_amd"))
    from cat_surface_gpu import _lib

    handle = _lib.lib()
    rc = handle.csg_resize_field(None, 0, 4, 4, 2, 2, 0, None, None)  # channels < 1: no launch
    assert rc == -1
    with pytest.raises(ValueError):
        _lib.check(rc, "csg_resize_field")
    rc = handle.csg_dartel_step(None, None, None, None, 1, _lib.doubles([1, 1, 0.1, 0, 0]), 1e-3,
                                3, 3, 7, 1, 8, 8, 1, 1, None, None, None, 0, None)
    assert rc == -3
    with pytest.raises(NotImplementedError):
        _lib.check(rc, "csg_dartel_step")
