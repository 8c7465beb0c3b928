"""The C-ABI library loads and exports every symbol include/imhd.h declares (no GPU
calls), and the package refuses to run without the built library."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "isothermal_mhd_amd", "libimhd.so")


def declared():
    src = open(os.path.join(ROOT, "include", "imhd.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(imhd_\w+)\s*\(", src, re.M)))


def test_header_matches_binding_list():
    from isothermal_mhd_amd import _lib
    assert declared() == sorted(_lib.EXPORTS)


@pytest.mark.skipif(not os.path.exists(LIB), reason="libimhd.so not built")
def test_library_exports_all_symbols():
    L = ctypes.CDLL(LIB)
    for name in declared():
        assert hasattr(L, name), name
    L.imhd_version.restype = ctypes.c_char_p
    assert b"gfx950" in L.imhd_version()


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    from isothermal_mhd_amd import _lib
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "absent.so"))
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(_lib.ImhdError):
        _lib.lib()
