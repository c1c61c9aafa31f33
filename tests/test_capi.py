"""The C-ABI library loads and exports every symbol include/matjes_hip.h declares (CPU; no compute)."""
import ctypes
import os
import re

from conftest import REPO


def _declared():
    text = open(os.path.join(REPO, "include", "matjes_hip.h")).read()
    return sorted(set(re.findall(r"^int\s+(mj_\w+)\s*\(", text, re.M)))


def _lib_path():
    from matjes_amd import _lib, build
    if not os.path.isfile(_lib.LIB_PATH):
        build.build_hip()
    return _lib.LIB_PATH


def test_library_exports_header_symbols():
    lib = ctypes.CDLL(_lib_path())
    names = _declared()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n


def test_python_binding_covers_header():
    from matjes_amd import _lib
    assert sorted(_lib.EXPORTS) == _declared()


def test_binding_fails_loudly_without_library(tmp_path, monkeypatch):
    from matjes_amd import _lib
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(_lib, "_lib", None)
    try:
        _lib.load()
    except _lib.MatjesError:
        return
    raise AssertionError("loading a missing library must raise")
