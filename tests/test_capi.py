"""The C-ABI library loads and exports every symbol include/*.h declares
(no compute calls: runs without a GPU)."""
import ctypes
import glob
import os
import re

import pytest

from numpyro_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        txt = open(h).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        names |= set(re.findall(r"\b(nax_[a-z0-9_]+)\s*\(", txt))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_symbols()
    assert "nax_create" in names and "nax_run" in names and len(names) >= 15


@pytest.mark.parametrize("name", declared_symbols())
def test_library_exports(name):
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libnumpyro_amd.so not built")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    assert hasattr(lib, name)


def test_python_binding_covers_header():
    assert set(declared_symbols()) <= set(_lib.SIGNATURES)


def test_no_cpu_fallback_when_library_missing(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        _lib.load()
