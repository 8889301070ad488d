"""CPU checks of the boundary: the library loads and exports every symbol the
public header declares, and the ctypes table matches the header."""
import os
import re

from cellweave_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(cw_[a-z0-9_]+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    lib = _abi.lib()
    names = declared_symbols()
    assert len(names) >= 12
    for name in sorted(names):
        assert hasattr(lib, name), name


def test_ctypes_table_covers_the_header():
    assert declared_symbols() == set(_abi.EXPORTS)


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(_abi, "LIB_PATH", str(tmp_path / "absent.so"))
    monkeypatch.setattr(_abi, "_lib", None)
    import pytest

    with pytest.raises(OSError, match="missing"):
        _abi.lib()
