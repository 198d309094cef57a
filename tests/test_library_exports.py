"""The C-ABI libraries load on a machine without a GPU and export every
function their headers declare (no compute calls)."""

import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header: str) -> list[str]:
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(\w+)\s*\(", text, flags=re.M)


def test_hip_library_exports_every_declared_symbol():
    from bmnd_amd import _lib
    L = _lib.lib()
    names = _declared(os.path.join(ROOT, "include", "bmnd_hip.h"))
    assert len(names) >= 10
    for name in names:
        assert hasattr(L, name), name
    conv = _declared(os.path.join(ROOT, "include", "bmnd_conv.h"))
    assert len(conv) >= 15
    for name in conv:
        assert hasattr(L, name), name
    assert set(_lib.EXPORTS) <= set(names) | set(conv)
    assert L.bmnd_hip_version() >= 100
    assert b"block=(4,4,4)" in L.bmnd_hip_configs()
    assert L.bmnd_conv_kpad(16) == 448 and L.bmnd_conv_kpad(96) == 2592


def test_oracle_library_exports_every_declared_symbol():
    import oracle
    L = oracle.lib()
    for name in _declared(os.path.join(ROOT, "oracle", "bmnd_oracle.h")):
        assert hasattr(L, name), name


def test_library_is_built_for_gfx950():
    from bmnd_amd import _lib
    data = open(_lib.LIB_PATH, "rb").read()
    assert b"gfx950" in data


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    from bmnd_amd import _lib
    monkeypatch.setattr(_lib, "_LIB", None)
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "absent.so"))
    with pytest.raises(_lib.HipError, match="not built"):
        _lib.lib()
