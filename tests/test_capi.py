"""The C-ABI library loads and exports every function include/*.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import glob
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(wl_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_library_exports_declared_symbols():
    import whirlaway_amd
    L = ctypes.CDLL(whirlaway_amd.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 15
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing


def test_python_layer_fails_loudly_without_library(tmp_path, monkeypatch):
    import whirlaway_amd
    monkeypatch.setattr(whirlaway_amd, "LIB_PATH", str(tmp_path / "absent.so"))
    monkeypatch.setattr(whirlaway_amd, "_lib", None)
    try:
        whirlaway_amd.lib()
    except whirlaway_amd.WhirlawayError as e:
        assert "missing" in str(e)
    else:
        raise AssertionError("expected WhirlawayError")
