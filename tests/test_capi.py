"""The C ABI library loads and exports every entry point declared in
include/*.h (no GPU work is launched here)."""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names |= set(re.findall(r"^\s*(?:int32_t|int64_t|const char\*|void)\s+(abc_\w+)\s*\(", src, re.M))
    return names


def test_header_declares_entry_points():
    names = declared()
    for must in ("abc_create", "abc_destroy", "abc_sample_prior", "abc_simulate", "abc_errors",
                 "abc_accept", "abc_last_error"):
        assert must in names


def test_library_exports_all_declared_symbols():
    from abc_inference_transcription_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libabc.so not built")
    L = _lib.load()
    missing = [n for n in sorted(declared()) if not hasattr(L, n)]
    assert not missing, missing


def test_missing_library_fails_loudly(tmp_path, monkeypatch):
    from abc_inference_transcription_amd import _lib
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "absent.so"))
    with pytest.raises(RuntimeError):
        _lib.load()
