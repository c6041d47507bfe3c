"""The C-ABI library loads and exports every symbol include/nmgpu.h declares
(no device calls)."""
import ctypes
import re
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def test_exports_every_declared_symbol():
    from normalmodes_amd._native import load_gpu
    L = load_gpu()
    hdr = (ROOT / "include" / "nmgpu.h").read_text()
    names = set(re.findall(r"\b(nmg_[a-z0-9_]+)\s*\(", hdr))
    assert len(names) > 30
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing


def test_host_only_entries_without_gpu():
    from normalmodes_amd._native import load_gpu
    L = load_gpu()
    L.nmg_last_error.restype = ctypes.c_char_p
    a, b = ctypes.c_int(), ctypes.c_int()
    assert L.nmg_version(ctypes.byref(a), ctypes.byref(b)) == 0
    # invalid handle is rejected, not dereferenced
    assert L.nmg_freepol(ctypes.c_int64(42)) != 0
    assert b"invalid handle" in L.nmg_last_error()
