"""The C-ABI library loads and exports every entry point include/*.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import re
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def declared_symbols():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(cb_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_symbols()
    assert "cb_grm_compute" in names and "cb_context_create" in names
    assert len(names) > 20


def test_library_exports_every_declared_symbol():
    from cosmicblup_amd import _lib

    L = ctypes.CDLL(str(_lib.LIB_PATH))
    missing = [n for n in declared_symbols() if not hasattr(L, n)]
    assert not missing, f"not exported: {missing}"


def test_python_binding_covers_header():
    from cosmicblup_amd import _lib

    assert set(declared_symbols()) == set(_lib.exported_names())


def test_last_error_without_device():
    from cosmicblup_amd import _lib

    L = _lib.lib()
    assert isinstance(L.cb_last_error(), bytes)
    # invalid arguments are rejected before any device work
    assert L.cb_grm_create(None, None, None) == _lib.CB_ERR_INVALID
    assert b"NULL" in L.cb_last_error()
