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


def test_dense_split_lookup_follows_tuned_table():
    """cb_tune_dense_lookup returns the first entry of the generated table
    (csrc/cb_dense_tuned.h) whose element size matches and whose q_max and
    r_max cover the shape, else the built-in 2048 blocks / 256 k."""
    import re

    from cosmicblup_amd import tune

    rows = [tuple(int(v) for v in m.groups()) for m in
            re.finditer(r"\{(\d+), (\d+)LL, (\d+), (\d+), (\d+)\}", tune.TABLE.read_text())]
    assert rows, "no tuned entries"

    def expect(elem, q, r):
        for e, q_max, r_max, tb, mk in rows:
            if e == elem and q <= q_max and min(r, 80) <= r_max:
                return tb, mk
        return tune.DEFAULT

    for elem in (2, 4, 8):
        for q in (64, 599, 600, 849, 850, 3000, 19999, 20000, 28284, 28285, 10 ** 6):
            for r in (3, 8, 9, 24, 25, 48, 49, 65, 80, 200):
                assert tune.lookup(elem, q, r) == expect(elem, q, r), (elem, q, r)
    assert tune.lookup(16, 1000, 10) == tune.DEFAULT


def test_tuned_table_rows_round_trip():
    """The tuner re-reads the rows it keeps (`--elems` tunes a subset) and
    writes them back unchanged."""
    from cosmicblup_amd import tune

    rows = tune.read_table()
    assert {e["elem"] for e in rows} == {2, 4, 8}
    old = [l for l in tune.TABLE.read_text().splitlines() if l.startswith("    {")]
    new = [l for l in tune.render(rows, "device").splitlines() if l.startswith("    {")]
    assert new == old
