"""CPU tests of the drop-in library: it loads, exports every function that
include/*.h declares, keeps the reference's struct layouts, and its host
utilities (row tables, movedir, lookup, settings, files) agree with the
oracle.  No kernel is launched here."""
import ctypes as C
import os
import random
import re

import numpy as np
import pytest

from tests import oracle_lib as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def cg():
    import cablegen_amd
    cablegen_amd.lib()
    return cablegen_amd


def declared_functions(path):
    txt = open(path).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    txt = re.sub(r"#define[^\n]*", "", txt)
    names = set()
    for m in re.finditer(r"([A-Za-z_][A-Za-z0-9_]*)\s*\(([^;{}]*)\)\s*;", txt):
        if m.group(1) not in ("if", "while", "sizeof"):
            names.add(m.group(1))
    return names


@pytest.mark.parametrize("header", ["cablegen.h", "cablegen_gpu.h"])
def test_exports_every_declared_symbol(cg, header):
    names = declared_functions(os.path.join(ROOT, "include", header))
    assert len(names) > 10
    L = cg.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_reference_api_names_present(cg):
    """Every function of api/cablegen.h:41-79 is declared in include/cablegen.h."""
    want = {"movedir", "generate_lut", "read_boards", "read_boards2", "change_config", "lookup",
            "get_settings_min", "init_settings", "generate", "solve", "malloc_errcheck",
            "init_darr", "init_sarr", "shrink_darr", "concat", "destroy_darr", "destroy_sarr",
            "push_back", "write_boards", "read_table", "pun_uint64", "test", "get_sum",
            "output_board", "get_version"}
    assert want <= declared_functions(os.path.join(ROOT, "include", "cablegen.h"))


def test_struct_layouts(cg):
    assert C.sizeof(O.DArr) == 32 and C.sizeof(O.SArr) == 24 and C.sizeof(O.Table) == 48
    assert C.sizeof(cg.MinSettings) == 80


def test_self_test_and_version(cg):
    assert cg.lib().test()
    assert cg.lib().get_version()


def test_row_tables_match_oracle(cg):
    for ig in (False, True):
        mv, _ = O.lut(False, ig)
        assert np.array_equal(cg.row_table(ig), mv[0])


def test_host_movedir_and_sum(cg, tmp_path):
    L = cg.lib()
    rng = random.Random(2)
    for ig in (False, True):
        with open(tmp_path / "c.conf", "w") as f:
            f.write(f"[Cablegen]\nignore_f = {'true' if ig else 'false'}\n")
        L.change_config(str(tmp_path / "c.conf").encode())
        assert L.get_settings_min().contents.ignore_f == ig
        cfg = O.make_cfg(ignore_f=ig)
        for _ in range(3000):
            fm = rng.choice([0x00000000FFFFFFFF, 0x00000000FFF0FFF0, 0])
            b = 0
            for x in range(16):
                sh = 60 - 4 * x
                b |= (0xf if (fm >> sh) & 0xf else (rng.randint(1, 9) if rng.random() > 0.4 else 0)) << sh
            assert L.get_sum(b) == O.orc().orc_get_sum(b)
            for d in range(4):
                x = C.c_uint64(b)
                ok = L.movedir(C.byref(x), d)
                ok_o, y = O.move(cfg, b, d)
                assert ok == ok_o
                if ok:
                    # movedir rotates up/down back; the oracle's stays rotated
                    assert O.canon(x.value) == O.canon(y)
                    if d < 2:
                        assert x.value == y


def test_host_lookup_known_answer(cg):
    L = cg.lib()
    k = np.arange(100, dtype=np.uint64)
    v = (99 - np.arange(100)).astype(np.float64)
    t = cg.Table(cg.SArr(True, k.ctypes.data_as(cg.U64P), 100),
                 cg.SArr(True, v.view(np.uint64).ctypes.data_as(cg.U64P), 100))
    for i in range(100):
        assert L.lookup(i, C.byref(t), False) == 99 - i
    assert L.lookup(100, C.byref(t), False) == 0.0
    e = cg.Table(cg.SArr(True, None, 0), cg.SArr(True, None, 0))
    assert L.lookup(5, C.byref(e), True) == 0.0


def test_settings_defaults_and_keys(cg, tmp_path):
    L = cg.lib()
    with open(tmp_path / "c.conf", "w") as f:
        f.write("; comment\n[CABLEGEN]\nCores = 7\nnox = 9\nfree_formation = TRUE\n"
                "[Generate]\nend = 900\ndir = /x/b/\n[Solve]\nend = 12\nwinstates = ./w\n")
    L.change_config(str(tmp_path / "c.conf").encode())
    s = L.get_settings_min().contents
    assert s.cores == 7 and s.nox == 9 and s.free_formation and s.end_gen == 900
    assert s.bdir == b"/x/b/" and s.end_solve == 12 and s.winstates == b"./w"
    assert s.tdir == b"./tables/" and s.initial == b"./initial"   # defaults kept
    r = cg.Rules()
    L.cg_rules_from_settings(C.byref(r))
    assert r.free_formation == 1 and r.nox == 9 and r.stsl == 200 and r.smallest_large == 6
    assert r.premove == 0
    with open(tmp_path / "p.conf", "w") as f:
        f.write("[Generate]\npremove = true\n")
    L.change_config(str(tmp_path / "p.conf").encode())
    L.cg_rules_from_settings(C.byref(r))
    assert r.premove == 1


def test_files_round_trip(cg, tmp_path):
    L = cg.lib()
    a = np.array([3, 1, 4, 1, 5, 9, 2, 6], dtype=np.uint64)
    sa = cg.SArr(True, a.ctypes.data_as(cg.U64P), len(a))
    L.write_boards.argtypes = [cg.SArr, C.c_char_p, C.c_int]
    L.write_boards(sa, str(tmp_path / "%d.boards").encode(), 42)
    L.read_boards.restype = cg.SArr
    L.read_boards.argtypes = [C.c_char_p]
    r = L.read_boards(str(tmp_path / "42.boards").encode())
    assert r.valid and r.size == 8
    assert np.array_equal(np.ctypeslib.as_array(r.bp, shape=(8,)), a)
    r = L.read_boards(str(tmp_path / "missing.boards").encode())
    assert not r.valid and r.size == 0
    k = np.arange(5, dtype=np.uint64)
    v = np.linspace(0, 1, 5)
    with open(tmp_path / "t.tables", "wb") as f:
        f.write(k.tobytes() + v.tobytes())
    t = cg.Table()
    L.read_table.argtypes = [C.POINTER(cg.Table), C.c_char_p]
    L.read_table(C.byref(t), str(tmp_path / "t.tables").encode())
    assert t.key.size == 5
    assert np.array_equal(np.ctypeslib.as_array(t.value.bp, shape=(5,)).view(np.float64), v)


def test_arrays(cg):
    L = cg.lib()
    L.init_darr.restype = O.DArr
    L.init_darr.argtypes = [C.c_bool, C.c_size_t]
    L.push_back.restype = C.c_bool
    L.push_back.argtypes = [C.POINTER(O.DArr), C.c_uint64]
    L.shrink_darr.restype = O.SArr
    L.shrink_darr.argtypes = [C.POINTER(O.DArr)]
    d = L.init_darr(False, 0)
    for i in range(1000):
        L.push_back(C.byref(d), i)
    s = L.shrink_darr(C.byref(d))
    assert s.valid and s.size == 1000
    assert np.array_equal(np.ctypeslib.as_array(s.bp, shape=(1000,)), np.arange(1000))


def test_stored_tuning_table_installs(cg):
    """The autotuner's stored MI355X table names only configurations that
    have kernels, installs through cg_tune_sort_entry and is what the sorts
    then choose (table lookups only; nothing runs on a device)."""
    import json
    if not os.path.exists(cg.TUNED_TABLE):
        pytest.skip("no stored table")
    rep = json.load(open(cg.TUNED_TABLE))
    kernels = {"8x32/16", "8x32/32", "8x16/16", "8x16/32", "8x8/16", "8x8/32", "9x16/16", "9x16/32",
               "9x32/16", "9x32/32", "9x7/16", "9x7/32", "10x8/16", "10x8/32"}
    assert len(rep["sort"]) == 3
    for c in rep["sort"]:
        assert set(c["pick"]) == {str(b) for b in range(1, 65)}
        assert set(c["pick"].values()) <= kernels
    try:
        cg.apply_tuning(rep)
        for cls, n in enumerate((1 << 18, 1 << 21, 1 << 25)):
            for nb in (1, 24, 40, 54, 64):
                b, k, w = cg.tune_sort_query(n, nb)
                assert rep["sort"][cls]["pick"][str(nb)] == "%dx%d/%d" % (b, k, w)
        assert cg.tune_get("dir_delta") == rep["dir"]["pick"]
        # forcing a field overrides the table; reset restores the defaults
        cg.tune_set("sort_bits", 10)
        assert cg.tune_sort_query(1 << 25, 40)[0] == 10
        cg.tune_set("builtin", 0)
        assert cg.tune_sort_query(1 << 25, 40) == (8, 32, 16)
        with pytest.raises(cg.CablegenError):
            cg.tune_set("sort_bits", 7)
        # the C loader reads the same file and "reset" returns to it
        assert cg.lib().cg_tune_load(cg.TUNED_TABLE.encode()) == 0
        for cls, n in enumerate((1 << 18, 1 << 21, 1 << 25)):
            b, k, w = cg.tune_sort_query(n, 48)
            assert rep["sort"][cls]["pick"]["48"] == "%dx%d/%d" % (b, k, w)
        assert cg.lib().cg_tune_load(os.path.join(ROOT, "bench.py").encode()) != 0
    finally:
        cg.tune_set("reset", 0)



def test_layer_index_refuses_2e32_boards(cg):
    """The directory, and the positions the solve and the adjoint derive from
    it, are 32-bit.  A layer of 2^32 boards or more must fail loudly (the
    check every layer build makes), not wrap; one below the limit plans a
    directory of at most 30 bits.  (Host-side planning only.)"""
    L = cg.lib()
    bits = C.c_int(-1)
    key_or, key_and = 0xFFFAFFFBFFFFFFFF, 0xFFF0FFF000000000   # a DPDGap layer's 40 varying bits
    assert L.cg_layer_index_probe((1 << 32) - 1, key_or, key_and, C.byref(bits)) == 0
    assert bits.value == 30
    for n in (1 << 32, (1 << 32) + 1, 1 << 40):
        bits.value = -1
        assert L.cg_layer_index_probe(n, key_or, key_and, C.byref(bits)) != 0
        assert bits.value == -1
        assert "2^32" in cg.last_error()
    assert L.cg_layer_index_probe(1000, key_or, key_and, C.byref(bits)) == 0 and 0 < bits.value <= 12
