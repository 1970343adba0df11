"""Parity of the HIP path against the oracle, through the C ABI (GPU tests).

Integer work (layers) must be bit-identical; FP64 values must be
bit-identical too (same summation order, no contraction), which is stricter
than the 1e-9 tolerance the values are specified with.  The adjoint is
checked against the oracle's FP64 gradient (1e-12 relative) and central
differences of GPU re-solves (1e-6 relative).
"""
import ctypes as C
import json
import os
import random

import numpy as np
import pytest

from tests import oracle_lib as O

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def cg():
    import cablegen_amd
    cablegen_amd.lib()
    assert torch.cuda.is_available(), "GPU tests need an MI355X"
    return cablegen_amd


def dev(a):
    """uint64 numpy -> device int64 tensor (plumbing only)."""
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.uint64).view(np.int64)).cuda()


def host(t, n=None):
    a = t.cpu().numpy().view(np.uint64)
    return a if n is None else a[:n]


def rand_boards(rng, fmask, n, p0=0.4, kmax=9):
    out = np.zeros(n, dtype=np.uint64)
    for x in range(16):
        sh = 60 - 4 * x
        if (fmask >> sh) & 0xf:
            out |= np.uint64(0xf << sh)
        else:
            v = rng.integers(1, kmax + 1, size=n).astype(np.uint64)
            v[rng.random(n) < p0] = 0
            out |= v << np.uint64(sh)
    return out


FORMATIONS = [0x00000000FFFFFFFF, 0x000000000000FFFF, 0x00000000FFF0FFF0, 0x00000000FF00FF00, 0]


# ---------------------------------------------------------------- primitives

@pytest.mark.parametrize("ff,ig,prune", [(0, 0, 0), (0, 1, 0), (1, 0, 0), (0, 1, 1)])
def test_move_kernel(cg, ff, ig, prune):
    rng = np.random.default_rng(1 + ff + 2 * ig + 4 * prune)
    cfg = O.make_cfg(free_formation=ff, ignore_f=ig, prune=prune, stsl=58, ltc=5, smallest_large=6)
    r = cg.rules(free_formation=ff, ignore_f=ig, prune=prune, stsl=58, ltc=5, smallest_large=6)
    for fm in FORMATIONS:
        b = rand_boards(rng, fm, 20000)
        exp = []
        for x in b:
            x = int(x)
            if prune and O.orc().orc_prune(x, 58, 5, 6):
                continue
            for d in range(4):
                ok, m = O.move(cfg, x, d)
                if ok and not (prune and O.orc().orc_prune(m, 58, 5, 6)):
                    exp.append(O.canon(m))
        din = dev(b)
        dout = torch.empty(4 * len(b) + 1, dtype=torch.int64, device="cuda")
        n = C.c_size_t()
        assert cg.lib().cg_dev_move(C.byref(r), din.data_ptr(), len(b), dout.data_ptr(),
                                    C.byref(n)) == 0
        got = np.sort(host(dout, n.value))
        assert np.array_equal(got, np.sort(np.array(exp, dtype=np.uint64)))


@pytest.mark.parametrize("stsl,ltc,sl", [(58, 5, 6), (0, 2, 3), (300, 16, 0), (100, 8, 13),
                                         (20, 4, 14), (2000, 16, 15), (1000, 16, -1)])
def test_move_kernel_prune_parameters(cg, stsl, ltc, sl):
    """prune_board (generate.c:98-131) over its parameter range: a value equal
    to smallest_large may occur 5 times, other large values twice; boards
    with 16 equal cells, 0xE tiles and 16 large tiles included."""
    rng = np.random.default_rng(100 + sl)
    cfg = O.make_cfg(ignore_f=1, prune=1, stsl=stsl, ltc=ltc, smallest_large=sl)
    r = cg.rules(ignore_f=1, prune=1, stsl=stsl, ltc=ltc, smallest_large=sl)
    b = np.concatenate([
        rand_boards(rng, 0, 3000, p0=0.2, kmax=13),
        rand_boards(rng, 0x000000000000FFFF, 1000, p0=0.1, kmax=14),
        np.array([u * 0x1111111111111111 for u in range(16)], dtype=np.uint64),
        # 15 equal cells and one different one, at every position
        np.array([(u * 0x1111111111111111) ^ (w << (4 * x)) for u in (0, 1, 6, 13)
                  for w in (1, 3) for x in range(16)], dtype=np.uint64),
        # five / six copies of one value among small tiles
        np.array([int(sum(v << (4 * x) for x, v in enumerate([k] * c + [1, 2] * 4)))
                  for k in (3, 6, 7, 13) for c in (2, 3, 5, 6)], dtype=np.uint64)])
    exp, pruned = [], 0
    for x in b:
        x = int(x)
        if O.orc().orc_prune(x, stsl, ltc, sl):
            pruned += 1
            continue
        for d in range(4):
            ok, m = O.move(cfg, x, d)
            if ok and not O.orc().orc_prune(m, stsl, ltc, sl):
                exp.append(O.canon(m))
    assert 0 < pruned < len(b)
    dout = torch.empty(4 * len(b) + 1, dtype=torch.int64, device="cuda")
    n = C.c_size_t()
    assert cg.lib().cg_dev_move(C.byref(r), dev(b).data_ptr(), len(b), dout.data_ptr(),
                                C.byref(n)) == 0
    got = np.sort(host(dout, n.value))
    assert np.array_equal(got, np.sort(np.array(exp, dtype=np.uint64)))


def test_generation_move_full_hash_sets(cg):
    """The generation step's move kernel takes 2048 boards per round into one
    8192-slot LDS hash set.  Boards with four empty cells and distinct tiles
    give about four distinct candidates each, so the sets run ~95% full;
    a block of one repeated board gives the opposite extreme.  The first layer
    of a one-step generation is uniq(canon(moves(initial))) (generate.c:409-550),
    checked against the oracle."""
    rng = np.random.default_rng(11)
    n = 1 << 19
    b = np.zeros(n, dtype=np.uint64)
    for x in range(16):
        b |= rng.integers(1, 12, size=n).astype(np.uint64) << np.uint64(60 - 4 * x)
    for _ in range(4):
        pos = rng.integers(0, 16, size=n).astype(np.uint64)
        b &= ~(np.uint64(0xF) << (np.uint64(60) - np.uint64(4) * pos))
    b[: 1 << 14] = b[0]
    cfg = O.make_cfg(cores=8)
    M, _, _ = O.gen_layer(cfg, b, np.zeros(0, np.uint64))
    assert len(M) > 3 * (n - (1 << 14))
    tb = cg.Tablebase(cg.rules())
    tb.generate(b, 100, 100)
    assert np.array_equal(tb.keys(0), M)


def test_spawn_kernel(cg):
    rng = np.random.default_rng(7)
    for fm in FORMATIONS:
        b = rand_boards(rng, fm, 5000)
        e2, e4 = [], []
        for x in b:
            x = int(x)
            for c in range(16):
                if (x >> (60 - 4 * c)) & 0xf == 0:
                    e2.append(O.canon(x | (1 << (60 - 4 * c))))
                    e4.append(O.canon(x | (2 << (60 - 4 * c))))
        d2 = torch.empty(16 * len(b), dtype=torch.int64, device="cuda")
        d4 = torch.empty(16 * len(b), dtype=torch.int64, device="cuda")
        n = C.c_size_t()
        assert cg.lib().cg_dev_spawn(dev(b).data_ptr(), len(b), d2.data_ptr(), d4.data_ptr(),
                                     C.byref(n)) == 0
        assert n.value == len(e2)
        # same order within a board (ascending cell); boards' blocks may interleave
        assert np.array_equal(np.sort(host(d2, n.value)), np.sort(np.array(e2, np.uint64)))
        assert np.array_equal(np.sort(host(d4, n.value)), np.sort(np.array(e4, np.uint64)))


@pytest.mark.parametrize("case", ["empty", "one", "two_equal", "all_equal", "random_dups",
                                  "full64", "one_bit", "high_bits", "big"])
def test_sort_unique(cg, case):
    rng = np.random.default_rng(hash(case) & 0xffff)
    if case == "empty":
        a = np.zeros(0, np.uint64)
    elif case == "one":
        a = np.array([12345], np.uint64)
    elif case == "two_equal":
        a = np.array([7, 7], np.uint64)
    elif case == "all_equal":
        a = np.full(100000, 0xFFF1FFF200000000, np.uint64)
    elif case == "random_dups":
        a = rng.integers(0, 50000, size=300001, dtype=np.uint64)
    elif case == "full64":
        a = rng.integers(0, 2**64 - 1, size=200000, dtype=np.uint64, endpoint=True)
        a[::7] = a[0]
        a[1] = 0
        a[2] = np.uint64(2**64 - 1)
    elif case == "one_bit":
        a = (rng.integers(0, 2, size=70000, dtype=np.uint64) << np.uint64(37)) | np.uint64(5)
    elif case == "high_bits":
        a = rng.integers(0, 2**20, size=100000, dtype=np.uint64) << np.uint64(44)
    else:
        a = rand_boards(rng, 0x00000000FFF0FFF0, 1 << 24)
    d = dev(a) if len(a) else torch.empty(1, dtype=torch.int64, device="cuda")
    out = torch.empty(max(len(a), 1), dtype=torch.int64, device="cuda")
    n = C.c_size_t()
    assert cg.lib().cg_dev_sort_unique(d.data_ptr(), len(a), out.data_ptr(), C.byref(n)) == 0
    assert np.array_equal(host(out, n.value), np.unique(a))


SORT_CONFIGS = [(8, 32, 16), (8, 32, 32), (8, 16, 16), (8, 16, 32), (8, 8, 16), (8, 8, 32),
                (9, 16, 16), (9, 16, 32), (9, 32, 16), (9, 32, 32), (9, 7, 16), (9, 7, 32),
                (10, 8, 16), (10, 8, 32)]


def _sort_check(cg, a):
    d = dev(a) if len(a) else torch.empty(1, dtype=torch.int64, device="cuda")
    out = torch.empty(max(len(a), 1), dtype=torch.int64, device="cuda")
    n = C.c_size_t()
    assert cg.lib().cg_dev_sort_unique(d.data_ptr(), len(a), out.data_ptr(), C.byref(n)) == 0
    assert np.array_equal(host(out, n.value), np.unique(a))


@pytest.mark.parametrize("cfg", SORT_CONFIGS, ids=lambda c: "%dx%d_w%d" % c)
def test_sort_every_configuration(cg, cfg):
    """Each (digit width, keys per thread, look-back window) kernel against
    numpy's unique: tile-boundary sizes, key widths that leave a partial last
    digit, 64-bit keys, heavy duplicates and board keys."""
    bits, kpt, lbw = cfg
    tile = (1024 if kpt <= 8 else 1 << bits) * kpt
    rng = np.random.default_rng(bits * 1000 + kpt * 10 + lbw)
    try:
        cg.tune_set("sample_sort", 0)   # the LSD kernels themselves
        cg.tune_set("sort_bits", bits)
        cg.tune_set("sort_kpt", kpt)
        cg.tune_set("sort_lbw", lbw)
        assert cg.tune_sort_query(1 << 22, 40) == cfg
        for n in (1, tile - 1, tile, tile + 1, 3 * tile + 17):
            _sort_check(cg, rng.integers(0, 1 << 13, size=n, dtype=np.uint64) << np.uint64(7))
        for nb in (bits + 1, 2 * bits - 1, 3 * bits, 45, 54):
            _sort_check(cg, rng.integers(0, 1 << nb, size=200003, dtype=np.uint64))
        full = rng.integers(0, 2**64 - 1, size=300000, dtype=np.uint64, endpoint=True)
        full[::5] = full[3]
        _sort_check(cg, full)
        _sort_check(cg, rand_boards(rng, 0x00000000FFF0FFF0, 1 << 22))
    finally:
        cg.tune_set("reset", 0)


def cell_keys(rng, n, cells, vmax, const=0xFFF0FFF000000000):
    """n keys equal to `const` outside `cells` (0 = most significant nibble),
    whose values are 0..vmax"""
    k = np.full(n, const, np.uint64)
    for c in cells:
        sh = np.uint64(60 - 4 * c)
        k &= ~(np.uint64(15) << sh)
        k |= rng.integers(0, vmax + 1, n).astype(np.uint64) << sh
    return k


def bounded(v):
    """a cell_max word bounding every cell by v"""
    return v * 0x1111111111111111


def _bounded_check(cg, a, cell_max):
    d = dev(a) if len(a) else torch.empty(1, dtype=torch.int64, device="cuda")
    out = torch.empty(max(len(a), 1), dtype=torch.int64, device="cuda")
    n = C.c_size_t()
    assert cg.lib().cg_dev_sort_unique_bounded(d.data_ptr(), len(a), out.data_ptr(), C.byref(n),
                                               cell_max) == 0, cg.last_error()
    assert np.array_equal(host(out, n.value), np.unique(a))


DPD_CELLS = [3, 7, 8, 9, 10, 11, 12, 13, 14, 15]
# per digit width: cells and bound for which the cell-pair code saves a pass
CODE_CASES = {8: (DPD_CELLS[:9], 10), 9: (DPD_CELLS, 10), 10: (DPD_CELLS, 7)}


@pytest.mark.parametrize("cfg", SORT_CONFIGS, ids=lambda c: "%dx%d_w%d" % c)
def test_sort_cell_code_every_configuration(cg, cfg):
    """Bounded sorts through the cell-pair code against numpy's unique, with
    each pass kernel encoding in its first pass: tile-boundary sizes and
    heavy duplicates.  Every call with enough keys for all the cells to vary
    must use the code (two keys may differ in too few cells to save a pass)."""
    bits, kpt, lbw = cfg
    tile = (1024 if kpt <= 8 else 1 << bits) * kpt
    cells, vmax = CODE_CASES[bits]
    rng = np.random.default_rng(7 + bits * 1000 + kpt * 10 + lbw)
    try:
        cg.tune_set("sample_sort", 0)
        cg.tune_set("sort_bits", bits)
        cg.tune_set("sort_kpt", kpt)
        cg.tune_set("sort_lbw", lbw)
        for n in (1, 2, tile + 1, 3 * tile + 17, 200003):
            before = cg.tune_get("cell_code_sorts")
            a = cell_keys(rng, n, cells, vmax)
            if n >= 3:
                a[::3] = a[1]
            _bounded_check(cg, a, bounded(vmax))
            if n > tile:
                assert cg.tune_get("cell_code_sorts") == before + 1
    finally:
        cg.tune_set("reset", 0)


def test_sort_cell_code_tuned_sizes(cg):
    """DPDGap-like keys (ten varying cells of values 0..10, constant 0xf
    cells) across the tuned table's size classes; bit packing and the code
    give the same unique keys."""
    rng = np.random.default_rng(11)
    try:
        cg.tune_set("sample_sort", 0)
        for n in (2, 1000, 70001, 3000017, (1 << 23) + 5):
            a = cell_keys(rng, n, DPD_CELLS, 10)
            before = cg.tune_get("cell_code_sorts")
            _bounded_check(cg, a, bounded(10))
            used = cg.tune_get("cell_code_sorts") - before
            assert used == 1 or n == 2
            _sort_check(cg, a)
            assert cg.tune_get("cell_code_sorts") == before + used
    finally:
        cg.tune_set("reset", 0)


def test_sort_cell_code_bound_violation(cg):
    """A key whose cell exceeds the stated bound makes the call fail (the
    histogram kernel checks every cell it codes); the next call succeeds."""
    rng = np.random.default_rng(12)
    a = cell_keys(rng, 100000, DPD_CELLS, 10)
    a[777] = (a[777] & ~np.uint64(0xF << 8)) | np.uint64(12 << 8)
    d, out, n = dev(a), torch.empty(len(a), dtype=torch.int64, device="cuda"), C.c_size_t()
    try:
        cg.tune_set("sample_sort", 0)   # the code is an LSD-path encoding
        assert cg.lib().cg_dev_sort_unique_bounded(d.data_ptr(), len(a), out.data_ptr(), C.byref(n),
                                                   bounded(10)) != 0
        assert "bound" in cg.last_error()
        _bounded_check(cg, cell_keys(rng, 5000, DPD_CELLS, 10), bounded(10))
    finally:
        cg.tune_set("reset", 0)


def test_sort_cell_code_per_cell_bases(cg):
    """Composite keys (a table id above ten free cells, as the batch builds
    them) with a bound per cell, against numpy's unique."""
    rng = np.random.default_rng(13)
    try:
        cg.tune_set("sample_sort", 0)
        for n in (5000, 300001):
            tid = rng.integers(0, 40, n).astype(np.uint64)
            a = cell_keys(rng, n, list(range(6, 16)), 8, const=0) | (tid << np.uint64(48))
            before = cg.tune_get("cell_code_sorts")
            _bounded_check(cg, a, (2 << 52) | (15 << 48) | (bounded(8) >> 16))
            assert cg.tune_get("cell_code_sorts") == before + 1
    finally:
        cg.tune_set("reset", 0)


@pytest.mark.parametrize("name", ["c1", "dpdgap_prune", "two_initial"])
def test_tablebase_canon_stabiliser_off(cg, name):
    """Generation with every canonical form the full max of 8 images (knob
    canon_stab 0) matches the oracle as the default (the formation's
    stabiliser, which every other tablebase test runs) does."""
    try:
        cg.tune_set("canon_stab", 0)
        assert cg.tune_get("canon_stab") == 0
        test_tablebase_vs_oracle(cg, name)
    finally:
        cg.tune_set("reset", 0)
    assert cg.tune_get("canon_stab") == 1


@pytest.mark.parametrize("name", ["c1", "dpdgap_prune", "two_initial"])
def test_tablebase_cell_code_on_and_off(cg, name):
    """Generation with the cell-pair code (the default; the initial boards fix
    each layer's tile sum and 0xf cells) and with bit packing only: both
    match the oracle, and the default run used the code."""
    try:
        cg.tune_set("sample_sort", 0)   # every layer sort on the LSD path
        before = cg.tune_get("cell_code_sorts")
        test_tablebase_vs_oracle(cg, name)
        assert cg.tune_get("cell_code_sorts") > before
        cg.tune_set("cell_code", 0)
        before = cg.tune_get("cell_code_sorts")
        test_tablebase_vs_oracle(cg, name)
        assert cg.tune_get("cell_code_sorts") == before
    finally:
        cg.tune_set("reset", 0)


@pytest.mark.parametrize("delta", [-3, -1, 1, 2])
def test_lookup_directory_depths(cg, delta):
    """Lookups through shallower and deeper prefix-bucket directories."""
    rng = np.random.default_rng(100 + delta)
    keys = np.unique(rand_boards(rng, 0x00000000FFF0FFF0, 200000))
    vals = rng.random(len(keys))
    q = np.concatenate([keys[rng.integers(0, len(keys), size=20000)],
                        rand_boards(rng, 0x00000000FFF0FFF0, 5000)])
    exp = np.where(np.isin(q, keys), vals[np.clip(np.searchsorted(keys, q), 0, len(keys) - 1)], 0.0)
    try:
        cg.tune_set("dir_delta", delta)
        kd, vd, qd = dev(keys), torch.from_numpy(vals).cuda(), dev(q)
        out = torch.empty(len(q), dtype=torch.float64, device="cuda")
        assert cg.lib().cg_dev_lookup(kd.data_ptr(), vd.data_ptr(), len(keys), qd.data_ptr(), len(q),
                                      out.data_ptr(), 0) == 0
        assert np.array_equal(out.cpu().numpy(), exp)
    finally:
        cg.tune_set("reset", 0)


def test_stored_table_installed_at_init(cg):
    """Initialisation installs cablegen_amd/tuned/mi355x.json, so every sort in
    these tests runs with the configurations the bench uses."""
    import json
    if not os.path.exists(cg.TUNED_TABLE) or os.environ.get("CABLEGEN_TUNING") is not None:
        pytest.skip("no stored table, or overridden")
    assert cg.lib().cg_dev_sync() == 0   # initialises the engine
    rep = json.load(open(cg.TUNED_TABLE))
    for cls, n in enumerate((1 << 18, 1 << 21, 1 << 25)):
        for nb in (24, 40, 48, 54):
            assert rep["sort"][cls]["pick"][str(nb)] == "%dx%d/%d" % cg.tune_sort_query(n, nb)
    assert cg.tune_get("dir_delta") == rep["dir"]["pick"]


def test_autotune(cg):
    """The autotuner fills the table with kernels that exist and sorts stay exact."""
    try:
        rep = cg.autotune(0)
        assert len(rep["sort"]) == 3 and all(len(c["pick"]) == 64 for c in rep["sort"])
        assert rep["dir"]["pick"] in (-3, -2, -1, 0, 1, 2)
        assert cg.tune_get("dir_delta") == rep["dir"]["pick"]
        for cls, n in enumerate((1 << 18, 1 << 21, 1 << 24)):
            for nb in (20, 40, 54):
                b, k, w = cg.tune_sort_query(n, nb)
                assert (b, k, w) in SORT_CONFIGS
                assert rep["sort"][cls]["pick"][str(nb)] == "%dx%d/%d" % (b, k, w)
        rng = np.random.default_rng(7)
        _sort_check(cg, rand_boards(rng, 0x00000000FFF0FFF0, 1 << 22))
        _sort_check(cg, rng.integers(0, 1 << 54, size=1 << 21, dtype=np.uint64))
    finally:
        cg.tune_set("reset", 0)


@pytest.mark.parametrize("n", [0, 1, 50, 1023, 1024, 1025, 300000])
def test_lookup_kernel(cg, n):
    rng = np.random.default_rng(n + 3)
    keys = np.unique(rand_boards(rng, 0x00000000FFF0FFF0, n) if n else np.zeros(0, np.uint64))
    keys = np.unique(np.array([O.canon(int(k)) for k in keys], dtype=np.uint64)) if n < 5000 else \
        np.unique(keys)
    vals = rng.random(len(keys))
    q = np.concatenate([keys[rng.integers(0, max(len(keys), 1), size=4000)] if len(keys) else
                        np.zeros(0, np.uint64), rand_boards(rng, 0x00000000FFF0FFF0, 1000)])
    exp = np.array([O.orc().orc_lookup(int(x), keys.ctypes.data_as(O.U64P),
                                       vals.ctypes.data_as(O.F64P), len(keys), 0) for x in q])
    dk = dev(keys) if len(keys) else torch.empty(1, dtype=torch.int64, device="cuda")
    dv = torch.from_numpy(vals).cuda() if len(keys) else torch.empty(1, dtype=torch.float64,
                                                                      device="cuda")
    dq = dev(q)
    do = torch.empty(len(q), dtype=torch.float64, device="cuda")
    assert cg.lib().cg_dev_lookup(dk.data_ptr(), dv.data_ptr(), len(keys), dq.data_ptr(), len(q),
                                  do.data_ptr(), 0) == 0
    assert np.array_equal(do.cpu().numpy().view(np.uint64), exp.view(np.uint64))


# ---------------------------------------------------------------- tablebases

def oracle_table(cfg, initial, win, start, end, end_solve=0):
    layers = O.generate(cfg, initial, start, end)
    return layers, O.solve(cfg, layers, start, end, end_solve, win)


def check_tb(tb, layers, vals):
    assert tb.num_layers == len(layers)
    for i in range(len(layers)):
        assert np.array_equal(tb.keys(i), layers[i]), i
        if vals[i] is not None:
            assert np.array_equal(tb.values(i).view(np.uint64), vals[i].view(np.uint64)), i


TB_CASES = {
    "c1": (dict(), [0x11000000FFFFFFFF], [0x7000000000000000], 4, 260),
    "c1_ignore_f": (dict(ignore_f=True), [0x11000000FFFFFFFF], [0x7000000000000000], 4, 260),
    "free_formation": (dict(free_formation=True), [0x11000000FFFFFFFF], [0x7000000000000000], 4, 40),
    "nox": (dict(ignore_f=True, nox=7), [0x11000000FFFFFFFF], [0x7000000000000000], 4, 200),
    "score": (dict(ignore_f=True, score=True), [0x11000000FFFFFFFF], [0x7000000000000000], 4, 120),
    "dpdgap_noprune": (dict(ignore_f=True), [0x00008765FFF3FFF2], [0x0000900000000000], 492, 560),
    "dpdgap_prune": (dict(ignore_f=True, prune=True, stsl=58, ltc=5, smallest_large=6),
                     [0x00008765FFF3FFF2], [0x0000900000000000], 492, 600),
    "two_initial": (dict(ignore_f=True), [0x11000000FFFFFFFF, 0x10100000FFFFFFFF],
                    [0x7000000000000000, 0x0000000000007000], 4, 160),
    "3x4": (dict(ignore_f=True), [0x1200000000000000 | 0xFFFF], [0x6000000000000000], 6, 90),
    # tight prune settings: about 20% (C1) and 85% (DPDGap) of the moved
    # boards fail the test, so the sorts drop boards on every path
    "c1_hard_prune": (dict(ignore_f=True, prune=True, stsl=0, ltc=3, smallest_large=4),
                      [0x11000000FFFFFFFF], [0x7000000000000000], 4, 260),
    "dpdgap_hard_prune": (dict(ignore_f=True, prune=True, stsl=10, ltc=4, smallest_large=5),
                          [0x00008765FFF3FFF2], [0x0000900000000000], 492, 600),
}


# the oracle's tables by case name: many tests run the same case with
# different knobs, and the oracle's answer does not depend on them
_ORACLE_TABLES = {}


def oracle_case(name):
    if name not in _ORACLE_TABLES:
        rules, initial, win, start, end = TB_CASES[name]
        cfg = O.make_cfg(cores=8, **rules)
        _ORACLE_TABLES[name] = oracle_table(cfg, initial, win, start, end)
    return _ORACLE_TABLES[name]


@pytest.mark.parametrize("name", list(TB_CASES))
def test_tablebase_vs_oracle(cg, name):
    rules, initial, win, start, end = TB_CASES[name]
    layers, vals = oracle_case(name)
    tb = cg.Tablebase(cg.rules(**rules))
    tb.generate(initial, start, end)
    tb.solve(win)
    check_tb(tb, layers, vals)
    tb.close()


PRUNE_PATHS = {
    "move_kernel": [("prune_in_sort", 0)],
    "bitonic_buckets": [("bucket_bitonic", 1)],
    "lsd": [("sample_sort", 0)],
    "oversize_repeat": [("sample_cap", 64)],
    "two_syncs": [("one_sync", 0)],
    "lsd_distinct": [("distinct_bitonic", 0)],
    "carry_sorted": [("carry_sort", 1)],
    "carry_sorted_lsd": [("carry_sort", 1), ("sample_sort", 0)],
}


@pytest.mark.parametrize("path", list(PRUNE_PATHS))
@pytest.mark.parametrize("name", ["dpdgap_prune", "c1_hard_prune", "dpdgap_hard_prune"])
def test_tablebase_prune_paths(cg, name, path):
    """With prune, the moved boards are tested by the candidate sort once per
    distinct board and the spawned boards by the spawn kernel once per board
    and value (the test depends only on the board's cells); every path the
    sort can take drops the same boards as the tests in the move kernel
    (prune_in_sort 0): the hash-set buckets, buckets sorted whole by the
    bitonic network, the LSD compaction, the LSD repeat of an oversize
    bucket, the sort sized on the host, LSD passes over the distinct values."""
    try:
        for k, v in PRUNE_PATHS[path]:
            cg.tune_set(k, v)
        test_tablebase_vs_oracle(cg, name)
    finally:
        cg.tune_set("reset", 0)
    assert cg.tune_get("prune_in_sort") == 1


@pytest.mark.parametrize("name", ["c1", "3x4", "dpdgap_noprune", "two_initial", "free_formation"])
@pytest.mark.parametrize("repeat", [False, True])
def test_tablebase_carry_sorted(cg, name, repeat):
    """The spawn-4 carry sorted and deduplicated before it is moved
    (carry_sort 1) instead of moved as the spawn kernel wrote it (the
    default): the pre-move sets are multisets either way and the layers are
    the same; also through the spawn-buffer repeat."""
    try:
        cg.tune_set("carry_sort", 1)
        if repeat:
            cg.tune_set("spawn_estimate", 1)
        test_tablebase_vs_oracle(cg, name)
    finally:
        cg.tune_set("reset", 0)
    assert cg.tune_get("carry_sort") == 0


@pytest.mark.parametrize("waves", [0, 5])
def test_tablebase_solve_waves(cg, waves):
    """The solve kernel with its registers fitted to 5 waves per SIMD or left
    to the compiler (4 waves), against the oracle; every other test runs the
    default, 6 waves."""
    try:
        cg.tune_set("solve_waves", waves)
        test_tablebase_vs_oracle(cg, "dpdgap_prune")
        test_tablebase_vs_oracle(cg, "score")
    finally:
        cg.tune_set("reset", 0)
    assert cg.tune_get("solve_waves") == 6


@pytest.mark.parametrize("count", [1, 0])
def test_tablebase_solve_profiled(cg, count):
    """A solve timed by the profiler: with prof_count 1 the kernel instance
    that counts its lookups (the only one the default runs leave out), with 0
    the production instance under the timers; values against the oracle."""
    try:
        cg.tune_set("prof_count", count)
        cg.prof_enable(True)
        cg.prof_select("solve")
        cg.prof_reset()
        test_tablebase_vs_oracle(cg, "dpdgap_prune")
        n, ms, by = cg.prof_query("solve")
        assert n > 0 and ms > 0
        if count:
            assert by > 0
    finally:
        cg.prof_enable(False)
        cg.prof_select(None)
        cg.tune_set("reset", 0)
    assert cg.tune_get("prof_count") == 1


@pytest.mark.parametrize("name", ["c1", "dpdgap_prune"])
def test_tablebase_directories_per_step(cg, name):
    """The layer directories built after each step (dir_deferred 0) instead
    of after the last one: the same tables and values."""
    try:
        cg.tune_set("dir_deferred", 0)
        assert cg.tune_get("dir_deferred") == 0
        test_tablebase_vs_oracle(cg, name)
    finally:
        cg.tune_set("reset", 0)
    assert cg.tune_get("dir_deferred") == 1


@pytest.mark.parametrize("name", ["c1", "dpdgap_prune"])
def test_tablebase_carry_sort_on_main_stream(cg, name):
    """The carry sort scheduled on the main stream (carry_stream 0) instead of
    the second one gives the same tables."""
    try:
        cg.tune_set("carry_sort", 1)
        cg.tune_set("carry_stream", 0)
        assert cg.tune_get("carry_stream") == 0
        test_tablebase_vs_oracle(cg, name)
    finally:
        cg.tune_set("reset", 0)
    assert cg.tune_get("carry_stream") == 1


@pytest.mark.parametrize("name", ["c1", "dpdgap_prune", "3x4"])
def test_tablebase_two_syncs_per_step(cg, name):
    """The layer step with the candidate count brought to the host before
    the M sort (one_sync 0, the path of steps whose bound is beyond the
    sample sort) gives the same tables as the default step, whose M sort is
    sized on the device and which synchronises once."""
    try:
        cg.tune_set("one_sync", 0)
        assert cg.tune_get("one_sync") == 0
        test_tablebase_vs_oracle(cg, name)
    finally:
        cg.tune_set("reset", 0)
    assert cg.tune_get("one_sync") == 1


@pytest.mark.parametrize("name", ["c1", "3x4"])
def test_tablebase_oversize_carry_on_main_stream(cg, name):
    """Every sample sort overflowing (capacity 64) with the carry sort on the
    main stream: the M sort's and the carry sort's repeats are told apart by
    their error words and both are redone on the LSD path."""
    try:
        cg.tune_set("sample_cap", 64)
        cg.tune_set("carry_sort", 1)
        cg.tune_set("carry_stream", 0)
        test_tablebase_vs_oracle(cg, name)
    finally:
        cg.tune_set("reset", 0)


def test_lsd_generation_after_oversize(cg):
    """An oversize bucket seen by one generation leaves nothing behind: a
    later generation on the LSD path (whose counter reads must not see that
    oversize again and repeat a carry sort whose input the LSD passes have
    overwritten) still matches the oracle."""
    try:
        cg.tune_set("sample_cap", 64)
        cg.tune_set("carry_sort", 1)
        test_tablebase_vs_oracle(cg, "3x4")   # every sample sort overflows and is repeated
    finally:
        cg.tune_set("reset", 0)
    assert cg.tune_get("sample_repeats") > 0
    try:
        cg.tune_set("sample_sort", 0)
        cg.tune_set("carry_sort", 1)
        test_tablebase_vs_oracle(cg, "c1")
        test_tablebase_vs_oracle(cg, "two_initial")
    finally:
        cg.tune_set("reset", 0)


@pytest.mark.parametrize("name", ["c1", "dpdgap_prune", "3x4"])
def test_tablebase_spawn_buffer_repeat(cg, name):
    """Undersized spawn buffers: every growing step overflows them and is
    repeated with exact buffers; layers and values stay identical."""
    try:
        cg.tune_set("spawn_estimate", 1)
        test_tablebase_vs_oracle(cg, name)
    finally:
        cg.tune_set("reset", 0)


def test_partial_solve_range(cg):
    rules, initial, win, start, end = TB_CASES["c1_ignore_f"]
    cfg = O.make_cfg(cores=8, **rules)
    layers, vals = oracle_table(cfg, initial, win, start, end, end_solve=150)
    tb = cg.Tablebase(cg.rules(**rules))
    tb.generate(initial, start, end)
    tb.solve(win, end_solve=150)
    for i in range(len(layers)):
        if vals[i] is None:
            continue
        assert np.array_equal(tb.values(i).view(np.uint64), vals[i].view(np.uint64))
    tb.close()


def test_c1_manifest(cg):
    man = json.load(open(os.path.join(HERE, "golden", "c1_ref.json")))
    for name, ig in (("default", False), ("ignore_f", True)):
        rows = man["variants"][name]
        tb = cg.Tablebase(cg.rules(ignore_f=ig))
        tb.generate([0x11000000FFFFFFFF], rows[0][0], rows[-1][0])
        tb.solve([0x7000000000000000])
        for i, (s, n, hk, hv) in enumerate(rows):
            assert tb.layer_size(i) == n
            assert "%016x" % O.hash_u64(tb.keys(i)) == hk, s
            assert "%016x" % O.hash_u64(tb.values(i).view(np.uint64)) == hv, s
        tb.close()


# ---------------------------------------------------------------- drop-in ABI

def test_dropin_generate_solve_files(cg, tmp_path):
    """generate()/solve() through the api/cablegen.h entry points, with the
    settings read from a config file, against the reference-made manifest."""
    d = str(tmp_path)
    os.makedirs(d + "/b")
    os.makedirs(d + "/t")
    with open(d + "/c.conf", "w") as f:
        f.write("[Cablegen]\nignore_f = true\ncores = 4\n[Generate]\nprune = false\n"
                f"dir = {d}/b/\nend = 260\n[Solve]\ndir = {d}/t/\nend = 0\n")
    L = cg.lib()
    L.change_config((d + "/c.conf").encode())
    init = np.array([0x11000000FFFFFFFF], np.uint64)
    win = np.array([0x7000000000000000], np.uint64)
    sa = cg.SArr(True, init.ctypes.data_as(cg.U64P), 1)
    L.generate(4, 260, f"{d}/b/%d.boards".encode(), C.byref(sa))
    wa = cg.SArr(True, win.ctypes.data_as(cg.U64P), 1)
    L.solve(260, 0, f"{d}/b/%d.boards".encode(), f"{d}/t/%d.tables".encode(), C.byref(wa))
    rows = json.load(open(os.path.join(HERE, "golden", "c1_ref.json")))["variants"]["ignore_f"]
    for s, n, hk, hv in rows:
        b = O.read_boards_file(f"{d}/b/{s}.boards")
        k, v = O.read_table_file(f"{d}/t/{s}.tables")
        assert len(b) == n and "%016x" % O.hash_u64(b) == hk, s
        assert np.array_equal(k, b)
        assert "%016x" % O.hash_u64(v.view(np.uint64)) == hv, s
    # layers below the start are solved as empty tables (missing files)
    assert os.path.getsize(f"{d}/t/2.tables") == 0


# ---------------------------------------------------------------- adjoint

def test_adjoint_vs_oracle_and_fd(cg):
    rules, initial, win, start, end = TB_CASES["c1_ignore_f"]
    cfg = O.make_cfg(cores=8, **rules)
    layers, vals = oracle_table(cfg, initial, win, start, end)
    Po, dPo, Do = O.adjoint(cfg, layers, start, vals, win, initial)
    tb = cg.Tablebase(cg.rules(**rules))
    tb.generate(initial, start, end)
    tb.solve(win, keep_argmax=True)
    P, dP, D = tb.adjoint(initial)
    assert P == Po
    assert abs(dP - dPo) <= 1e-12 * abs(dPo)
    assert np.allclose(D, Do, rtol=1e-10, atol=1e-15)

    def p_at(p4):
        tb.solve(win, p4=p4, keep_argmax=True)
        return tb.adjoint(initial)[0]

    h = 1e-6
    fd = (p_at(0.1 + h) - p_at(0.1 - h)) / (2 * h)
    assert abs(fd - dP) <= 1e-6 * abs(dP)
    tb.close()


@pytest.mark.parametrize("name", ["c1_ignore_f", "dpdgap_prune", "c1_hard_prune", "dpdgap_hard_prune"])
def test_tablebase_premove_vs_oracle(cg, name):
    """premove = 1 (the initial boards are positions after the move, taken in
    any orientation): layers and values against the oracle's premove
    generation, and the GPU's premove run from a plain run's first layer
    equals the plain run."""
    rules, initial, win, start, end = TB_CASES[name]
    end = min(end, start + 80)
    cfg = O.make_cfg(cores=8, **rules)
    plain = O.generate(cfg, initial, start, end)
    b0 = [O.orc().orc_rotate_cw(int(x)) for x in plain[0][::3]] + [int(x) for x in plain[0]]
    cfgp = O.make_cfg(cores=8, premove=True, **rules)
    layers = O.generate(cfgp, b0, start, end)
    vals = O.solve(cfgp, layers, start, end, 0, win)
    tb = cg.Tablebase(cg.rules(premove=True, **rules))
    tb.generate(b0, start, end)
    tb.solve(win, keep_argmax=True)
    assert tb.num_layers == len(layers)
    for i in range(len(layers)):
        assert np.array_equal(tb.keys(i), layers[i]), i
        assert np.array_equal(tb.keys(i), plain[i]), i
        assert np.array_equal(tb.values(i).view(np.uint64), vals[i].view(np.uint64)), i
    Po, dPo, Do = O.adjoint(cfgp, layers, start, vals, win, b0)
    P, dP, D = tb.adjoint(b0)
    assert P == Po
    assert abs(dP - dPo) <= 1e-12 * abs(dPo)
    assert np.allclose(D, Do, rtol=1e-10, atol=1e-15)
    tb.close()


def test_batch_rejects_premove(cg):
    r = cg.rules(ignore_f=True, premove=True)
    assert cg.lib().cg_batch_create(C.byref(r)) is None
    assert "premove" in cg.last_error()


def test_adjoint_autograd(cg):
    from cablegen_amd.autograd import win_probability
    rules, initial, win, start, end = TB_CASES["c1_ignore_f"]
    tb = cg.Tablebase(cg.rules(**rules))
    tb.generate(initial, start, 150)
    p4 = torch.tensor(0.1, dtype=torch.float64, requires_grad=True)
    P = win_probability(p4, tb, initial, win)
    P.backward()
    tb.solve(win, keep_argmax=True)
    _, dP, _ = tb.adjoint(initial)
    # the occupation push uses FP64 atomics, so two sweeps may differ in the last ulps
    assert abs(p4.grad.item() - dP) <= 1e-12 * abs(dP)
    assert torch.autograd.gradcheck(lambda x: win_probability(x, tb, initial, win),
                                    (torch.tensor(0.1, dtype=torch.float64, requires_grad=True),),
                                    eps=1e-6, atol=1e-8, rtol=1e-6)
    tb.close()


# ---------------------------------------------------------------- full size C2

C2 = dict(rules=dict(ignore_f=True, prune=True, stsl=58, ltc=5, smallest_large=6),
          initial=[0x00008765FFF3FFF2], win=[0x0000900000000000], end=1100)


@pytest.fixture(scope="module")
def c2_table(cg):
    tb = cg.Tablebase(cg.rules(**C2["rules"]))
    tb.generate(C2["initial"], 492, C2["end"])
    tb.solve(C2["win"], keep_argmax=True)
    yield tb
    tb.close()


def test_c2_properties(cg, c2_table):
    """Size-independent properties at full size: every layer sorted, unique,
    canonical prefix constant, tile sums match; values in [0, 1]."""
    tb = c2_table
    assert tb.num_layers == 305
    for i in range(0, tb.num_layers, 7):
        k = tb.keys(i)
        if len(k) == 0:
            continue
        assert np.all(k[1:] > k[:-1])
        assert np.all((k & np.uint64(0xFFF0FFF000000000)) == np.uint64(0xFFF0FFF000000000))
        s = 492 + 2 * i
        for x in k[:: max(1, len(k) // 200)]:
            assert O.orc().orc_get_sum(int(x)) == s
            assert O.canon(int(x)) == int(x)
        v = tb.values(i)
        assert np.all((v >= 0) & (v <= 1))


def test_c2_sampled_layers_vs_oracle(cg, c2_table):
    """Full-size layers: one oracle generation step and solve of mid-table
    layers, from the GPU's own neighbouring layers, must match bit for bit."""
    tb = c2_table
    cfg = O.make_cfg(cores=16, **C2["rules"])
    W = O.expand_winstates(C2["win"])
    for i in (60, 150):
        s2a, s4a = O.spawn_sets(cfg, tb.keys(i - 1))
        _, s4b = O.spawn_sets(cfg, tb.keys(i - 2))
        P = np.union1d(s2a, s4b).astype(np.uint64)
        M, P2, C2_ = O.gen_layer(cfg, P, s4a)
        assert np.array_equal(M, tb.keys(i))
        sel = np.arange(0, len(M), 13)
        V = O.solve_layer(cfg, M[sel], tb.keys(i + 1), tb.values(i + 1), tb.keys(i + 2),
                          tb.values(i + 2), W)
        assert np.array_equal(V.view(np.uint64), tb.values(i)[sel].view(np.uint64))


@pytest.mark.skipif(not os.path.exists(os.path.join(HERE, "golden", "c2_oracle.json")),
                    reason="C2 manifest not generated")
def test_c2_manifest(cg, c2_table):
    """Every layer and every value of the full DPDGap-512 table against the
    oracle's manifest (keys and value bits)."""
    tb = c2_table
    rows = json.load(open(os.path.join(HERE, "golden", "c2_oracle.json")))["layers"]
    assert len(rows) == tb.num_layers
    for i, (s, n, hk, hv) in enumerate(rows):
        assert tb.layer_size(i) == n, s
        assert "%016x" % O.hash_u64(tb.keys(i)) == hk, s
        assert "%016x" % O.hash_u64(tb.values(i).view(np.uint64)) == hv, s


@pytest.mark.skipif(not os.path.exists(os.path.join(HERE, "golden", "c5_oracle.json")),
                    reason="C5 oracle gradient not generated")
def test_c5_adjoint_vs_oracle(cg, c2_table):
    """C5: dP/dp4 and the per-layer sensitivities D_s on the full DPDGap-512
    table against the oracle's FP64 adjoint (tests/golden/c5_oracle.json,
    written by make_manifests.py c5): P bit for bit (the solve is exact),
    dP/dp4 and every D_s within 1e-6 relative (the GPU pushes occupation with
    FP64 atomics, so its sums run in another order)."""
    tb = c2_table
    ref = json.load(open(os.path.join(HERE, "golden", "c5_oracle.json")))
    P, dP, D = tb.adjoint(C2["initial"])
    assert P == ref["P"]
    assert abs(dP - ref["dP_dp4"]) <= 1e-6 * abs(ref["dP_dp4"]), (dP, ref["dP_dp4"])
    exp = np.zeros(tb.num_layers)
    for s, d in ref["D"]:
        exp[(s - 492) // 2] = d
    assert len(ref["D"]) > 100
    assert np.allclose(D, exp, rtol=1e-6, atol=1e-300), np.max(np.abs(D - exp))
    assert np.all((D == 0) == (exp == 0))


def test_c2_adjoint_fd(cg, c2_table):
    """dP/dp4 on the DPDGap-512 table against central differences of FP64
    re-solves at p4 = 0.1 +- 1e-5."""
    tb = c2_table
    P, dP, D = tb.adjoint(C2["initial"])
    assert np.isclose(D.sum(), dP, rtol=1e-12)

    def p_at(p4):
        tb.solve(C2["win"], p4=p4, keep_argmax=True)
        return tb.adjoint(C2["initial"])[0]

    h = 1e-5
    fd = (p_at(0.1 + h) - p_at(0.1 - h)) / (2 * h)
    tb.solve(C2["win"], keep_argmax=True)
    assert abs(fd - dP) <= 1e-6 * abs(dP), (fd, dP)


# ---------------------------------------------------------------- batched tables

def test_batch_matches_single_tables(cg):
    from cablegen_amd.batch_configs import batch
    cfgs = batch(12, lo=3, hi=12)
    r = cg.rules(ignore_f=True)
    B = cg.Batch(r)
    for c in cfgs:
        B.add(c["initial"], c["start"], c["end"], c["winstates"])
    before = cg.tune_get("sample_sorts")
    g0 = cg.tune_get("graph_steps")
    B.generate()
    # the batch's layer sorts take the sample sort (the default path), and a
    # layer step is one graph launch (those whose candidate bound is beyond
    # the sample sort read the candidate count first, outside a graph)
    assert cg.tune_get("sample_sorts") > before
    steps = max(B.num_layers(t) for t in range(len(cfgs)))
    assert steps // 2 <= cg.tune_get("graph_steps") - g0 <= steps
    B.solve(use_graph=False)
    eager = [[B.values(t, i) for i in range(B.num_layers(t))] for t in range(len(cfgs))]
    for t, c in enumerate(cfgs):
        tb = cg.Tablebase(r)
        tb.generate(c["initial"], c["start"], c["end"])
        tb.solve(c["winstates"])
        assert B.num_layers(t) == tb.num_layers
        for i in range(tb.num_layers):
            assert np.array_equal(B.keys(t, i), tb.keys(i)), (t, i)
            assert np.array_equal(eager[t][i].view(np.uint64), tb.values(i).view(np.uint64)), (t, i)
        tb.close()
        if t < 3:   # and against the oracle
            cfg = O.make_cfg(cores=8, ignore_f=True)
            layers, vals = oracle_table(cfg, c["initial"], c["winstates"], c["start"], c["end"])
            for i in range(len(layers)):
                assert np.array_equal(B.keys(t, i), layers[i])
                assert np.array_equal(eager[t][i].view(np.uint64), vals[i].view(np.uint64))
    # graph replays, with eager work and synchronisation between them
    for rep in range(3):
        B.solve(use_graph=True)
        for t in range(len(cfgs)):
            for i in range(B.num_layers(t)):
                assert np.array_equal(B.values(t, i).view(np.uint64), eager[t][i].view(np.uint64))
        B.solve(use_graph=False)
    B.close()


def test_batch_spawn_buffer_repeat(cg):
    """Batched generation with undersized spawn buffers (every growing step is
    repeated) gives the same layers as the single-table path."""
    from cablegen_amd.batch_configs import batch
    cfgs = batch(8, lo=3, hi=14)
    r = cg.rules(ignore_f=True)
    B = cg.Batch(r)
    for c in cfgs:
        B.add(c["initial"], c["start"], c["end"], c["winstates"])
    try:
        cg.tune_set("spawn_estimate", 1)
        B.generate()
        B.solve(use_graph=False)
    finally:
        cg.tune_set("reset", 0)
    for t, c in enumerate(cfgs):
        tb = cg.Tablebase(r)
        tb.generate(c["initial"], c["start"], c["end"])
        tb.solve(c["winstates"])
        assert B.num_layers(t) == tb.num_layers
        for i in range(tb.num_layers):
            assert np.array_equal(B.keys(t, i), tb.keys(i)), (t, i)
            assert np.array_equal(B.values(t, i).view(np.uint64), tb.values(i).view(np.uint64)), (t, i)
        tb.close()
    B.close()


BATCH_MODES = {
    "eager": [("batch_graph", 0)],
    "two_syncs": [("one_sync", 0)],
    "carry_sorted": [("carry_sort", 1)],
    "carry_sorted_main": [("carry_sort", 1), ("carry_stream", 0)],
    "carry_sorted_eager": [("carry_sort", 1), ("batch_graph", 0)],
    "carry_sorted_two_syncs": [("carry_sort", 1), ("one_sync", 0)],
}


@pytest.mark.parametrize("mode", list(BATCH_MODES))
def test_batch_step_modes(cg, mode):
    """The batched step's launches enqueued one by one (batch_graph 0), with
    the candidate count read before the M sort (one_sync 0), with the
    spawn-4 carry sorted before it is moved (carry_sort 1; on the second
    stream, a branch of the step's graph, or on the main stream), and their
    combinations: the same layers and values as each table alone."""
    from cablegen_amd.batch_configs import batch
    cfgs = batch(10, lo=3, hi=14)
    r = cg.rules(ignore_f=True)
    B = cg.Batch(r)
    for c in cfgs:
        B.add(c["initial"], c["start"], c["end"], c["winstates"])
    try:
        for k, v in BATCH_MODES[mode]:
            cg.tune_set(k, v)
        g0 = cg.tune_get("graph_steps")
        B.generate()
        launched = cg.tune_get("graph_steps") - g0
        B.solve(use_graph=False)
    finally:
        cg.tune_set("reset", 0)
    assert (launched == 0) == any(k in ("batch_graph", "one_sync") for k, _ in BATCH_MODES[mode])
    for t, c in enumerate(cfgs):
        tb = cg.Tablebase(r)
        tb.generate(c["initial"], c["start"], c["end"])
        tb.solve(c["winstates"])
        assert B.num_layers(t) == tb.num_layers
        for i in range(tb.num_layers):
            assert np.array_equal(B.keys(t, i), tb.keys(i)), (t, i)
            assert np.array_equal(B.values(t, i).view(np.uint64), tb.values(i).view(np.uint64)), (t, i)
        tb.close()
    B.close()


def test_batch_rejects_bad_tables(cg):
    B = cg.Batch(cg.rules(ignore_f=True))
    with pytest.raises(cg.CablegenError):
        B.add([0x1200000000000000], 4, 20, [0x5000000000000000])   # 16 free cells
    B.close()


# ---------------------------------------------------------------- sample sort

def _sample_sort_check(cg, a, expect_sample=True):
    before = cg.tune_get("sample_sorts")
    _sort_check(cg, a)
    if expect_sample:
        assert cg.tune_get("sample_sorts") > before


@pytest.mark.parametrize("n", [2, 3, 1000, 8191, 8192, 8193, 16383, 16384, 16385, 70001, 1000003, 5000011,
                               9500009, 14000009])
def test_sample_sort_sizes(cg, n):
    """The sample sort (splitters from a sample, one partition pass into
    LDS-sized buckets, each bucket deduplicated and sorted in LDS) against
    numpy's unique: one-block sizes around 8192 keys and multi-bucket sizes
    up to 1.4e7, on DPDGap boards with the duplicate rate of a move-candidate
    set and on full 64-bit keys (up to 2048 buckets of 5000 to 7000 keys)."""
    rng = np.random.default_rng(n)
    b = rand_boards(rng, 0x00000000FFF0FFF0, max(n // 3, 1))
    a = b[rng.integers(0, len(b), size=n)]
    _sample_sort_check(cg, a)
    if n <= 1000003:
        _sample_sort_check(cg, rng.integers(0, 2**64 - 1, size=n, dtype=np.uint64, endpoint=True))


@pytest.mark.parametrize("n", [12000007, 20000003, 30000001])
def test_sample_sort_large(cg, n):
    """Sorts above 2048 buckets of 5000 keys: 2048 buckets up to 14.3M keys,
    then buckets of 7000 keys (up to 8192 here): splitters from 16384, 32768
    or 65536 samples (the larger samples ranked in groups of runs), the
    partition with up to 8192 64-bit splitters in LDS; against numpy's unique on
    DPDGap boards with repeats.  With at most 2048 buckets (knob
    sample_bmax) the largest take the LSD path."""
    rng = np.random.default_rng(n + 1)
    b = rand_boards(rng, 0x00000000FFF0FFF0, n // 3)
    a = b[rng.integers(0, len(b), size=n)]
    _sample_sort_check(cg, a)
    try:
        cg.tune_set("sample_bmax", 2048)
        _sample_sort_check(cg, a, expect_sample=n <= 14336000)
    finally:
        cg.tune_set("reset", 0)
    assert cg.tune_get("sample_bmax") == 16384


@pytest.mark.parametrize("n", [60000011, 100000007])
def test_sample_sort_16384_buckets(cg, n):
    """Sorts above 8192 buckets of 7000 keys (the default sample_bmax
    16384): 131072 samples ranked in groups of runs, the partition holding
    its 16383 splitters in LDS as 32-bit sample values (exact: a splitter is
    min + (value << z)); against numpy's unique on DPDGap boards with repeats
    and, for the smaller size, on distinct full 64-bit keys (buckets spanning
    2^32 or more).  With sample_bmax 8192 they take the LSD path."""
    rng = np.random.default_rng(n + 2)
    b = rand_boards(rng, 0x00000000FFF0FFF0, n // 3)
    a = b[rng.integers(0, len(b), size=n)]
    del b
    assert cg.tune_get("sample_bmax") == 16384
    rep = cg.tune_get("sample_repeats")
    _sample_sort_check(cg, a)
    print("n=%d boards: LSD repeats %d" % (n, cg.tune_get("sample_repeats") - rep))
    if n < 80000000:
        a = rng.integers(0, 2**64 - 1, size=n, dtype=np.uint64, endpoint=True)
        a[:3] = (0, 2**64 - 1, 1)
        rep = cg.tune_get("sample_repeats")
        _sample_sort_check(cg, a)
        print("n=%d 64-bit keys: LSD repeats %d" % (n, cg.tune_get("sample_repeats") - rep))
    try:
        cg.tune_set("sample_bmax", 8192)
        before = cg.tune_get("sample_sorts")
        _sort_check(cg, a[: 60000000])
        assert cg.tune_get("sample_sorts") == before
    finally:
        cg.tune_set("reset", 0)
    assert cg.tune_get("sample_bmax") == 16384


@pytest.mark.parametrize("blocks,prefetch", [(0, 1), (1, 1), (7, 1), (8, 1), (512, 0), (7, 0)])
def test_sample_sort_part_blocks(cg, blocks, prefetch):
    """The partition's grid (knob part_blocks): one block per 8192-key tile
    (0), or fewer blocks striding over the tiles (one block for every tile,
    and counts that are and are not a multiple of the 8 region groups), with
    the next tile's keys loaded ahead (knob part_prefetch, the default) or
    not, at 2048, 8192 and 16384 buckets, against numpy's unique."""
    rng = np.random.default_rng(blocks + 77 + prefetch)
    try:
        cg.tune_set("part_blocks", blocks)
        cg.tune_set("part_prefetch", prefetch)
        assert cg.tune_get("part_blocks") == blocks
        assert cg.tune_get("part_prefetch") == prefetch
        for n in (3000017, 20000003, 60000011):
            if blocks == 1 and n > 20000003:
                continue
            b = rand_boards(rng, 0x00000000FFF0FFF0, n // 3)
            _sample_sort_check(cg, b[rng.integers(0, len(b), size=n)])
    finally:
        cg.tune_set("reset", 0)
    assert cg.tune_get("part_blocks") == 512
    assert cg.tune_get("part_prefetch") == 1


def test_sample_sort_key_shapes(cg):
    """All-equal keys, two values, keys spanning the whole 64-bit range with
    the extremes, keys differing only in bit 63 and only in bit 0, and a
    bucket whose keys span more than 2^32 (relative keys are 64-bit)."""
    rng = np.random.default_rng(5)
    n = 200001
    _sample_sort_check(cg, np.full(n, 0x1234, np.uint64))
    _sample_sort_check(cg, np.where(rng.random(n) < 0.5, 7, 2**64 - 1).astype(np.uint64))
    x = rng.integers(0, 2**64 - 1, size=n, dtype=np.uint64, endpoint=True)
    x[:3] = [0, 2**64 - 1, 2**63]
    _sample_sort_check(cg, x)
    _sample_sort_check(cg, (rng.integers(0, 2, n).astype(np.uint64) << np.uint64(63)) | np.uint64(5))
    _sample_sort_check(cg, rng.integers(0, 2, n).astype(np.uint64) | np.uint64(0xABC0))
    # sparse keys: a few clusters far apart
    c = rng.integers(0, 4, n).astype(np.uint64) << np.uint64(50)
    _sample_sort_check(cg, c | rng.integers(0, 1 << 40, n).astype(np.uint64))
    # buckets of nearly all distinct keys spanning 2^32 or more (one block of
    # 8190, and 5000-key buckets): the LDS hash set gives up and the bucket is
    # sorted whole, equal runs dropped; a few repeats among them
    for m in (8190, 40000):
        x = rng.integers(0, 2**64 - 1, size=m, dtype=np.uint64, endpoint=True)
        x[rng.integers(0, m, 50)] = x[0]
        _sample_sort_check(cg, x)


@pytest.mark.parametrize("n", [700, 1500, 3000, 8190, 16000, 40000, 300000])
def test_sample_sort_bitonic_buckets(cg, n):
    """Buckets sorted whole by the bitonic network (knob bucket_bitonic; by
    default only buckets whose hash set fills take it): every row count per
    thread (1 to 16 for 32-bit values, up to 8 for 64-bit ones), one-block
    and multi-bucket sorts, with repeats, against numpy's unique."""
    rng = np.random.default_rng(n + 7)
    b = rand_boards(rng, 0x00000000FFF0FFF0, max(n // 2, 1))
    a = b[rng.integers(0, len(b), size=n)]
    x = rng.integers(0, 2**64 - 1, size=n, dtype=np.uint64, endpoint=True)
    x[rng.integers(0, n, n // 10)] = x[1]
    try:
        cg.tune_set("bucket_bitonic", 1)
        _sample_sort_check(cg, a)
        _sample_sort_check(cg, x)
        if n >= 40000:
            # buckets of ~12000 32-bit values: 16 rows per thread
            cg.tune_set("sample_bucket", 12000)
            _sample_sort_check(cg, a)
    finally:
        cg.tune_set("reset", 0)
    assert cg.tune_get("bucket_bitonic") == 0


@pytest.mark.parametrize("n", [700, 3000, 8190, 40000, 300000, 2000003])
def test_sample_sort_bitonic_distinct(cg, n):
    """A bucket's distinct values sorted by the bitonic network instead of
    LSD passes (knobs distinct_bitonic, distinct_bitonic_bits; by default
    up to 2048 values spanning 9 bits or more): every row count per thread
    from 1 to 16 (distinct counts up to ~12000 in buckets of 12000 keys),
    32- and 64-bit values, values spanning few bits (one LSD pass) and many,
    against numpy's unique, with the threshold at every count, at 2048, and
    with LSD passes only."""
    rng = np.random.default_rng(n + 11)
    b = rand_boards(rng, 0x00000000FFF0FFF0, max(n // 3, 1))
    a = b[rng.integers(0, len(b), size=n)]
    x = rng.integers(0, 2**64 - 1, size=n, dtype=np.uint64, endpoint=True)
    x[rng.integers(0, n, n // 10)] = x[1]
    narrow = (rng.integers(0, 200, size=n).astype(np.uint64) << np.uint64(20)) | np.uint64(0x5000000000)
    try:
        for dmax, dnb in ((16384, 0), (2048, 9)):
            cg.tune_set("distinct_bitonic", dmax)
            cg.tune_set("distinct_bitonic_bits", dnb)
            _sample_sort_check(cg, a)
            _sample_sort_check(cg, x)
            _sample_sort_check(cg, narrow)
            if n >= 40000:
                cg.tune_set("sample_bucket", 12000)
                _sample_sort_check(cg, a)
                _sample_sort_check(cg, np.unique(a))
                cg.tune_set("sample_bucket", 5000)
        cg.tune_set("distinct_bitonic", 0)
        _sample_sort_check(cg, a)
        _sample_sort_check(cg, x)
    finally:
        cg.tune_set("reset", 0)
    assert cg.tune_get("distinct_bitonic") == 2048 and cg.tune_get("distinct_bitonic_bits") == 9


def test_sample_sort_oversize_repeats_on_lsd(cg):
    """A bucket above the capacity (one key repeated 60000 times, or every
    bucket once the capacity is lowered) makes the library repeat the sort on
    the LSD path: the result is exact either way."""
    rng = np.random.default_rng(6)
    a = rand_boards(rng, 0x00000000FFF0FFF0, 300000)
    a[rng.integers(0, len(a), 60000)] = a[17]
    _sort_check(cg, a)
    try:
        cg.tune_set("sample_cap", 64)
        for n in (1000, 70001, 2000003):
            _sort_check(cg, rand_boards(rng, 0x00000000FFF0FFF0, n))
    finally:
        cg.tune_set("reset", 0)


@pytest.mark.parametrize("name", ["c1", "dpdgap_prune"])
def test_tablebase_sample_sort_repeat_path(cg, name):
    """Generation with every sample sort overflowing (capacity 64): each
    step's M sort (and, with carry_sort 1, its carry sort) is repeated on the
    LSD path and the table still matches the oracle."""
    for carry in (0, 1):
        try:
            cg.tune_set("sample_cap", 64)
            cg.tune_set("carry_sort", carry)
            test_tablebase_vs_oracle(cg, name)
        finally:
            cg.tune_set("reset", 0)


def test_batch_sample_sort_repeat_path(cg):
    from cablegen_amd.batch_configs import batch
    cfgs = batch(6, lo=3, hi=10)
    r = cg.rules(ignore_f=True)
    B = cg.Batch(r)
    for c in cfgs:
        B.add(c["initial"], c["start"], c["end"], c["winstates"])
    singles = []
    for c in cfgs:
        tb = cg.Tablebase(r)
        tb.generate(c["initial"], c["start"], c["end"])
        tb.solve(c["winstates"])
        singles.append([(tb.keys(i), tb.values(i).view(np.uint64)) for i in range(tb.num_layers)])
        tb.close()
    for carry in (0, 1):
        try:
            cg.tune_set("sample_cap", 64)
            cg.tune_set("carry_sort", carry)
            B.generate()
            B.solve(use_graph=False)
        finally:
            cg.tune_set("reset", 0)
        for t, layers in enumerate(singles):
            for i, (k, v) in enumerate(layers):
                assert np.array_equal(B.keys(t, i), k), (carry, t, i)
                assert np.array_equal(B.values(t, i).view(np.uint64), v), (carry, t, i)
    B.close()
