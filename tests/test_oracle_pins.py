"""Pins the oracle (oracle/liborc.so) before it is trusted as the parity checker.

Three independent sources:
  * the reference library compiled from its own sources (oracle/_ref; tests
    skip where it was not built), called function by function and end to end;
  * closed forms: the D4 group on a 4x4 nibble grid, an independent pure-Python
    row rule checked exhaustively over all 65,536 rows, a set-based Python BFS
    for the layer recurrence, hand-computable solve values;
  * the committed manifest tests/golden/c1_ref.json made by the reference.
"""
import ctypes as C
import json
import os
import random

import numpy as np
import pytest

from tests import oracle_lib as O

HERE = os.path.dirname(os.path.abspath(__file__))
needs_ref = pytest.mark.skipif(not O.have_ref(), reason="reference library not built")

FORMATIONS = {
    "2x4": 0x00000000FFFFFFFF,
    "3x4": 0x000000000000FFFF,
    "DPDGap": 0x00000000FFF0FFF0,
    "LL": 0x00000000FF00FF00,
    "none": 0,
}


def rand_board(rng, fmask, p0=0.4, kmax=9):
    b = 0
    for x in range(16):
        sh = 60 - 4 * x
        if (fmask >> sh) & 0xf:
            b |= 0xf << sh
        elif rng.random() >= p0:
            b |= rng.randint(1, kmax) << sh
    return b


def py_tile(b, x):
    return (b >> (60 - 4 * x)) & 0xf


# ---------------------------------------------------------------- row rules

def py_row_move(cells, ignore_f):
    """Independent restatement of the row move: returns the new cells of a row
    given in move order (toward index 0)."""
    cells = list(cells)
    keep = len(cells)
    if ignore_f:
        while keep and cells[keep - 1] == 0xf:
            keep -= 1
    head, tail = cells[:keep], cells[keep:]
    tiles = [c for c in head if c]
    out = []
    while tiles:
        if len(tiles) > 1 and tiles[0] == tiles[1] and tiles[0] < 0xe:
            out.append(tiles[0] + 1)
            tiles = tiles[2:]
        else:
            out.append(tiles.pop(0))
    return out + [0] * (keep - len(out)) + tail


def test_row_tables_exhaustive_closed_form():
    for ff in (False, True):
        for ig in (False, True):
            mv, lk = O.lut(ff, ig)
            for r in range(65536):
                c = [(r >> 12) & 15, (r >> 8) & 15, (r >> 4) & 15, r & 15]
                left = py_row_move(c, ig)
                right = py_row_move(c[::-1], ig)[::-1]
                for d, o in ((0, left), (1, right)):
                    val = o[0] << 12 | o[1] << 8 | o[2] << 4 | o[3]
                    assert mv[d][r] == val, (ff, ig, d, hex(r))
                    vac = any(c[i] == 0xf and o[i] != 0xf for i in range(4))
                    assert lk[d][r] == int(ff or not vac), (ff, ig, d, hex(r))


def test_row_mirror_identity():
    """right(r) = rev(left(rev(r))) including the lock table (board.c:105-121)."""
    rev = lambda r: ((r & 0xf) << 12) | ((r & 0xf0) << 4) | ((r >> 4) & 0xf0) | (r >> 12)
    for ig in (False, True):
        mv, lk = O.lut(False, ig)
        r = np.arange(65536)
        rr = np.array([rev(x) for x in range(65536)])
        assert np.array_equal(mv[1], rr[mv[0][rr]] if False else np.array(
            [rev(int(mv[0][rev(x)])) for x in range(65536)]))
        assert np.array_equal(lk[1], lk[0][rr])
        del r


@needs_ref
@pytest.mark.parametrize("ff,ig", [(False, False), (False, True), (True, False), (True, True)])
def test_row_tables_vs_reference(tmp_path, ff, ig):
    L = O.ref()
    O.ref_configure(str(tmp_path), free_formation=ff, ignore_f=ig)
    mvr = np.ctypeslib.as_array((C.c_uint16 * 131072).in_dll(L, "_move_lut")).reshape(2, 65536)
    lkr = np.ctypeslib.as_array((C.c_bool * 131072).in_dll(L, "_locked_lut")).reshape(2, 65536)
    mv, lk = O.lut(ff, ig)
    assert np.array_equal(mv, mvr)
    assert np.array_equal(lk.astype(bool), lkr)


# ---------------------------------------------------------------- symmetries

def test_d4_closed_form():
    rng = random.Random(5)
    for _ in range(3000):
        b = rng.getrandbits(64)
        r = O.orc().orc_rotate_cw(b)
        for rr in range(4):
            for c in range(4):
                assert py_tile(r, 4 * rr + c) == py_tile(b, 4 * (3 - c) + rr)
        f = O.orc().orc_flip(b)
        for x in range(16):
            assert py_tile(f, x) == py_tile(b, 4 * (3 - x // 4) + x % 4)
        imgs = O.all_rots(b)
        c = O.canon(b)
        assert c == max(imgs)
        assert all(O.canon(i) == c for i in imgs)
        assert all(O.orc().orc_get_sum(i) == O.orc().orc_get_sum(b) for i in imgs)
        x = b
        for _ in range(4):
            x = O.orc().orc_rotate_cw(x)
        assert x == b


@needs_ref
def test_symmetries_vs_reference():
    L = O.ref()
    rng = random.Random(7)
    for _ in range(5000):
        b = rng.getrandbits(64)
        for fn, ofn in (("rotate_clockwise", "orc_rotate_cw"), ("flip", "orc_flip"),
                        ("canonicalize_b", "orc_canon")):
            x = C.c_uint64(b)
            getattr(L, fn)(C.byref(x))
            assert x.value == getattr(O.orc(), ofn)(b), fn
        p = L.get_all_rots(b)
        assert [p[i] for i in range(8)] == O.all_rots(b)
        assert L.get_sum(b) == O.orc().orc_get_sum(b)


# ---------------------------------------------------------------- moves

@needs_ref
@pytest.mark.parametrize("ff,ig", [(False, False), (False, True), (True, False)])
def test_moves_vs_reference(tmp_path, ff, ig):
    L = O.ref()
    O.ref_configure(str(tmp_path), free_formation=ff, ignore_f=ig)
    cfg = O.make_cfg(free_formation=ff, ignore_f=ig)
    rng = random.Random(11)
    for name, fm in FORMATIONS.items():
        for _ in range(800):
            b = rand_board(rng, fm)
            for d in range(4):
                x = C.c_uint64(b)
                ok_r = L.movedir_unstable(C.byref(x), d)
                ok_o, y = O.move(cfg, b, d)
                assert bool(ok_r) == ok_o, (name, hex(b), d)
                if ok_o:
                    assert x.value == y, (name, hex(b), d)


# ---------------------------------------------------------------- prune / cmpbrd

def py_prune(b, stsl, ltc, sl):
    vals = [py_tile(b, x) for x in range(16)]
    large = [v for v in vals if sl <= v < 0xe]
    sts = sum(1 << v for v in vals if v < sl and v < 0xe)
    for v in set(large):
        if large.count(v) > (5 if v == sl else 2):
            return True
    return sts > stsl + 64 or len(large) > ltc


def test_prune_closed_form():
    rng = random.Random(3)
    for _ in range(20000):
        b = rand_board(rng, FORMATIONS[rng.choice(list(FORMATIONS))], p0=0.3, kmax=13)
        stsl, ltc, sl = rng.choice([58, 100, 200]), rng.randint(1, 8), rng.randint(3, 9)
        assert O.orc().orc_prune(b, stsl, ltc, sl) == int(py_prune(b, stsl, ltc, sl))


@needs_ref
def test_prune_vs_reference():
    L = O.ref()
    rng = random.Random(4)
    for _ in range(20000):
        b = rand_board(rng, FORMATIONS[rng.choice(list(FORMATIONS))], p0=0.3, kmax=13)
        stsl, ltc, sl = rng.choice([58, 100, 200]), rng.randint(1, 8), rng.randint(3, 9)
        assert bool(L.prune_board(b, stsl, ltc, sl)) == bool(O.orc().orc_prune(b, stsl, ltc, sl))


@needs_ref
def test_cmpbrd_vs_reference():
    L = O.ref()
    rng = random.Random(9)
    for _ in range(20000):
        b = rng.getrandbits(64)
        w = b & rng.getrandbits(64) if rng.random() < 0.5 else rng.getrandbits(64)
        w &= sum(0xf << (4 * i) for i in range(16) if rng.random() < 0.3)
        assert bool(L.cmpbrd(b, w)) == bool(O.orc().orc_cmpbrd(b, w))


# ---------------------------------------------------------------- lookup / merge

def test_lookup_known_answer():
    """test.c:22-52: key i -> value 99 - i; misses give 0."""
    k = np.arange(100, dtype=np.uint64)
    v = (99 - np.arange(100)).astype(np.float64)
    for i in range(100):
        assert O.orc().orc_lookup(i, k.ctypes.data_as(O.U64P), v.ctypes.data_as(O.F64P), 100, 0) == 99 - i
    assert O.orc().orc_lookup(100, k.ctypes.data_as(O.U64P), v.ctypes.data_as(O.F64P), 100, 0) == 0.0


@needs_ref
def test_lookup_vs_reference():
    L = O.ref()
    rng = np.random.default_rng(1)
    for n in (0, 1, 7, 49, 50, 51, 1000, 100000):
        keys = np.unique(rng.integers(0, 2**63, size=n, dtype=np.uint64))
        vals = rng.random(len(keys))
        t = O.Table(O.SArr(True, keys.ctypes.data_as(O.U64P), len(keys)),
                    O.SArr(True, vals.view(np.uint64).ctypes.data_as(O.U64P), len(keys)))
        qs = np.concatenate([keys[:200], rng.integers(0, 2**63, size=200, dtype=np.uint64)])
        for q in qs:
            a = L.lookup(int(q), C.byref(t), False)
            b = O.orc().orc_lookup(int(q), keys.ctypes.data_as(O.U64P), vals.ctypes.data_as(O.F64P),
                                   len(keys), 0)
            assert a == b


@needs_ref
def test_kway_dedupe_vs_reference():
    L = O.ref()
    rng = np.random.default_rng(2)
    for k in (1, 2, 3, 9):
        arrs = [np.sort(rng.integers(0, 5000, size=rng.integers(0, 3000), dtype=np.uint64))
                for _ in range(k)]
        darrs = (O.DArr * k)()
        keep = []
        for i, a in enumerate(arrs):
            buf = (C.c_uint64 * max(len(a), 1))(*[int(x) for x in a])
            keep.append(buf)
            p = C.cast(buf, O.U64P)
            darrs[i] = O.DArr(True, p, C.cast(C.addressof(buf) + 8 * len(a), O.U64P), len(a))
        r = L.deduplicate_threads(darrs, k)
        n = (C.cast(r.sp, C.c_void_p).value - C.cast(r.bp, C.c_void_p).value) // 8
        ref = np.ctypeslib.as_array(r.bp, shape=(n,)).copy() if n else np.zeros(0, np.uint64)
        ptrs = (O.U64P * k)(*[a.ctypes.data_as(O.U64P) for a in arrs])
        szs = (C.c_size_t * k)(*[len(a) for a in arrs])
        nn = C.c_size_t()
        out = O.orc().orc_kway_dedupe(ptrs, szs, k, C.byref(nn))
        got = O._take(out, nn.value)
        assert np.array_equal(ref, got)
        assert np.array_equal(got, np.unique(np.concatenate(arrs)))


def test_sort_u64():
    rng = np.random.default_rng(3)
    for n in (0, 1, 2, 25, 1000, 200000):
        a = rng.integers(0, 2**64 - 1, size=n, dtype=np.uint64, endpoint=True)
        a[: n // 3] = a[0] if n else 0
        b = a.copy()
        O.orc().orc_sort_u64(b.ctypes.data_as(O.U64P), n)
        assert np.array_equal(b, np.sort(a))


# ---------------------------------------------------------------- recurrence

def py_canon(b):
    return max(O.all_rots(b))


def py_moves(cfg, b):
    out = set()
    for d in range(4):
        ok, m = O.move(cfg, b, d)
        if ok:
            out.add(py_canon(m))
    return out


def test_generate_brute_force_bfs():
    """Set-based restatement of the layer recurrence (generate.c:551-592) on a
    small formation, against the oracle's threaded sort/merge implementation."""
    cfg = O.make_cfg(cores=3, ignore_f=True)
    initial = [0x120000FFFFFFFFFF]   # six free cells: row 0 and half of row 1
    start = O.orc().orc_get_sum(initial[0])
    end = start + 40
    layers = O.generate(cfg, initial, start, end)
    P, Cc = set(initial), set()
    for i, s in enumerate(range(start, end + 1, 2)):
        M = set()
        for b in P:
            M |= py_moves(cfg, b)
        assert np.array_equal(layers[i], np.array(sorted(M), dtype=np.uint64)), s
        S2, S4 = set(), set()
        for m in M:
            for x in range(16):
                if py_tile(m, x) == 0:
                    S2.add(py_canon(m | (1 << (60 - 4 * x))))
                    S4.add(py_canon(m | (2 << (60 - 4 * x))))
        P, Cc = S2 | Cc, S4


# ---------------------------------------------------------------- solve

def test_solve_closed_forms():
    cfg = O.make_cfg(ignore_f=True)
    W = O.expand_winstates([0x7000000000000000])
    # top layer: no successors -> 1 if satisfied else 0
    M = np.array(sorted({0xFFFFFFFF70000000, 0xFFFFFFFF12340000}), dtype=np.uint64)
    V = O.solve_layer(cfg, M, [], [], [], [], W)
    for b, v in zip(M, V):
        assert v == (1.0 if any((int(b) & O.orc().orc_cmpbrd(int(b), int(w)) * 0) == 0 and
                                O.orc().orc_cmpbrd(int(b), int(w)) for w in W) else 0.0)
    # one empty cell whose every spawn leads (after any valid move) to a layer
    # where all successors are worth 1: value = 0.9 * 1 + 0.1 * 1
    b = 0xFFFFFFFF1234567F & ~0xF  # 2x4 free row pair with a single empty cell (tile 15)
    succ2 = set()
    succ4 = set()
    for v, S in ((1, succ2), (2, succ4)):
        for d in range(4):
            ok, m = O.move(cfg, b | v, d)
            if ok:
                S.add(O.canon(m))
    K2 = np.array(sorted(succ2), dtype=np.uint64)
    K4 = np.array(sorted(succ4), dtype=np.uint64)
    V = O.solve_layer(cfg, np.array([b], np.uint64), K2, np.ones(len(K2)), K4, np.ones(len(K4)),
                      np.zeros(0, np.uint64))
    exp = (0.9 * (1.0 if len(K2) else 0.0) / 1) + (0.1 * (1.0 if len(K4) else 0.0) / 1)
    assert V[0] == exp


@needs_ref
def test_expectimax_vs_reference(tmp_path):
    """expectimax (solve.c:257-281) on random boards against random solved
    successor tables, bit-exact."""
    L = O.ref()
    O.ref_configure(str(tmp_path), ignore_f=True)
    cfg = O.make_cfg(ignore_f=True)
    rng = random.Random(21)
    nrng = np.random.default_rng(21)
    Ms = [rand_board(rng, FORMATIONS["DPDGap"], p0=0.5, kmax=6) for _ in range(400)]
    succ = set()
    for b in Ms:
        for x in range(16):
            if py_tile(b, x) == 0:
                for v in (1, 2):
                    for d in range(4):
                        ok, m = O.move(cfg, b | (v << (60 - 4 * x)), d)
                        if ok and rng.random() < 0.7:
                            succ.add(O.canon(m))
    K = np.array(sorted(succ), dtype=np.uint64)
    V = nrng.random(len(K))
    half = len(K) // 2
    K2, V2, K4, V4 = K, V, K[half:], V[half:].copy()
    W = np.zeros(0, np.uint64)
    mine = O.solve_layer(cfg, np.array(Ms, np.uint64), K2, V2, K4, V4, W)
    t2 = O.Table(O.SArr(True, K2.ctypes.data_as(O.U64P), len(K2)),
                 O.SArr(True, V2.view(np.uint64).ctypes.data_as(O.U64P), len(K2)))
    t4 = O.Table(O.SArr(True, K4.ctypes.data_as(O.U64P), len(K4)),
                 O.SArr(True, V4.view(np.uint64).ctypes.data_as(O.U64P), len(K4)))
    ws = O.SArr(True, W.ctypes.data_as(O.U64P), 0)
    for b, v in zip(Ms, mine):
        r = L.expectimax(b, C.byref(t2), C.byref(t4), C.byref(ws), 0, False)
        assert np.float64(r).view(np.uint64) == np.float64(v).view(np.uint64)


# ---------------------------------------------------------------- end to end

@needs_ref
@pytest.mark.parametrize("variant", [
    dict(ignore_f=False), dict(ignore_f=True), dict(ignore_f=True, cores=3),
    dict(free_formation=True), dict(ignore_f=True, nox=7),
])
def test_generate_solve_vs_reference(tmp_path, variant):
    v = dict(variant)
    cores = v.pop("cores", 1)
    # free_formation lets the formation cells move: the layers grow far faster
    initial, win = [0x11000000FFFFFFFF], [0x7000000000000000]
    end = 36 if v.get("free_formation") else 120
    d, start = O.ref_generate_solve(initial, win, end, workdir=str(tmp_path), cores=cores, **v)
    cfg = O.make_cfg(cores=cores, **v)
    layers = O.generate(cfg, initial, start, end)
    vals = O.solve(cfg, layers, start, end, 0, win)
    for i, s in enumerate(range(start, end + 1, 2)):
        b = O.read_boards_file(f"{d}/b/{s}.boards")
        k, val = O.read_table_file(f"{d}/t/{s}.tables")
        assert np.array_equal(b, layers[i]), s
        assert np.array_equal(k, layers[i]), s
        assert np.array_equal(val.view(np.uint64), vals[i].view(np.uint64)), s


@needs_ref
def test_dpdgap_no_prune_vs_reference(tmp_path):
    initial, win, end = [0x00008765FFF3FFF2], [0x0000900000000000], 540
    d, start = O.ref_generate_solve(initial, win, end, workdir=str(tmp_path), cores=4, ignore_f=True)
    cfg = O.make_cfg(cores=4, ignore_f=True)
    layers = O.generate(cfg, initial, start, end)
    vals = O.solve(cfg, layers, start, end, 0, win)
    assert sum(len(x) for x in layers) > 1000
    for i, s in enumerate(range(start, end + 1, 2)):
        k, val = O.read_table_file(f"{d}/t/{s}.tables")
        assert np.array_equal(k, layers[i]), s
        assert np.array_equal(val.view(np.uint64), vals[i].view(np.uint64)), s


def test_c1_manifest_from_reference():
    """Committed manifest made by the reference build (works without it)."""
    man = json.load(open(os.path.join(HERE, "golden", "c1_ref.json")))
    initial, win = [0x11000000FFFFFFFF], [0x7000000000000000]
    for name, ig in (("default", False), ("ignore_f", True)):
        cfg = O.make_cfg(cores=4, ignore_f=ig)
        rows = man["variants"][name]
        start, end = rows[0][0], rows[-1][0]
        layers = O.generate(cfg, initial, start, end)
        vals = O.solve(cfg, layers, start, end, 0, win)
        assert len(rows) == len(layers) == 129
        for (s, n, hk, hv), k, v in zip(rows, layers, vals):
            assert n == len(k)
            assert "%016x" % O.hash_u64(k) == hk, s
            assert "%016x" % O.hash_u64(v.view(np.uint64)) == hv, s


# ---------------------------------------------------------------- adjoint

def test_adjoint_vs_finite_differences():
    cfg = O.make_cfg(cores=4, ignore_f=True)
    initial, win = [0x11000000FFFFFFFF], [0x7000000000000000]
    start, end = 4, 200
    layers = O.generate(cfg, initial, start, end)
    vals = O.solve(cfg, layers, start, end, 0, win, p4=0.1)
    P, dP, D = O.adjoint(cfg, layers, start, vals, win, initial, p4=0.1)
    assert P > 0.0 and abs(dP) > 0.0
    assert np.isclose(D.sum(), dP, rtol=1e-12, atol=0)

    def p_at(p4):
        v = O.solve(cfg, layers, start, end, 0, win, p4=p4)
        return O.adjoint(cfg, layers, start, v, win, initial, p4=p4)[0]

    # this table has an argmax switch (a kink of P(p4)) within 1e-5 of 0.1, so
    # the central difference is taken at 1e-6, where P is smooth on both sides
    h = 1e-6
    fd = (p_at(0.1 + h) - p_at(0.1 - h)) / (2 * h)
    assert abs(fd - dP) <= 1e-6 * abs(dP), (fd, dP)
    assert p_at(0.1) == P


def test_spawn_sets_consistent_with_gen_layer():
    cfg = O.make_cfg(cores=3, ignore_f=True, prune=True, stsl=58, ltc=5, smallest_large=6)
    P, Cc = np.array([0x00008765FFF3FFF2], np.uint64), np.zeros(0, np.uint64)
    for _ in range(12):
        M, P2, C2 = O.gen_layer(cfg, P, Cc)
        S2, S4 = O.spawn_sets(cfg, M)
        assert np.array_equal(P2, np.union1d(S2, Cc))
        assert np.array_equal(C2, S4)
        P, Cc = P2, C2


# ---------------------------------------------------------------- premove

@pytest.mark.parametrize("rules,initial,start,end", [
    (dict(ignore_f=True), [0x11000000FFFFFFFF], 4, 80),
    (dict(ignore_f=True, prune=True, stsl=58, ltc=5, smallest_large=6), [0x00008765FFF3FFF2], 492, 530),
])
def test_premove_identity(rules, initial, start, end):
    """premove (the initial boards are positions after the move; the first
    layer is their canonical set): started from the first layer of a plain
    run, it gives that run's layers exactly -- the plain recurrence's state
    after its first move step is the premove run's state at the start."""
    cfg = O.make_cfg(cores=4, **rules)
    plain = O.generate(cfg, initial, start, end)
    # the first layer's boards in non-canonical orientations, with repeats
    b0 = [O.orc().orc_rotate_cw(int(x)) for x in plain[0][::3]] + [int(x) for x in plain[0]]
    pre = O.generate(O.make_cfg(cores=4, premove=True, **rules), b0, start, end)
    assert len(pre) == len(plain)
    for a, b in zip(plain, pre):
        assert np.array_equal(a, b)


def test_premove_adjoint():
    """With premove the win probability is the mean value of the initial
    boards' canonical forms in the first layer, and its p4-derivative matches
    central differences."""
    rules = dict(ignore_f=True)
    cfgp = O.make_cfg(cores=4, premove=True, **rules)
    win, start, end = [0x7000000000000000], 4, 200
    initial = [0x1100000FFFFFFFF0 & 0xFFFFFFFFFFFFFFFF, 0x11000000FFFFFFFF, 0x10100000FFFFFFFF]
    initial = [x for x in initial if O.orc().orc_get_sum(x) == 4]
    layers = O.generate(cfgp, initial, start, end)
    vals = O.solve(cfgp, layers, start, end, 0, win, p4=0.1)
    P, dP, D = O.adjoint(cfgp, layers, start, vals, win, initial, p4=0.1)
    pos = np.searchsorted(layers[0], [O.canon(x) for x in initial])
    assert P == sum(vals[0][pos] / len(initial)) or np.isclose(P, vals[0][pos].mean(), rtol=1e-15)
    assert np.isclose(D.sum(), dP, rtol=1e-12, atol=0)

    def p_at(p4):
        v = O.solve(cfgp, layers, start, end, 0, win, p4=p4)
        return O.adjoint(cfgp, layers, start, v, win, initial, p4=p4)[0]

    h = 1e-6
    fd = (p_at(0.1 + h) - p_at(0.1 - h)) / (2 * h)
    assert abs(fd - dP) <= 1e-6 * abs(dP), (fd, dP)
