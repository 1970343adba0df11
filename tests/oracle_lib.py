"""ctypes access to the oracle (oracle/liborc.so) and, where it was built, the
reference library compiled from its own sources (oracle/_ref/libcablegen_ref.so).

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module.  Both libraries are checkers; neither is
ever the thing measured or shipped.
"""
import ctypes as C
import os
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORC_PATH = os.path.join(ROOT, "oracle", "liborc.so")
REF_PATH = os.path.join(ROOT, "oracle", "_ref", "libcablegen_ref.so")

U64P = C.POINTER(C.c_uint64)
F64P = C.POINTER(C.c_double)


class OrcCfg(C.Structure):
    _fields_ = [
        ("free_formation", C.c_int), ("ignore_f", C.c_int), ("cores", C.c_int),
        ("nox", C.c_int), ("prune", C.c_int), ("stsl", C.c_long), ("ltc", C.c_long),
        ("smallest_large", C.c_long), ("score", C.c_int), ("premove", C.c_int),
    ]


def make_cfg(free_formation=False, ignore_f=False, cores=1, nox=0, prune=False,
             stsl=200, ltc=5, smallest_large=6, score=False, premove=False):
    return OrcCfg(int(free_formation), int(ignore_f), int(cores), int(nox), int(prune),
                  int(stsl), int(ltc), int(smallest_large), int(score), int(premove))


def _u64(a):
    a = np.ascontiguousarray(a, dtype=np.uint64)
    return a, a.ctypes.data_as(U64P)


_orc = None


def orc():
    """Load the oracle library (raises if it has not been built)."""
    global _orc
    if _orc is None:
        if not os.path.exists(ORC_PATH):
            raise FileNotFoundError(f"{ORC_PATH} not built (run make -f oracle/Makefile)")
        L = C.CDLL(ORC_PATH)
        L.orc_build_lut.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p]
        for f in ("orc_rotate_cw", "orc_flip", "orc_canon"):
            getattr(L, f).restype = C.c_uint64
            getattr(L, f).argtypes = [C.c_uint64]
        L.orc_all_rots.argtypes = [C.c_uint64, U64P]
        L.orc_get_sum.restype = C.c_int
        L.orc_get_sum.argtypes = [C.c_uint64]
        L.orc_move.restype = C.c_int
        L.orc_move.argtypes = [C.POINTER(OrcCfg), U64P, C.c_int]
        L.orc_prune.restype = C.c_int
        L.orc_prune.argtypes = [C.c_uint64, C.c_long, C.c_long, C.c_long]
        L.orc_cmpbrd.restype = C.c_int
        L.orc_cmpbrd.argtypes = [C.c_uint64, C.c_uint64]
        L.orc_lookup.restype = C.c_double
        L.orc_lookup.argtypes = [C.c_uint64, U64P, F64P, C.c_size_t, C.c_int]
        L.orc_sort_u64.argtypes = [U64P, C.c_size_t]
        L.orc_kway_dedupe.restype = U64P
        L.orc_kway_dedupe.argtypes = [C.POINTER(U64P), C.POINTER(C.c_size_t), C.c_size_t,
                                      C.POINTER(C.c_size_t)]
        L.orc_gen_layer.argtypes = [C.POINTER(OrcCfg), U64P, C.c_size_t, U64P, C.c_size_t,
                                    C.POINTER(U64P), C.POINTER(C.c_size_t),
                                    C.POINTER(U64P), C.POINTER(C.c_size_t),
                                    C.POINTER(U64P), C.POINTER(C.c_size_t)]
        L.orc_spawn_sets.argtypes = [C.POINTER(OrcCfg), U64P, C.c_size_t, C.POINTER(U64P),
                                     C.POINTER(C.c_size_t), C.POINTER(U64P), C.POINTER(C.c_size_t)]
        L.orc_generate.argtypes = [C.POINTER(OrcCfg), U64P, C.c_size_t, C.c_int, C.c_int,
                                   C.POINTER(C.POINTER(U64P)), C.POINTER(C.POINTER(C.c_size_t)),
                                   C.POINTER(C.c_int)]
        L.orc_solve_layer.argtypes = [C.POINTER(OrcCfg), U64P, C.c_size_t, U64P, F64P,
                                      C.c_size_t, U64P, F64P, C.c_size_t, U64P, C.c_size_t,
                                      C.c_double, F64P]
        L.orc_solve.argtypes = [C.POINTER(OrcCfg), C.POINTER(U64P), C.POINTER(C.c_size_t),
                                C.c_int, C.c_int, C.c_int, C.c_int, U64P, C.c_size_t,
                                C.c_double, C.POINTER(F64P)]
        L.orc_expand_winstates.restype = C.c_size_t
        L.orc_expand_winstates.argtypes = [U64P, C.c_size_t, U64P]
        L.orc_adjoint.argtypes = [C.POINTER(OrcCfg), C.POINTER(U64P), C.POINTER(C.c_size_t),
                                  C.c_int, C.c_int, C.POINTER(F64P), U64P, C.c_size_t,
                                  U64P, C.c_size_t, C.c_double, F64P, F64P, F64P]
        L.orc_hash_u64.restype = C.c_uint64
        L.orc_hash_u64.argtypes = [U64P, C.c_size_t]
        L.orc_free.argtypes = [C.c_void_p]
        _orc = L
    return _orc


def _take(ptr, n, dtype=np.uint64):
    """Copy a malloc'd C array into numpy and free it."""
    if n:
        out = np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)
    else:
        out = np.zeros(0, dtype=dtype)
    orc().orc_free(C.cast(ptr, C.c_void_p))
    return out


def lut(free_formation=False, ignore_f=False):
    mv = np.zeros((2, 65536), dtype=np.uint16)
    lk = np.zeros((2, 65536), dtype=np.uint8)
    orc().orc_build_lut(int(free_formation), int(ignore_f), mv.ctypes.data, lk.ctypes.data)
    return mv, lk


def canon(b):
    return orc().orc_canon(int(b))


def all_rots(b):
    out = (C.c_uint64 * 8)()
    orc().orc_all_rots(int(b), out)
    return list(out)


def move(cfg, b, d):
    x = C.c_uint64(int(b))
    ok = orc().orc_move(C.byref(cfg), C.byref(x), int(d))
    return bool(ok), x.value


def expand_winstates(ws):
    ws, p = _u64(ws)
    out = np.zeros(8 * len(ws), dtype=np.uint64)
    orc().orc_expand_winstates(p, len(ws), out.ctypes.data_as(U64P))
    return out


def gen_layer(cfg, P, Cset):
    P, pp = _u64(P)
    Cset, cp = _u64(Cset)
    M, P2, C2 = U64P(), U64P(), U64P()
    nM, nP2, nC2 = C.c_size_t(), C.c_size_t(), C.c_size_t()
    orc().orc_gen_layer(C.byref(cfg), pp, len(P), cp, len(Cset), C.byref(M), C.byref(nM),
                        C.byref(P2), C.byref(nP2), C.byref(C2), C.byref(nC2))
    return _take(M, nM.value), _take(P2, nP2.value), _take(C2, nC2.value)


def spawn_sets(cfg, M):
    """(S2, S4): sorted unique spawn sets of a post-move layer."""
    M, mp = _u64(M)
    S2, S4 = U64P(), U64P()
    n2, n4 = C.c_size_t(), C.c_size_t()
    orc().orc_spawn_sets(C.byref(cfg), mp, len(M), C.byref(S2), C.byref(n2), C.byref(S4),
                         C.byref(n4))
    return _take(S2, n2.value), _take(S4, n4.value)


def generate(cfg, initial, start, end):
    """Returns a list of sorted unique layers M_s, s = start, start+2, ..., end."""
    init, ip = _u64(initial)
    L = C.POINTER(U64P)()
    S = C.POINTER(C.c_size_t)()
    nl = C.c_int()
    orc().orc_generate(C.byref(cfg), ip, len(init), int(start), int(end), C.byref(L),
                       C.byref(S), C.byref(nl))
    layers = [_take(L[i], S[i]) for i in range(nl.value)]
    orc().orc_free(C.cast(L, C.c_void_p))
    orc().orc_free(C.cast(S, C.c_void_p))
    return layers


def solve_layer(cfg, M, K2, V2, K4, V4, W, p4=0.1):
    M, mp = _u64(M)
    K2, k2p = _u64(K2)
    K4, k4p = _u64(K4)
    W, wp = _u64(W)
    V2 = np.ascontiguousarray(V2, dtype=np.float64)
    V4 = np.ascontiguousarray(V4, dtype=np.float64)
    V = np.zeros(len(M), dtype=np.float64)
    orc().orc_solve_layer(C.byref(cfg), mp, len(M), k2p, V2.ctypes.data_as(F64P), len(K2),
                          k4p, V4.ctypes.data_as(F64P), len(K4), wp, len(W), float(p4),
                          V.ctypes.data_as(F64P))
    return V


def solve(cfg, layers, start_gen, end_gen, end_solve, winstates, p4=0.1):
    """Solves the stored layers; returns a list of value arrays (None outside range)."""
    nl = len(layers)
    keep = [np.ascontiguousarray(x, dtype=np.uint64) for x in layers]
    LP = (U64P * max(nl, 1))(*[k.ctypes.data_as(U64P) for k in keep])
    SZ = (C.c_size_t * max(nl, 1))(*[len(k) for k in keep])
    ws, wp = _u64(winstates)
    VP = (F64P * max(nl, 1))()
    orc().orc_solve(C.byref(cfg), LP, SZ, nl, int(start_gen), int(end_gen), int(end_solve),
                    wp, len(ws), float(p4), VP)
    out = []
    for i in range(nl):
        if VP[i]:
            out.append(_take(VP[i], len(keep[i]), np.float64))
        else:
            out.append(None)
    return out


def adjoint(cfg, layers, start_gen, values, winstates, initial, p4=0.1):
    """Returns (P, dP/dp4, per-layer D array)."""
    nl = len(layers)
    keep = [np.ascontiguousarray(x, dtype=np.uint64) for x in layers]
    vals = [None if v is None else np.ascontiguousarray(v, dtype=np.float64) for v in values]
    LP = (U64P * max(nl, 1))(*[k.ctypes.data_as(U64P) for k in keep])
    SZ = (C.c_size_t * max(nl, 1))(*[len(k) for k in keep])
    VP = (F64P * max(nl, 1))(*[(v.ctypes.data_as(F64P) if v is not None else F64P())
                               for v in vals])
    W = expand_winstates(winstates)
    init, ip = _u64(initial)
    P, dP = C.c_double(), C.c_double()
    D = np.zeros(max(nl, 1), dtype=np.float64)
    orc().orc_adjoint(C.byref(cfg), LP, SZ, nl, int(start_gen), VP, W.ctypes.data_as(U64P),
                      len(W), ip, len(init), float(p4), C.byref(P), C.byref(dP),
                      D.ctypes.data_as(F64P))
    return P.value, dP.value, D[:nl]


def hash_u64(a):
    a, p = _u64(a)
    return orc().orc_hash_u64(p, len(a))


# --------------------------------------------------------------------------
# reference library (built from /root/reference's sources by oracle/Makefile.ref)
# --------------------------------------------------------------------------

class SArr(C.Structure):
    _fields_ = [("valid", C.c_bool), ("bp", U64P), ("size", C.c_size_t)]


class DArr(C.Structure):
    _fields_ = [("valid", C.c_bool), ("bp", U64P), ("sp", U64P), ("size", C.c_size_t)]


class Table(C.Structure):
    _fields_ = [("key", SArr), ("value", SArr)]


_ref = None


def have_ref():
    return os.path.exists(REF_PATH)


def ref():
    """Load the reference library built from its own sources, or None."""
    global _ref
    if _ref is None and have_ref():
        L = C.CDLL(REF_PATH)
        L.set_log_level(4)
        L.canonicalize_b.argtypes = [U64P]
        L.rotate_clockwise.argtypes = [U64P]
        L.flip.argtypes = [U64P]
        L.movedir_unstable.restype = C.c_bool
        L.movedir_unstable.argtypes = [U64P, C.c_int]
        L.movedir.restype = C.c_bool
        L.movedir.argtypes = [U64P, C.c_int]
        L.prune_board.restype = C.c_bool
        L.prune_board.argtypes = [C.c_uint64, C.c_long, C.c_long, C.c_long]
        L.get_sum.restype = C.c_int
        L.get_sum.argtypes = [C.c_uint64]
        L.cmpbrd.restype = C.c_bool
        L.cmpbrd.argtypes = [C.c_uint64, C.c_uint64]
        L.lookup.restype = C.c_double
        L.lookup.argtypes = [C.c_uint64, C.POINTER(Table), C.c_bool]
        L.get_all_rots.restype = U64P
        L.get_all_rots.argtypes = [C.c_uint64]
        L.expectimax.restype = C.c_double
        L.expectimax.argtypes = [C.c_uint64, C.POINTER(Table), C.POINTER(Table), C.POINTER(SArr),
                                 C.c_char, C.c_bool]
        L.deduplicate_threads.restype = DArr
        L.deduplicate_threads.argtypes = [C.POINTER(DArr), C.c_size_t]
        L.change_config.argtypes = [C.c_char_p]
        L.generate.argtypes = [C.c_int, C.c_int, C.c_char_p, C.POINTER(SArr)]
        L.solve.argtypes = [C.c_uint, C.c_uint, C.c_char_p, C.c_char_p, C.POINTER(SArr)]
        _ref = L
    return _ref


def ref_configure(workdir, free_formation=False, ignore_f=False, cores=1, nox=0, prune=False,
                  stsl=200, ltc=5, smallest_large=6, end_gen=1200, score=False):
    """Writes a full config file and points the reference's settings at it
    (change_config re-reads every key on the next get_settings())."""
    tf = lambda v: "true" if v else "false"
    path = os.path.join(workdir, "cablegen.conf")
    with open(path, "w") as f:
        f.write(f"""[Cablegen]
free_formation = {tf(free_formation)}
ignore_f = {tf(ignore_f)}
cores = {cores}
nox = {nox}
use_gpu = false
[Generate]
premove = false
dir = {workdir}/b/
end = {end_gen}
stsl = {stsl}
ltc = {ltc}
smallest_large = {smallest_large}
prune = {tf(prune)}
[Solve]
dir = {workdir}/t/
end = 0
score = {tf(score)}
delete_boards = false
""")
    ref().change_config(path.encode())
    ref().generate_lut()


def ref_generate_solve(initial, winstates, end_gen, end_solve=0, workdir=None, **cfg):
    """Runs the reference's generate() then solve() into workdir; returns
    (workdir, start) with %d.boards and %d.tables files."""
    L = ref()
    if workdir is None:
        workdir = tempfile.mkdtemp(prefix="cgref_")
    os.makedirs(workdir + "/b", exist_ok=True)
    os.makedirs(workdir + "/t", exist_ok=True)
    ref_configure(workdir, end_gen=end_gen, **cfg)
    init, ip = _u64(initial)
    start = L.get_sum(int(init[0]))
    L.generate(start, end_gen, f"{workdir}/b/%d.boards".encode(), C.byref(SArr(True, ip, len(init))))
    ws, wp = _u64(winstates)
    L.solve(end_gen, end_solve, f"{workdir}/b/%d.boards".encode(),
            f"{workdir}/t/%d.tables".encode(), C.byref(SArr(True, wp, len(ws))))
    return workdir, start


def read_boards_file(path):
    return np.fromfile(path, dtype=np.uint64) if os.path.exists(path) else np.zeros(0, np.uint64)


def read_table_file(path):
    raw = np.fromfile(path, dtype=np.uint64) if os.path.exists(path) else np.zeros(0, np.uint64)
    n = len(raw) // 2
    return raw[:n].copy(), raw[n:2 * n].view(np.float64).copy()


# --------------------------------------------------------------------------
# table consumers: the CLI's best() and lookup output (frontends/cli/main.c)
# --------------------------------------------------------------------------

MOVE_NAMES = ("left", "right", "up", "down")


def tile(b, i):
    return (int(b) >> (60 - 4 * i)) & 0xF


def set_tile(b, i, v):
    sh = 60 - 4 * i
    return (int(b) & ~(0xF << sh)) | (int(v) << sh)


def tile_sum(b):
    return sum(1 << tile(b, i) for i in range(16) if 0 < tile(b, i) < 0xE)


def table_lookup(b, K, V):
    """lookup(b, t, canonicalize=true) (solve.c:108-164): 0.0 on a miss"""
    c = canon(b)
    i = int(np.searchsorted(K, np.uint64(c)))
    return float(V[i]) if i < len(K) and int(K[i]) == c else 0.0


def best(cfg, b, K, V):
    """main.c:174-189: the first move whose value is strictly larger wins,
    starting from left at 0"""
    maxd, maxp = 0, 0.0
    for d in range(4):
        ok, t = move(cfg, b, d)
        if not ok:
            continue
        p = table_lookup(t, K, V)
        if p > maxp:
            maxd, maxp = d, p
    return maxd, maxp


def lookup_lines(cfg, q, tabs, spawn=False):
    """The "Best move" lines of the CLI's lookup (spawn=False) or
    lookup_spawn (spawn=True) of board q (main.c:191-265); tabs maps a tile
    sum to its (keys, values)."""
    b = int(q)
    if spawn:
        for i in range(16):
            if tile(b, i) <= 2:
                b = set_tile(b, i, 0)
                break
    s = tile_sum(b)
    empty = (np.zeros(0, np.uint64), np.zeros(0, np.float64))
    out = []
    for i in range(16):
        if tile(b, i):
            continue
        for v in (1, 2):
            c = set_tile(b, i, v)
            if spawn and c != int(q):
                continue
            K, V = tabs.get(s + 2 * v, empty)
            md, mp = best(cfg, c, K, V)
            out.append("Best move: %s (%.10f)" % (MOVE_NAMES[md], mp))
    return out
