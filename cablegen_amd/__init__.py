"""cablegen_amd -- Python access to the MI355X tablebase engine (libcablegen.so).

Mirrors the reference's C interface for this path (api/cablegen.h: generate,
solve, lookup, settings) plus the engine's extensions (include/cablegen_gpu.h:
device-resident tablebase, adjoint, device primitives, kernel timers).  All
computation runs in the HIP library; if it has not been built, loading raises.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcablegen.so")

U64P = C.POINTER(C.c_uint64)
F64P = C.POINTER(C.c_double)


class Rules(C.Structure):
    """cg_rules: the settings that shape layers and values (inc/settings.h:15-31)."""
    _fields_ = [("free_formation", C.c_int), ("ignore_f", C.c_int), ("prune", C.c_int),
                ("stsl", C.c_long), ("ltc", C.c_long), ("smallest_large", C.c_long),
                ("nox", C.c_int), ("score", C.c_int), ("premove", C.c_int)]


def rules(free_formation=False, ignore_f=False, prune=False, stsl=200, ltc=5, smallest_large=6,
          nox=0, score=False, premove=False):
    """Defaults are the reference's (src/settings.c:35-59).  premove: the
    initial boards are positions after the move (the first layer is their
    canonical set)."""
    return Rules(int(free_formation), int(ignore_f), int(prune), int(stsl), int(ltc),
                 int(smallest_large), int(nox), int(score), int(premove))


class SArr(C.Structure):
    _fields_ = [("valid", C.c_bool), ("bp", U64P), ("size", C.c_size_t)]


class Table(C.Structure):
    _fields_ = [("key", SArr), ("value", SArr)]


class MinSettings(C.Structure):
    _fields_ = [("free_formation", C.c_bool), ("ignore_f", C.c_bool), ("cores", C.c_longlong),
                ("nox", C.c_longlong), ("bdir", C.c_char_p), ("initial", C.c_char_p),
                ("end_gen", C.c_longlong), ("tdir", C.c_char_p), ("winstates", C.c_char_p),
                ("end_solve", C.c_longlong), ("delete_boards", C.c_bool)]


class CablegenError(RuntimeError):
    pass


_lib = None


def lib():
    """Loads libcablegen.so; raises if it is missing (there is no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise CablegenError(f"{LIB_PATH} is not built; run python -m cablegen_amd.build")
    L = C.CDLL(LIB_PATH)
    L.cg_last_error.restype = C.c_char_p
    L.cg_tb_create.restype = C.c_void_p
    L.cg_tb_create.argtypes = [C.POINTER(Rules)]
    L.cg_tb_free.argtypes = [C.c_void_p]
    L.cg_tb_generate.argtypes = [C.c_void_p, U64P, C.c_size_t, C.c_int, C.c_int]
    L.cg_tb_generate_bounded.argtypes = [C.c_void_p, U64P, C.c_size_t, C.c_int, C.c_int, C.c_size_t,
                                         C.POINTER(C.c_int)]
    L.cg_tb_solve.argtypes = [C.c_void_p, U64P, C.c_size_t, C.c_int, C.c_double, C.c_int]
    L.cg_tb_adjoint.argtypes = [C.c_void_p, U64P, C.c_size_t, F64P, F64P, F64P]
    L.cg_tb_num_layers.argtypes = [C.c_void_p]
    L.cg_tb_start.argtypes = [C.c_void_p]
    L.cg_tb_layer_size.restype = C.c_size_t
    L.cg_tb_layer_size.argtypes = [C.c_void_p, C.c_int]
    L.cg_tb_total_boards.restype = C.c_size_t
    L.cg_tb_total_boards.argtypes = [C.c_void_p]
    L.cg_tb_copy_layer.argtypes = [C.c_void_p, C.c_int, U64P, F64P]
    L.cg_tb_layer_device.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_void_p),
                                     C.POINTER(C.c_void_p)]
    L.cg_tb_write_files.argtypes = [C.c_void_p, C.c_char_p, C.c_char_p]
    L.cg_dev_move.argtypes = [C.POINTER(Rules), C.c_void_p, C.c_size_t, C.c_void_p,
                              C.POINTER(C.c_size_t)]
    L.cg_dev_spawn.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p,
                               C.POINTER(C.c_size_t)]
    L.cg_dev_sort_unique.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.POINTER(C.c_size_t)]
    L.cg_dev_sort_unique_bounded.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p,
                                             C.POINTER(C.c_size_t), C.c_uint64]
    L.cg_sort_code_probe.argtypes = [C.c_void_p, C.c_size_t, C.c_uint64, C.c_uint64, C.c_void_p,
                                     C.c_void_p]
    L.cg_layer_index_probe.argtypes = [C.c_size_t, C.c_uint64, C.c_uint64, C.POINTER(C.c_int)]
    L.cg_canon_sel_probe.argtypes = [C.c_uint64, C.c_int, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p,
                                     C.c_void_p]
    L.cg_dev_lookup.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t,
                                C.c_void_p, C.c_int]
    L.cg_dev_lookup_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t,
                                        C.c_void_p, C.c_int]
    L.cg_dev_solve_layer.argtypes = [C.POINTER(Rules), C.c_void_p, C.c_size_t, C.c_void_p,
                                     C.c_void_p, C.c_void_p, C.c_size_t,
                                     C.c_void_p, C.c_void_p, C.c_size_t,
                                     U64P, C.c_size_t, C.c_double]
    L.cg_dev_random_boards.argtypes = [C.c_void_p, C.c_size_t, C.c_uint64, C.c_uint64, C.c_double,
                                       C.c_int, C.c_int]
    L.cg_dev_canon.argtypes = [C.c_void_p, C.c_size_t]
    L.cg_batch_create.restype = C.c_void_p
    L.cg_batch_create.argtypes = [C.POINTER(Rules)]
    L.cg_batch_free.argtypes = [C.c_void_p]
    L.cg_batch_add.argtypes = [C.c_void_p, U64P, C.c_size_t, C.c_int, C.c_int, U64P, C.c_size_t]
    L.cg_batch_generate.argtypes = [C.c_void_p]
    L.cg_batch_solve.argtypes = [C.c_void_p, C.c_double, C.c_int]
    L.cg_batch_num_tables.argtypes = [C.c_void_p]
    L.cg_batch_num_layers.argtypes = [C.c_void_p, C.c_int]
    L.cg_batch_layer_size.restype = C.c_size_t
    L.cg_batch_layer_size.argtypes = [C.c_void_p, C.c_int, C.c_int]
    L.cg_batch_total_boards.restype = C.c_size_t
    L.cg_batch_total_boards.argtypes = [C.c_void_p]
    L.cg_batch_copy_layer.argtypes = [C.c_void_p, C.c_int, C.c_int, U64P, F64P]
    L.cg_tune_set.argtypes = [C.c_char_p, C.c_int]
    L.cg_tune_get.argtypes = [C.c_char_p]
    L.cg_tune_sort_entry.argtypes = [C.c_int] * 5
    L.cg_tune_sort_query.argtypes = [C.c_size_t, C.c_int, C.POINTER(C.c_int),
                                     C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.cg_autotune.argtypes = [C.c_int, C.c_char_p, C.c_size_t]
    L.cg_tune_load.argtypes = [C.c_char_p]
    L.cg_prof_enable.argtypes = [C.c_int]
    L.cg_prof_query.restype = C.c_long
    L.cg_prof_query.argtypes = [C.c_char_p, F64P, F64P]
    L.cg_prof_select.argtypes = [C.c_char_p]
    L.cg_prof_query_stream.restype = C.c_long
    L.cg_prof_query_stream.argtypes = [C.c_char_p, C.c_int, F64P, F64P]
    # drop-in API (api/cablegen.h)
    L.movedir.restype = C.c_bool
    L.movedir.argtypes = [U64P, C.c_int]
    L.lookup.restype = C.c_double
    L.lookup.argtypes = [C.c_uint64, C.POINTER(Table), C.c_bool]
    L.get_sum.restype = C.c_int
    L.get_sum.argtypes = [C.c_uint64]
    L.change_config.argtypes = [C.c_char_p]
    L.get_settings_min.restype = C.POINTER(MinSettings)
    L.generate.argtypes = [C.c_int, C.c_int, C.c_char_p, C.POINTER(SArr)]
    L.solve.argtypes = [C.c_uint, C.c_uint, C.c_char_p, C.c_char_p, C.POINTER(SArr)]
    L.test.restype = C.c_bool
    L.get_version.restype = C.c_char_p
    L.cge_build_row_table.argtypes = [C.c_int, C.c_void_p]
    _lib = L
    return L


def last_error():
    """the library's message for its last failed call"""
    return lib().cg_last_error().decode()


def _check(rc, what):
    if rc != 0:
        raise CablegenError(f"{what}: {lib().cg_last_error().decode()}")


def _u64(a):
    a = np.ascontiguousarray(a, dtype=np.uint64)
    return a, a.ctypes.data_as(U64P)


def win_images(winstates):
    """The 8 D4 images of each winstate (get_all_rots, src/board.c:349-374)."""
    out = []
    M = (1 << 64) - 1

    def cw(b):
        b = (((b & 0xff00ff0000000000) >> 8) | ((b & 0x00ff00ff00000000) >> 32) |
             ((b & 0x00000000ff00ff00) << 32) | ((b & 0x0000000000ff00ff) << 8)) & M
        return (((b & 0xf0f00000f0f00000) >> 4) | ((b & 0x0f0f00000f0f0000) >> 16) |
                ((b & 0x0000f0f00000f0f0) << 16) | ((b & 0x00000f0f00000f0f) << 4)) & M

    def flip(b):
        b = ((b >> 32) | (b << 32)) & M
        return (((b & 0xffff0000ffff0000) >> 16) | ((b & 0x0000ffff0000ffff) << 16)) & M

    for w in winstates:
        w = int(w)
        x = w
        out.append(w)
        for _ in range(3):
            x = cw(x)
            out.append(x)
        x = flip(w)
        out.append(x)
        for _ in range(3):
            x = cw(x)
            out.append(x)
    return np.array(out, dtype=np.uint64)


class Tablebase:
    """Device-resident tablebase (cg_tb_*): every layer stays in HBM."""

    def __init__(self, r):
        self.L = lib()
        self.r = r
        self.h = self.L.cg_tb_create(C.byref(r))
        if not self.h:
            raise CablegenError(self.L.cg_last_error().decode())

    def close(self):
        if self.h:
            self.L.cg_tb_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def generate(self, initial, start, end):
        a, p = _u64(initial)
        _check(self.L.cg_tb_generate(self.h, p, len(a), int(start), int(end)), "cg_tb_generate")

    def generate_bounded(self, initial, start, end, max_boards):
        """generates up to end, or stops after the first layer that takes the
        total above max_boards; returns the tile sum of the last layer"""
        a, p = _u64(initial)
        last = C.c_int()
        _check(self.L.cg_tb_generate_bounded(self.h, p, len(a), int(start), int(end), int(max_boards),
                                             C.byref(last)), "cg_tb_generate_bounded")
        return last.value

    def solve(self, winstates, end_solve=0, p4=0.1, keep_argmax=False):
        a, p = _u64(winstates)
        _check(self.L.cg_tb_solve(self.h, p, len(a), int(end_solve), float(p4), int(keep_argmax)),
               "cg_tb_solve")

    def adjoint(self, initial):
        a, p = _u64(initial)
        P, dP = C.c_double(), C.c_double()
        D = np.zeros(max(self.num_layers, 1), dtype=np.float64)
        _check(self.L.cg_tb_adjoint(self.h, p, len(a), C.byref(P), C.byref(dP),
                                    D.ctypes.data_as(F64P)), "cg_tb_adjoint")
        return P.value, dP.value, D[:self.num_layers]

    @property
    def num_layers(self):
        return self.L.cg_tb_num_layers(self.h)

    @property
    def start(self):
        return self.L.cg_tb_start(self.h)

    def layer_size(self, i):
        return self.L.cg_tb_layer_size(self.h, int(i))

    def total_boards(self):
        return self.L.cg_tb_total_boards(self.h)

    def keys(self, i):
        n = self.layer_size(i)
        k = np.zeros(n, dtype=np.uint64)
        _check(self.L.cg_tb_copy_layer(self.h, int(i), k.ctypes.data_as(U64P), None), "copy")
        return k

    def values(self, i):
        n = self.layer_size(i)
        v = np.zeros(n, dtype=np.float64)
        _check(self.L.cg_tb_copy_layer(self.h, int(i), None, v.ctypes.data_as(F64P)), "copy")
        return v

    def write_files(self, boards_fmt=None, tables_fmt=None):
        _check(self.L.cg_tb_write_files(self.h, boards_fmt.encode() if boards_fmt else None,
                                        tables_fmt.encode() if tables_fmt else None), "write")


def tune_set(knob, value):
    _check(lib().cg_tune_set(knob.encode(), int(value)), "cg_tune_set")


def tune_get(knob):
    return lib().cg_tune_get(knob.encode())


def tune_sort_query(n, nbits):
    """(digit bits, keys per thread, look-back window) a sort of n keys with
    nbits varying bits uses."""
    b, k, w = C.c_int(), C.c_int(), C.c_int()
    _check(lib().cg_tune_sort_query(int(n), int(nbits), C.byref(b), C.byref(k), C.byref(w)),
           "cg_tune_sort_query")
    return b.value, k.value, w.value


TUNED_TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned", "mi355x.json")


def apply_tuning(rep):
    """Installs the picks of an autotuner report (sort table, directory depth)."""
    L = lib()
    tune_set("builtin", 0)
    for cls, c in enumerate(rep["sort"]):
        for nb, cfg in c["pick"].items():
            bits, rest = cfg.split("x")
            kpt, lbw = rest.split("/")
            _check(L.cg_tune_sort_entry(cls, int(nb), int(bits), int(kpt), int(lbw)),
                   "cg_tune_sort_entry")
    tune_set("dir_delta", rep["dir"]["pick"])


def load_tuning(path=TUNED_TABLE):
    """Reads and installs an autotuner report saved as JSON; returns it."""
    import json
    with open(path) as f:
        rep = json.load(f)
    apply_tuning(rep)
    return rep


def autotune(effort=0):
    """Runs the autotuner (sort configuration per size class and key width,
    directory depth) on the GPU, installs its picks and returns its report."""
    import json
    buf = C.create_string_buffer(1 << 16)
    _check(lib().cg_autotune(int(effort), buf, len(buf)), "cg_autotune")
    return json.loads(buf.value.decode())


class Batch:
    """Many small tables generated and solved together (cg_batch_*)."""

    def __init__(self, r):
        self.L = lib()
        self.h = self.L.cg_batch_create(C.byref(r))
        if not self.h:
            raise CablegenError(self.L.cg_last_error().decode())

    def close(self):
        if self.h:
            self.L.cg_batch_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, initial, start, end, winstates):
        a, p = _u64(initial)
        w, wp = _u64(winstates)
        t = self.L.cg_batch_add(self.h, p, len(a), int(start), int(end), wp, len(w))
        if t < 0:
            raise CablegenError(self.L.cg_last_error().decode())
        return t

    def generate(self):
        _check(self.L.cg_batch_generate(self.h), "cg_batch_generate")

    def solve(self, p4=0.1, use_graph=True):
        _check(self.L.cg_batch_solve(self.h, float(p4), int(use_graph)), "cg_batch_solve")

    @property
    def num_tables(self):
        return self.L.cg_batch_num_tables(self.h)

    def num_layers(self, t):
        return self.L.cg_batch_num_layers(self.h, int(t))

    def layer_size(self, t, i):
        return self.L.cg_batch_layer_size(self.h, int(t), int(i))

    def total_boards(self):
        return self.L.cg_batch_total_boards(self.h)

    def total_boards_table(self, t):
        return sum(self.layer_size(t, i) for i in range(self.num_layers(t)))

    def keys(self, t, i):
        k = np.zeros(self.layer_size(t, i), dtype=np.uint64)
        _check(self.L.cg_batch_copy_layer(self.h, int(t), int(i), k.ctypes.data_as(U64P), None),
               "copy")
        return k

    def values(self, t, i):
        v = np.zeros(self.layer_size(t, i), dtype=np.float64)
        _check(self.L.cg_batch_copy_layer(self.h, int(t), int(i), None, v.ctypes.data_as(F64P)),
               "copy")
        return v


def prof_enable(on=True):
    lib().cg_prof_enable(int(on))


def prof_reset():
    lib().cg_prof_reset()


def prof_query(kernel):
    """(launches, total_ms, total_algorithmic_bytes) for one kernel class."""
    ms, b = C.c_double(), C.c_double()
    n = lib().cg_prof_query(kernel.encode(), C.byref(ms), C.byref(b))
    return n, ms.value, b.value


def prof_select(kernel=None):
    """time only one kernel's launches (None: every kernel; "<kernel>@main":
    its launches on the main stream only)"""
    if lib().cg_prof_select(kernel.encode() if kernel else None) != 0:
        raise RuntimeError("cg_prof_select(%r) failed" % (kernel,))


def prof_query_stream(kernel, stream):
    """prof_query for the launches of one stream (0 main, 1 carry-sort stream)."""
    ms, b = C.c_double(), C.c_double()
    n = lib().cg_prof_query_stream(kernel.encode(), stream, C.byref(ms), C.byref(b))
    return n, ms.value, b.value


def row_table(ignore_f=False):
    t = np.zeros(65536, dtype=np.uint16)
    lib().cge_build_row_table(int(ignore_f), t.ctypes.data)
    return t
