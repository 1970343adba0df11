"""The reference's own CLI (frontends/cli/main.c, unchanged, linked to this
library) as a consumer of tables: its `lookup` command reads the `%d.tables`
files of the board's sum, sum + 2 and sum + 4 and prints, for every empty
cell and spawn, the best move and its probability (`best()`, main.c:174-189:
a move wins only when strictly better, starting from left at 0).

The tables here are written by the oracle, so the test runs on the CPU: it
pins the library's host `read_table`, `lookup`, `movedir` and `get_sum` as the
CLI uses them.  The expected lines come from a restatement of `best()` over the
oracle's moves and canonical forms."""
import os
import subprocess

import numpy as np
import pytest

from tests import oracle_lib as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "oracle", "_ref", "cablegen_cli_mi355x")
NAMES = O.MOVE_NAMES

pytestmark = pytest.mark.skipif(not os.path.exists(CLI), reason="reference CLI not built")


tile, set_tile, best = O.tile, O.set_tile, O.best


@pytest.fixture(scope="module")
def tables(tmp_path_factory):
    d = tmp_path_factory.mktemp("cli")
    (d / "t").mkdir()
    cfg = O.make_cfg(ignore_f=True, cores=8)
    start, end = 6, 80
    initial = [0x1200000000000000 | 0xFFFF]
    layers = O.generate(cfg, initial, start, end)
    vals = O.solve(cfg, layers, start, end, 0, [0x6000000000000000])
    tabs = {}
    for i, (k, v) in enumerate(zip(layers, vals)):
        s = start + 2 * i
        if v is None:
            continue
        k = np.ascontiguousarray(k, np.uint64)
        v = np.ascontiguousarray(v, np.float64)
        with open(d / "t" / ("%d.tables" % s), "wb") as f:
            f.write(k.tobytes())
            f.write(v.tobytes())
        tabs[s] = (k, v)
    with open(d / "c.conf", "w") as f:
        f.write("[Cablegen]\nignore_f = true\n[Solve]\ndir = %s/\n" % (d / "t"))
    return d, cfg, tabs


def test_lookup_command_best_moves(tables):
    d, cfg, tabs = tables
    rng = np.random.default_rng(3)
    checked = nonzero = 0
    for s in (62, 68, 74):
        K, _ = tabs[s]
        K2, V2 = tabs[s + 2]
        K4, V4 = tabs[s + 4]
        for b in K[rng.integers(0, len(K), size=6)]:
            b = int(b)
            exp = []
            for i in range(16):
                if tile(b, i):
                    continue
                for v, (Kt, Vt) in ((1, (K2, V2)), (2, (K4, V4))):
                    md, mp = best(cfg, set_tile(b, i, v), Kt, Vt)
                    exp.append("Best move: %s (%.10f)" % (NAMES[md], mp))
                    nonzero += mp > 0
            r = subprocess.run([CLI, "-C=%s" % (d / "c.conf"), "lookup", "%016x" % b], cwd=str(d),
                               capture_output=True, text=True, timeout=60)
            assert r.returncode == 0, r.stdout + r.stderr
            got = [l.strip() for l in r.stdout.splitlines() if l.startswith("Best move:")]
            assert got == exp, hex(b)
            checked += 1
    assert checked == 18 and nonzero > 20


def test_lookup_spawn_command(tables):
    """`lookup_spawn BOARD` (main.c:191-246 with spawn = true): the first cell
    holding at most 2 (an empty cell counts) is cleared, and only the spawn
    that rebuilds the given board prints its best move."""
    d, cfg, tabs = tables
    rng = np.random.default_rng(4)
    checked = 0
    for s in (62, 68, 74):
        K, _ = tabs[s]
        K2, V2 = tabs[s + 2]
        K4, V4 = tabs[s + 4]
        for b in K[rng.integers(0, len(K), size=4)]:
            # a board of sum s + 2 or s + 4: a tile 2 or 4 placed in an empty cell of b
            b = int(b)
            empty = [i for i in range(16) if not tile(b, i)]
            i0 = empty[int(rng.integers(0, len(empty)))]
            v0 = 1 + int(rng.integers(0, 2))
            q = set_tile(b, i0, v0)
            # the command clears the first cell with a value <= 2, then tries every spawn
            base = q
            for i in range(16):
                if tile(q, i) <= 2:
                    base = set_tile(q, i, 0)
                    break
            s0 = sum(1 << tile(base, i) for i in range(16) if 0 < tile(base, i) < 0xE)
            if s0 != s:
                continue   # the cleared cell was not the spawn: other tables
            exp = []
            for i in range(16):
                if tile(base, i):
                    continue
                for v, (Kt, Vt) in ((1, (K2, V2)), (2, (K4, V4))):
                    if set_tile(base, i, v) == q:
                        md, mp = best(cfg, q, Kt, Vt)
                        exp.append("Best move: %s (%.10f)" % (NAMES[md], mp))
            r = subprocess.run([CLI, "-C=%s" % (d / "c.conf"), "lookup_spawn", "%016x" % q],
                               cwd=str(d), capture_output=True, text=True, timeout=60)
            assert r.returncode == 0, r.stdout + r.stderr
            got = [l.strip() for l in r.stdout.splitlines() if l.startswith("Best move:")]
            assert got == exp, hex(q)
            checked += 1
    assert checked >= 4
