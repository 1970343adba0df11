"""The drop-in C ABI driven from C: tests/native/dropin_caller.c (our own
caller, linked to libcablegen.so, built by build()) runs generate() and
solve() the way the reference's CLI does (frontends/cli/main.c:79-139), then
consumes the GPU-written tables as the CLI's lookup / lookup_spawn / explore
do (main.c:174-287).  Expected layers come from the reference-made manifest
(tests/golden/c1_ref.json) and expected consumer lines from the oracle's
tables through a restatement of best() (tests/oracle_lib.py).

The reference's own CLI (oracle/_ref, built from its sources in the build
container only) is exercised on the CPU: it links to the library and its
host-only commands (write, read, explore) run on oracle-written files."""
import json
import os
import subprocess

import numpy as np
import pytest

from tests import oracle_lib as O

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CALLER = os.path.join(HERE, "native", "_build", "dropin_caller")
REF_CLI = os.path.join(ROOT, "oracle", "_ref", "cablegen_cli_mi355x")

needs_caller = pytest.mark.skipif(not os.path.exists(CALLER),
                                  reason="tests/native/dropin_caller not built (run build())")


def write_conf(d, initial, win, end_gen, end_solve=0, ignore_f=True, delete_boards=False, premove=False):
    os.makedirs(f"{d}/b", exist_ok=True)
    os.makedirs(f"{d}/t", exist_ok=True)
    np.array(initial, np.uint64).tofile(f"{d}/initial")
    np.array(win, np.uint64).tofile(f"{d}/winstates")
    conf = f"{d}/c.conf"
    with open(conf, "w") as f:
        f.write(f"[Cablegen]\nignore_f = {'true' if ignore_f else 'false'}\ncores = 2\n"
                f"[Generate]\ndir = {d}/b/\ninitial = {d}/initial\nend = {end_gen}\n"
                f"premove = {'true' if premove else 'false'}\n"
                f"[Solve]\ndir = {d}/t/\nwinstates = {d}/winstates\nend = {end_solve}\n"
                f"delete_boards = {'true' if delete_boards else 'false'}\n")
    return conf


def run(*args, timeout=180):
    print("dropin_caller", *args[1:], flush=True)
    r = subprocess.run(list(args), capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def parse_explore(out):
    vals, boards, rows = [], [], []
    for line in out.splitlines():
        if line.startswith("Board("):
            vals.append(float(line[6:line.index(")")]))
            rows = []
        elif line.strip() or rows:
            rows.append(line)
            if len(rows) == 4:
                boards.append(rows)
    return vals, boards


# ---------------------------------------------------------------- GPU: generate/solve + consumers

@pytest.mark.gpu
@needs_caller
def test_native_caller_c1_manifest_and_consumers(tmp_path):
    """C1 (2x4, ignore_f) through generate()/solve() from C: every table
    matches the reference-made manifest; lookup, lookup_spawn and explore
    on the GPU-written tables give the oracle's best moves and values."""
    d = str(tmp_path)
    conf = write_conf(d, [0x11000000FFFFFFFF], [0x7000000000000000], 260)
    run(CALLER, "gensolve", conf)
    rows = json.load(open(os.path.join(HERE, "golden", "c1_ref.json")))["variants"]["ignore_f"]
    for s, n, hk, hv in rows:
        k, v = O.read_table_file(f"{d}/t/{s}.tables")
        assert len(k) == n and "%016x" % O.hash_u64(k) == hk, s
        assert "%016x" % O.hash_u64(v.view(np.uint64)) == hv, s
        assert np.array_equal(O.read_boards_file(f"{d}/b/{s}.boards"), k), s
    _check_consumers(d, conf, O.make_cfg(ignore_f=True, cores=8), [0x11000000FFFFFFFF],
                     [0x7000000000000000], 4, 260, (100, 140, 180))


@pytest.mark.gpu
@needs_caller
def test_native_caller_3x4_consumers(tmp_path):
    """A 3x4 table (bottom row 0xf) written by the GPU through the C ABI, read
    back by the consumers."""
    d = str(tmp_path)
    init, win = [0x1200000000000000 | 0xFFFF], [0x6000000000000000]
    conf = write_conf(d, init, win, 80)
    run(CALLER, "gensolve", conf)
    _check_consumers(d, conf, O.make_cfg(ignore_f=True, cores=8), init, win, 6, 80, (62, 68, 74))


def _check_consumers(d, conf, cfg, initial, win, start, end, sums):
    layers = O.generate(cfg, initial, start, end)
    vals = O.solve(cfg, layers, start, end, 0, win)
    tabs = {start + 2 * i: (np.asarray(k, np.uint64), np.asarray(v, np.float64))
            for i, (k, v) in enumerate(zip(layers, vals)) if v is not None}
    for s, (K, V) in tabs.items():   # the GPU's tables are the oracle's
        k, v = O.read_table_file(f"{d}/t/{s}.tables")
        assert np.array_equal(k, K) and np.array_equal(v.view(np.uint64), V.view(np.uint64)), s
    rng = np.random.default_rng(5)
    nonzero = 0
    for s in sums:
        K, V = tabs[s]
        for b in K[rng.integers(0, len(K), size=5)]:
            exp = O.lookup_lines(cfg, int(b), tabs)
            got = [l.strip() for l in run(CALLER, "lookup", conf, "%016x" % int(b), "0").splitlines()
                   if l.startswith("Best move:")]
            assert got == exp, hex(int(b))
            nonzero += sum(not l.endswith("(0.0000000000)") for l in exp)
            # the same board with a 2 spawned in its first empty cell
            i0 = next(i for i in range(16) if O.tile(int(b), i) == 0)
            q = O.set_tile(int(b), i0, 1)
            exp = O.lookup_lines(cfg, q, tabs, spawn=True)
            got = [l.strip() for l in run(CALLER, "lookup", conf, "%016x" % q, "1").splitlines()
                   if l.startswith("Best move:")]
            assert got == exp, hex(q)
        # explore: every board of the table with its value
        out = run(CALLER, "explore", f"{d}/t/{s}.tables")
        got_v, got_b = parse_explore(out)
        assert len(got_v) == len(K) and len(got_b) == len(K)
        assert np.allclose(got_v, V, rtol=0, atol=5e-20)
    assert nonzero > 10


@pytest.mark.gpu
@needs_caller
def test_native_caller_delete_boards(tmp_path):
    """[Solve] delete_boards = true: solve() removes each %d.boards file once
    it has read it (src/solve.c:206-207); the tables are unchanged."""
    d = str(tmp_path)
    conf = write_conf(d, [0x11000000FFFFFFFF], [0x7000000000000000], 120, delete_boards=True)
    run(CALLER, "gensolve", conf)
    assert sorted(os.listdir(f"{d}/b")) == []
    rows = json.load(open(os.path.join(HERE, "golden", "c1_ref.json")))["variants"]["ignore_f"]
    checked = 0
    for s, n, hk, hv in rows:
        if s > 120:
            continue
        k, v = O.read_table_file(f"{d}/t/{s}.tables")
        assert len(k) == n and "%016x" % O.hash_u64(k) == hk, s
        checked += 1
    assert checked == 59
    # without the key the files stay
    d2 = str(tmp_path / "keep")
    os.makedirs(d2)
    conf2 = write_conf(d2, [0x11000000FFFFFFFF], [0x7000000000000000], 120)
    run(CALLER, "gensolve", conf2)
    assert len(os.listdir(f"{d2}/b")) == 59


@pytest.mark.gpu
@needs_caller
def test_native_caller_premove(tmp_path):
    """[Generate] premove = true through generate()/solve() from C: the
    initial boards are positions after the move, so the first layer is their
    canonical set; every table equals the oracle's premove run."""
    cfg = O.make_cfg(ignore_f=True, cores=8)
    plain = O.generate(cfg, [0x11000000FFFFFFFF], 4, 120)
    b0 = [O.orc().orc_rotate_cw(int(x)) for x in plain[0]]
    cfgp = O.make_cfg(ignore_f=True, cores=8, premove=True)
    layers = O.generate(cfgp, b0, 4, 120)
    vals = O.solve(cfgp, layers, 4, 120, 0, [0x7000000000000000])
    d = str(tmp_path)
    conf = write_conf(d, b0, [0x7000000000000000], 120, premove=True)
    run(CALLER, "gensolve", conf)
    for i, (K, V) in enumerate(zip(layers, vals)):
        k, v = O.read_table_file(f"{d}/t/{4 + 2 * i}.tables")
        assert np.array_equal(k, K) and np.array_equal(v.view(np.uint64), V.view(np.uint64)), i
        assert np.array_equal(O.read_boards_file(f"{d}/b/{4 + 2 * i}.boards"), K), i
    assert np.array_equal(layers[0], plain[0])


# ---------------------------------------------------------------- CPU: host-only consumers

@pytest.fixture(scope="module")
def oracle_tables(tmp_path_factory):
    d = tmp_path_factory.mktemp("host")
    (d / "t").mkdir()
    cfg = O.make_cfg(ignore_f=True, cores=8)
    layers = O.generate(cfg, [0x1200000000000000 | 0xFFFF], 6, 40)
    vals = O.solve(cfg, layers, 6, 40, 0, [0x6000000000000000])
    s = 30
    k, v = np.asarray(layers[(s - 6) // 2], np.uint64), np.asarray(vals[(s - 6) // 2], np.float64)
    path = str(d / "t" / f"{s}.tables")
    with open(path, "wb") as f:
        f.write(k.tobytes())
        f.write(v.tobytes())
    return d, path, k, v


@needs_caller
def test_native_caller_explore_host(oracle_tables):
    _, path, k, v = oracle_tables
    got_v, got_b = parse_explore(run(CALLER, "explore", path))
    assert len(got_b) == len(k) and np.allclose(got_v, v, rtol=0, atol=5e-20)


@pytest.mark.skipif(not os.path.exists(REF_CLI), reason="reference CLI not built (build container only)")
def test_reference_cli_host_commands(oracle_tables):
    """The reference's unchanged CLI, compiled against include/cablegen.h and
    linked to libcablegen.so, runs its host-only commands on this library:
    write + read round trip of boards (main.c:140-152, 289-299) and explore
    of a table (main.c:275-287), whose output equals our caller's."""
    d, path, k, v = oracle_tables
    boards = ["%016x" % int(x) for x in k[:5]]
    r = subprocess.run([REF_CLI, "write", str(d / "w.boards")] + boards, capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert np.array_equal(O.read_boards_file(str(d / "w.boards")), k[:5])
    r = subprocess.run([REF_CLI, "read", str(d / "w.boards")], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 0 and r.stdout.count("----") == 4, r.stdout + r.stderr
    r = subprocess.run([REF_CLI, "explore", path], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    got_v, got_b = parse_explore(r.stdout)
    assert len(got_b) == len(k) and np.allclose(got_v, v, rtol=0, atol=5e-20)
    if os.path.exists(CALLER):
        assert r.stdout == run(CALLER, "explore", path)
