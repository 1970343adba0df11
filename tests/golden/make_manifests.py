"""Writes the layer manifests under tests/golden/.

Each manifest lists, per tile-sum layer s: the board count, the word-wise
FNV-1a-64 hash of the sorted u64 keys and of the f64 value bits, exactly as the %d.boards /
%d.tables files hold them.

  c1_ref.json      config 1 (2x4, initial 0x11000000FFFFFFFF, winstate 128 at
                   tile 0, end 260), ignore_f false and true: produced by the
                   reference library built from its own sources
                   (oracle/_ref/libcablegen_ref.so, prune=false so its CPU path
                   is well defined).  Needs /root/reference's build.
  c2_oracle.json   config 2 (DPDGap 0x00008765FFF3FFF2, prune stsl=58 ltc=5
                   smallest_large=6, ignore_f, end 1100, winstate 512 at
                   (1,0)): produced by the oracle (oracle/liborc.so), keys and
                   values.

  c5_oracle.json   config 5: the gradient dP/dp4 at p4 = 0.1 on the C2
                   table and the per-layer sensitivities D_s, by the oracle's
                   adjoint (FP64).

    python tests/golden/make_manifests.py c1
    python tests/golden/make_manifests.py c2 [cores]
    python tests/golden/make_manifests.py c5 [cores]
"""
import json
import os
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import numpy as np  # noqa: E402

from tests import oracle_lib as O  # noqa: E402

C1 = dict(initial=[0x11000000FFFFFFFF], win=[0x7000000000000000], end=260)
C2 = dict(initial=[0x00008765FFF3FFF2], win=[0x0000900000000000], end=1100,
          rules=dict(ignore_f=True, prune=True, stsl=58, ltc=5, smallest_large=6))


def dump(obj):
    """JSON with one layer row per line."""
    txt = json.dumps(obj, indent=1)
    out, depth_row = [], False
    for line in txt.splitlines():
        st = line.strip()
        if depth_row:
            out[-1] += st
            if st.startswith("]"):
                depth_row = False
            continue
        out.append(line)
        if st == "[" and out and len(out) > 1 and out[-2].strip().endswith(("[", "],")):
            depth_row = True
    return "\n".join(out) + "\n"


def make_c1():
    out = {"config": "2x4 initial 0x11000000FFFFFFFF winstate 0x7000000000000000 end 260",
           "source": "reference library built from its sources", "variants": {}}
    for ig in (False, True):
        d = tempfile.mkdtemp(prefix="cgref_")
        _, start = O.ref_generate_solve(C1["initial"], C1["win"], C1["end"], workdir=d,
                                        cores=2, ignore_f=ig)
        layers = []
        for s in range(start, C1["end"] + 1, 2):
            b = O.read_boards_file(f"{d}/b/{s}.boards")
            k, v = O.read_table_file(f"{d}/t/{s}.tables")
            assert np.array_equal(b, k)
            layers.append([s, int(len(k)), "%016x" % O.hash_u64(k),
                           "%016x" % O.hash_u64(v.view(np.uint64))])
        out["variants"]["ignore_f" if ig else "default"] = layers
    with open(os.path.join(HERE, "c1_ref.json"), "w") as f:
        f.write(dump(out))


def make_c2(cores=8):
    cfg = O.make_cfg(cores=cores, **C2["rules"])
    start = O.orc().orc_get_sum(C2["initial"][0])
    t0 = time.time()
    layers = O.generate(cfg, C2["initial"], start, C2["end"])
    t1 = time.time()
    print("generated", sum(len(x) for x in layers), "boards in %.1fs" % (t1 - t0), flush=True)
    vals = O.solve(cfg, layers, start, C2["end"], 0, C2["win"])
    t2 = time.time()
    print("solved in %.1fs" % (t2 - t1), flush=True)
    rows = []
    for i, k in enumerate(layers):
        rows.append([start + 2 * i, int(len(k)), "%016x" % O.hash_u64(k),
                     "%016x" % O.hash_u64(vals[i].view(np.uint64))])
    out = {"config": "DPDGap initial 0x00008765FFF3FFF2 winstate 0x0000900000000000 ignore_f "
                     "prune stsl=58 ltc=5 smallest_large=6 end 1100",
           "source": "oracle", "cores": cores, "gen_seconds": t1 - t0, "solve_seconds": t2 - t1,
           "p_initial_layer_max": float(max((v.max() for v in vals if len(v)), default=0.0)),
           "layers": rows}
    with open(os.path.join(HERE, "c2_oracle.json"), "w") as f:
        f.write(dump(out))


def make_c5(cores=8):
    """C5: the exact gradient dP/dp4 of the win probability at p4 = 0.1 on the
    C2 table, with the per-layer sensitivities D_s, by the oracle's adjoint
    (orc_adjoint: occupation pushed down the optimal-play successors,
    following the solve of src/solve.c:246-281); P is the initial board's
    value."""
    cfg = O.make_cfg(cores=cores, **C2["rules"])
    start = O.orc().orc_get_sum(C2["initial"][0])
    t0 = time.time()
    layers = O.generate(cfg, C2["initial"], start, C2["end"])
    vals = O.solve(cfg, layers, start, C2["end"], 0, C2["win"])
    t1 = time.time()
    print("generated and solved in %.1fs" % (t1 - t0), flush=True)
    P, dP, D = O.adjoint(cfg, layers, start, vals, C2["win"], C2["initial"], p4=0.1)
    t2 = time.time()
    print("adjoint in %.1fs: P %.17g dP/dp4 %.17g" % (t2 - t1, P, dP), flush=True)
    out = {"config": "C2 table (DPDGap initial 0x00008765FFF3FFF2 winstate 0x0000900000000000 "
                     "ignore_f prune stsl=58 ltc=5 smallest_large=6 end 1100), p4 = 0.1",
           "source": "oracle (orc_adjoint, FP64, one thread)", "cores": cores,
           "gen_solve_seconds": t1 - t0, "adjoint_seconds": t2 - t1,
           "P": P, "dP_dp4": dP,
           "D": [[start + 2 * i, float(d)] for i, d in enumerate(D) if d != 0.0]}
    with open(os.path.join(HERE, "c5_oracle.json"), "w") as f:
        f.write(dump(out))


if __name__ == "__main__":
    what = sys.argv[1] if len(sys.argv) > 1 else "c1"
    if what == "c1":
        make_c1()
    elif what == "c5":
        make_c5(int(sys.argv[2]) if len(sys.argv) > 2 else 8)
    else:
        make_c2(int(sys.argv[2]) if len(sys.argv) > 2 else 8)
