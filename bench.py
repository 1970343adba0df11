"""Benchmark: boards/s of layered 2048 tablebase generate + solve on one MI355X.

Workload (one step): the DPDGap-512 tablebase -- initial board
0x00008765FFF3FFF2 (tile sum 492), formation cells as walls (ignore_f),
prune stsl=58 ltc=5 smallest_large=6, generation to tile sum 1100, solve
from 1100 down to 0 with the winstate 0x0000900000000000 (512 next to the
formation, all 8 images).  Every layer stays resident in HBM; the input (one
board) is on the device before the timed region starts.

value = total boards over all layers / (T_generate + T_solve), per step.

Also reported:
  roofline      the stage with the largest total time (move, spawn, solve, or
                the sort stage = every kernel of a sort + dedupe), from one
                untimed step with every kernel timed, then timed in the timed
                steps with HIP events on the stream each launch runs on.
                achieved = algorithmic bytes / time over its main-stream units
                (with --carry-sort 1 the carry sorts, on the second stream,
                share the GPU with the next step's moves and are reported
                beside it).  Algorithmic
                bytes: the payload read once and written once -- a sort
                8 B per key + 8 B per unique key (extra radix passes are
                implementation overhead), the move kernel 8 B per board + 8 B
                per candidate, the spawn 8 B per board + 16 B per spawn, the
                solve 16 B per board + 24 B per lookup + 8 B per hit.
                traffic = calibrated rocprofv3 HBM bytes per unit
  cpu_baseline  the oracle (a pthreads C port of the reference's CPU
                algorithm) on min(nproc, OMP_NUM_THREADS, 127) threads, on
                windows of three consecutive layers spread over the table
                (about --cpu-seconds of CPU work; --cpu-full: the whole
                table), outputs checked bit-for-bit against the GPU's
  adjoint       time of the exact dP(win)/dp4 sweep relative to a plain
                solve, mean of --adjoint-reps sweeps
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

METRIC = "boards/sec gen+solve, DPDGap-512, 1 GPU vs host CPU; adjoint/solve time ratio"
INITIAL = [0x00008765FFF3FFF2]
WINSTATES = [0x0000900000000000]
END_GEN = 1100
RULES = dict(ignore_f=True, prune=True, stsl=58, ltc=5, smallest_large=6)
HBM_PEAK_GBS = 8000.0
KERNELS = ["move", "spawn", "sort_hist", "sort_pass", "unique", "sort_split", "sort_part",
           "sort_local", "solve"]
# the sort stage's kernels (LSD radix: histogram, passes, compaction; sample
# sort: splitters, partition, bucket sort + dedupe); the last one of each sort
# carries the stage's algorithmic bytes 8 B per key in + 8 B per unique key out
SORT_KINDS = ["sort_hist", "sort_pass", "unique", "sort_split", "sort_part", "sort_local"]
SORT_FINAL = ["unique", "sort_local"]
# roofline candidates: the move, spawn and solve kernels, and the sort stage
STAGES = ["move", "spawn", "sort", "solve"]


def host_cpu():
    """CPU model, the cores this process may run on (what nproc prints), the
    machine's logical CPUs, and the thread budget of the job (OMP_NUM_THREADS
    when the environment sets one)"""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    nproc = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    share = os.environ.get("OMP_NUM_THREADS")
    share = int(share) if share and share.isdigit() and int(share) > 0 else None
    return model, nproc, os.cpu_count() or nproc, share


def baseline_cores():
    """Every core the job may use, capped at 127: the reference merges
    cores + 1 sorted arrays and allows at most 128 (src/array.c:247-251)"""
    _, nproc, _, share = host_cpu()
    return max(1, min(nproc, share or nproc, 127))


def cpu_baseline(tb, cg, cores, budget_s=20.0, full=False):
    """The oracle (the reference's CPU algorithm, -O2, pthreads on `cores`
    threads) timed on windows of consecutive full-size layers spread over
    the table, until about budget_s of CPU work; each window's pre-move
    layer and carry are rebuilt from the GPU's two previous layers, and every
    output is compared bit-for-bit with the GPU's.  full=True times the
    whole table instead (generate() then solve(), as the reference runs it)."""
    from tests import oracle_lib as O
    cfg = O.make_cfg(cores=cores, **RULES)
    model, nproc, ncpu, share = host_cpu()
    total = tb.total_boards()
    W = O.expand_winstates(WINSTATES)
    host = {"cpu_model": model, "nproc": nproc, "cpu_count": ncpu, "omp_num_threads": share,
            "cores_rule": "min(nproc, OMP_NUM_THREADS, 127)"}
    if full:
        start = tb.start
        t0 = time.perf_counter()
        layers = O.generate(cfg, INITIAL, start, END_GEN)
        t1 = time.perf_counter()
        vals = O.solve(cfg, layers, start, END_GEN, 0, WINSTATES)
        t2 = time.perf_counter()
        exact = len(layers) == tb.num_layers and all(
            np.array_equal(layers[i], tb.keys(i)) and
            np.array_equal(vals[i].view(np.uint64), tb.values(i).view(np.uint64))
            for i in range(len(layers)))
        boards = sum(len(x) for x in layers)
        return dict(host, value=boards / (t2 - t0), unit="boards/s", cores=cores, kind="port",
                    sample=f"the whole table: generate {t1 - t0:.1f}s + solve {t2 - t1:.1f}s of "
                           f"{boards} boards; outputs bit-identical to the GPU's: {exact}",
                    sample_fraction=boards / total)
    sizes = [tb.layer_size(i) for i in range(tb.num_layers)]
    nl = tb.num_layers
    # window starts spread over the table's bulk (layers of at least a tenth of the largest)
    big = [i for i in range(4, nl - 3) if sizes[i] >= max(sizes) / 10]
    order = []
    if big:
        lo, hi = big[0], big[-1] - 2
        for k in range(8):
            i0 = lo + (hi - lo) * ((5 * k) % 8) // 7   # a spread order: middle first
            if i0 not in order:
                order.append(i0)
    t_gen = t_solve = 0.0
    boards = 0
    exact = True
    done = []
    for i0 in order:
        if t_gen + t_solve >= budget_s:
            break
        # pre-move layer P_s = S2(M_{s-2}) u S4(M_{s-4}) and the carry S4(M_{s-2})
        s2_a, s4_a = O.spawn_sets(cfg, tb.keys(i0 - 1))
        _, s4_b = O.spawn_sets(cfg, tb.keys(i0 - 2))
        P = np.union1d(s2_a, s4_b).astype(np.uint64)
        Cc = s4_a
        for i in (i0, i0 + 1, i0 + 2):
            t0 = time.perf_counter()
            M, P, Cc = O.gen_layer(cfg, P, Cc)
            t_gen += time.perf_counter() - t0
            exact &= bool(np.array_equal(M, tb.keys(i)))
            K2, V2 = tb.keys(i + 1), tb.values(i + 1)
            K4, V4 = tb.keys(i + 2), tb.values(i + 2)
            t0 = time.perf_counter()
            V = O.solve_layer(cfg, M, K2, V2, K4, V4, W)
            t_solve += time.perf_counter() - t0
            exact &= bool(np.array_equal(V.view(np.uint64), tb.values(i).view(np.uint64)))
            boards += len(M)
        done.append(tb.start + 2 * i0)
    return dict(host, value=boards / (t_gen + t_solve), unit="boards/s", cores=cores, kind="port",
                sample=f"generation step + solve of 3 consecutive layers from each of s={done} "
                       f"({boards} of the table's {total} boards, {100.0 * boards / total:.1f}%; "
                       f"gen {t_gen:.2f}s, solve {t_solve:.2f}s); outputs bit-identical to the "
                       f"GPU's: {exact}",
                sample_fraction=boards / total)


def micro(args):
    """Kernel roofline microbench (seed 1): 2^30 random boards under the
    DPDGap mask -> 4 moves + canonical form -> sort + dedupe; then 2^28
    queries (80% hits) into a sorted 2^30-entry layer.  HBM GB/s per kernel
    from the library's HIP-event timers, algorithmic bytes as documented."""
    import ctypes as C
    import torch
    import cablegen_amd as cg
    L = cg.lib()
    log2n = args.micro_log2
    n = 1 << log2n
    nq = n >> 2
    fmask = 0x00000000FFF0FFF0
    r = cg.rules(ignore_f=True)
    dev = torch.device("cuda")
    out = {"metric": "kernel roofline microbench (HBM GB/s per kernel)", "n_boards": n,
           "queries": nq, "seed": 1, "formation_mask": "0x00000000FFF0FFF0",
           "rules": "ignore_f, no prune", "kernels": {}}
    out["tuning"] = run_autotune(cg, args)

    def timed(name, fn):
        cg.prof_enable(True)
        cg.prof_reset()
        fn()
        res = {}
        for k in ("move", "sort_hist", "sort_pass", "unique", "lookup"):
            cnt, ms, b = cg.prof_query(k)
            if cnt:
                res[k] = {"launches": cnt, "ms": ms, "GB_s": b / ms / 1e6 if ms else 0.0,
                          "frac_of_8TBs": (b / ms / 1e6) / HBM_PEAK_GBS if ms else 0.0}
        cg.prof_enable(False)
        out["kernels"][name] = res

    # calibration for HBM counters: a plain device copy of 2^27 u64 (1 GiB read, 1 GiB written)
    cx = torch.ones(1 << 27, dtype=torch.int64, device=dev)
    cy = torch.empty_like(cx)
    cy.copy_(cx)
    torch.cuda.synchronize()
    out["calibration_copy_bytes_each_way"] = 8 << 27
    del cx, cy
    boards = torch.empty(n, dtype=torch.int64, device=dev)
    assert L.cg_dev_random_boards(boards.data_ptr(), n, 1, fmask, 0.4, 9, 0) == 0
    cand = torch.empty(4 * n, dtype=torch.int64, device=dev)
    ncand = C.c_size_t()
    for _ in range(2):   # warm-up + measured
        timed("move", lambda: L.cg_dev_move(C.byref(r), boards.data_ptr(), n, cand.data_ptr(),
                                            C.byref(ncand)))
    uniq = torch.empty(ncand.value, dtype=torch.int64, device=dev)
    nu = C.c_size_t()
    t0 = time.perf_counter()
    timed("sort_unique", lambda: L.cg_dev_sort_unique(cand.data_ptr(), ncand.value, uniq.data_ptr(),
                                                      C.byref(nu)))
    out["candidates"] = ncand.value
    out["unique"] = nu.value
    out["sort_unique_wall_ms"] = (time.perf_counter() - t0) * 1e3
    sk = out["kernels"]["sort_unique"]
    tot_ms = sum(v["ms"] for v in sk.values())
    out["sort_unique_stage_GB_s"] = (8.0 * ncand.value + 8.0 * nu.value) / tot_ms / 1e6
    del cand, uniq
    torch.cuda.empty_cache()
    # Lookup layer: the skewed distribution above has far fewer than 2^30
    # distinct boards (2^30 draws give ~4e7), so the layer is drawn with free
    # cells uniform in 0..13 (about 3.6e10 canonical boards), canonicalised,
    # sorted, deduplicated and cut to its first 2^log2n entries.
    extra = n + n // 8
    lay = torch.empty(extra, dtype=torch.int64, device=dev)
    assert L.cg_dev_random_boards(lay.data_ptr(), extra, 2, fmask, 0.0, 13, 1) == 0
    assert L.cg_dev_canon(lay.data_ptr(), extra) == 0
    keys = torch.empty(extra, dtype=torch.int64, device=dev)
    nk = C.c_size_t()
    assert L.cg_dev_sort_unique(lay.data_ptr(), extra, keys.data_ptr(), C.byref(nk)) == 0
    del lay
    nl = min(nk.value, n)
    keys = keys[:nl].contiguous()
    vals = torch.arange(nl, dtype=torch.float64, device=dev) / nl + 1.0 / (4 * nl)
    g = torch.Generator(device=dev)
    g.manual_seed(1)
    nh = (nq * 4) // 5
    q = torch.empty(nq, dtype=torch.int64, device=dev)
    q[:nh] = keys[torch.randint(0, nl, (nh,), device=dev, generator=g)]
    # 20% verified-absent canonical boards: draw, look up, keep the misses
    need = nq - nh
    cand_q = torch.empty(2 * need, dtype=torch.int64, device=dev)
    assert L.cg_dev_random_boards(cand_q.data_ptr(), 2 * need, 3, fmask, 0.0, 13, 1) == 0
    assert L.cg_dev_canon(cand_q.data_ptr(), 2 * need) == 0
    probe = torch.empty(2 * need, dtype=torch.float64, device=dev)
    assert L.cg_dev_lookup(keys.data_ptr(), vals.data_ptr(), nl, cand_q.data_ptr(), 2 * need,
                           probe.data_ptr(), 0) == 0
    absent = cand_q[probe == 0]
    assert absent.numel() >= need
    q[nh:] = absent[:need]
    del cand_q, probe, absent
    q = q[torch.randperm(nq, device=dev, generator=g)].contiguous()
    res = torch.empty(nq, dtype=torch.float64, device=dev)
    for _ in range(2):
        timed("lookup", lambda: L.cg_dev_lookup(keys.data_ptr(), vals.data_ptr(), nl, q.data_ptr(),
                                                nq, res.data_ptr(), 0))
    hits = int((res > 0).sum().item())
    out["layer_entries"] = nl
    out["hit_fraction"] = hits / nq
    # algorithmic bytes of the lookup: 24 B per query (query in, one key probe,
    # result out) + 8 B per hit (value)
    lk = out["kernels"]["lookup"]["lookup"]
    lk["algorithmic_GB_s"] = (24.0 * nq + 8.0 * hits) / lk["ms"] / 1e6
    lk["frac_of_8TBs"] = lk["algorithmic_GB_s"] / HBM_PEAK_GBS
    print(json.dumps(out))


PMC_KERNELS = {"move": "::k_move", "spawn": "::k_spawn(", "sort_hist": "::k_os_hist",
               "sort_pass": "::k_os_pass", "unique": "::k_os_unique",
               "sort_split:gather": "::k_ss_gather(", "sort_split:runs": "::k_ss_runs(",
               "sort_split:select": "::k_ss_select(", "sort_split:rank": "::k_ss_rank(",
               "sort_split:pick": "::k_ss_pick(",
               "sort_part": "::k_ss_part", "sort_local": "::k_ss_local(", "solve": "::k_solve<",
               "lookup": "::k_lookup(", "adjoint": "::k_adjoint("}
# SQ counter sets (one rocprofv3 run each) and the ratios printed from them:
# per-wave cycles issuing VALU / LDS / VMEM, waiting on LDS or on anything,
# LDS bank-conflict cycles per LDS-active cycle, instructions per wave
SQ_SETS = [["SQ_WAVE_CYCLES", "SQ_BUSY_CYCLES", "SQ_ACTIVE_INST_VALU", "SQ_ACTIVE_INST_LDS",
            "SQ_ACTIVE_INST_VMEM", "SQ_WAIT_INST_LDS", "SQ_WAIT_ANY", "SQ_LDS_BANK_CONFLICT"],
           ["SQ_WAVES", "SQ_INSTS_VALU", "SQ_INSTS_LDS", "SQ_INSTS_VMEM_RD", "SQ_INSTS_VMEM_WR",
            "SQ_INSTS_SALU", "SQ_INSTS_SMEM", "SQ_LDS_ADDR_CONFLICT"]]


def _csvs(d, suffix):
    import glob
    return sorted(glob.glob(os.path.join(d, "**", "*" + suffix), recursive=True))


def prof_summary(args):
    """Summaries of rocprofv3 output (printed as markdown):
    --prof-summary DIR            kernel table of DIR's *_kernel_stats.csv;
    --prof-summary DIR --pmc      per-kernel HBM traffic and L2 hit rate from the
                                  FETCH_SIZE / WRITE_SIZE / TCC_HIT_sum+TCC_MISS_sum
                                  runs in DIR/*FETCH_SIZE*, *WRITE_SIZE*, *TCC_HIT*;
                                  the calibration copy (__amd_rocclr_copyBuffer of
                                  1 GiB) is taken from --pmc-calib DIR (default DIR);
                                  --pmc-json FILE also writes per-launch averages;
                                  with runs of the SQ_SETS counters (DIR/*SQ_WAVE_CYCLES*,
                                  *SQ_WAVES*), a table of per-wave issue and wait ratios."""
    import csv
    import collections
    d = args.prof_summary
    if not args.pmc:
        for f in _csvs(d, "_kernel_stats.csv"):
            rows = list(csv.DictReader(open(f)))
            print("| kernel | calls | total ms | avg us | % | min us | max us |")
            print("|---|---:|---:|---:|---:|---:|---:|")
            for r in rows[:args.prof_rows]:
                print("| %s | %s | %.1f | %.2f | %.2f | %.2f | %.2f |" % (
                    r["Name"][:70].replace("|", "/"), r["Calls"], float(r["TotalDurationNs"]) / 1e6,
                    float(r["AverageNs"]) / 1e3, float(r["Percentage"]), float(r["MinNs"]) / 1e3,
                    float(r["MaxNs"]) / 1e3))
        # per stream: the main stream is the one that runs the move kernel;
        # the second one runs the carry sort beside the next step's moves
        for f in _csvs(d, "_kernel_trace.csv"):
            rows = list(csv.DictReader(open(f)))
            main = collections.Counter(r["Stream_Id"] for r in rows if "::k_move" in r["Kernel_Name"])
            if not main:
                continue
            main_id = main.most_common(1)[0][0]
            print("\nBy stream (main = the stream running `k_move`; times from the kernel trace):\n")
            print("| kernel | main: calls | main: avg us | second: calls | second: avg us |")
            print("|---|---:|---:|---:|---:|")
            for name, pat in (("sort_split", "::k_ss_gather"), ("sort_split", "::k_ss_runs"),
                              ("sort_split", "::k_ss_select"), ("sort_part", "::k_ss_part"),
                              ("sort_local", "::k_ss_local"), ("sort_hist", "::k_os_hist"),
                              ("sort_pass", "::k_os_pass"), ("unique", "::k_os_unique"),
                              ("move", "::k_move_gen"), ("spawn", "::k_spawn("), ("solve", "::k_solve<")):
                name = "%s (%s)" % (name, pat.strip(":(")) if name.startswith("sort_split") else name
                acc = {True: [], False: []}
                for r in rows:
                    if pat in r["Kernel_Name"]:
                        acc[r["Stream_Id"] == main_id].append(
                            (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)
                mn, sn = acc[True], acc[False]
                print("| %s | %d | %.2f | %d | %.2f |" % (name, len(mn), sum(mn) / max(len(mn), 1),
                                                          len(sn), sum(sn) / max(len(sn), 1)))
        return

    def load(sub, name):
        out = {}
        for f in _csvs(os.path.join(d), "_counter_collection.csv"):
            if sub not in f or (args.pmc_tag and args.pmc_tag not in f):
                continue
            for r in csv.DictReader(open(f)):
                if r["Counter_Name"] == name:
                    out[int(r["Dispatch_Id"])] = (r["Kernel_Name"], float(r["Counter_Value"]),
                                                  r["Queue_Id"])
        return out

    def main_queue(m):
        """queue of the move kernel's dispatches (the library's main stream)"""
        q = collections.Counter(v[2] for v in m.values() if "::k_move" in v[0])
        return q.most_common(1)[0][0] if q else None
    F, W = load("FETCH_SIZE", "FETCH_SIZE"), load("WRITE_SIZE", "WRITE_SIZE")
    H, M = load("TCC_HIT", "TCC_HIT_sum"), load("TCC_HIT", "TCC_MISS_sum")
    calib_dir = args.pmc_calib or d
    cF = max(v for k, v in load_calib(calib_dir, "FETCH_SIZE"))
    cW = max(v for k, v in load_calib(calib_dir, "WRITE_SIZE"))
    gib_kib = float(1 << 20)
    kf, kw = gib_kib / cF, gib_kib / cW
    print("calibration: a 1 GiB device copy reads FETCH_SIZE %.0f KiB and WRITE_SIZE %.0f KiB; "
          "factors %.3f (fetch) and %.3f (write)\n" % (cF, cW, kf, kw))
    qF, qW, qH = main_queue(F), main_queue(W), main_queue(H)
    print("| kernel | launches | HBM read / launch (MB) | HBM write / launch (MB) | L2 hit rate "
          "| main stream: launches | read (MB) | write (MB) |")
    print("|---|---:|---:|---:|---:|---:|---:|---:|")
    js = {"calibration": {"fetch_factor": kf, "write_factor": kw}, "kernels": {},
          "kernels_main_stream": {}}

    def per_launch(T, pat, k, q=None):
        ids = [i for i, v in T.items() if pat in v[0] and (q is None or v[2] == q)]
        return len(ids), (sum(T[i][1] for i in ids) * 1024 * k / len(ids) if ids else 0.0)
    for name, pat in PMC_KERNELS.items():
        nl, fb = per_launch(F, pat, kf)
        if not nl:
            continue
        _, wb = per_launch(W, pat, kw)
        hl = [i for i, v in H.items() if pat in v[0]]
        h = sum(H[i][1] for i in hl)
        m = sum(M[i][1] for i in hl if i in M)
        hit = h / (h + m) if h + m else float("nan")
        nm, fm = per_launch(F, pat, kf, qF)
        _, wm = per_launch(W, pat, kw, qW)
        print("| %s | %d | %.3f | %.3f | %.3f | %d | %.3f | %.3f |" % (
            name, nl, fb / 1e6, wb / 1e6, hit, nm, fm / 1e6, wm / 1e6))
        js["kernels"][name] = {"launches": nl, "read_bytes_per_launch": fb,
                               "write_bytes_per_launch": wb, "l2_hit_rate": hit}
        js["kernels_main_stream"][name] = {"launches": nm, "read_bytes_per_launch": fm,
                                           "write_bytes_per_launch": wm}
    if args.pmc_json:
        js["workload"] = args.pmc_workload
        with open(args.pmc_json, "w") as f:
            json.dump(js, f, indent=1)
    sq = {}
    for grp in SQ_SETS:
        for c in grp:
            sq[c] = load(grp[0], c)
    if not sq["SQ_WAVE_CYCLES"] or not sq["SQ_WAVES"]:
        return
    print("\nPer-wave issue and wait (SQ counters summed over a kernel's dispatches; fractions of "
          "SQ_WAVE_CYCLES; bank conflicts per SQ_ACTIVE_INST_LDS cycle; instructions per wave):\n")
    print("| kernel | VALU | LDS | VMEM | wait LDS | wait any | LDS bank confl. | VALU / wave "
          "| LDS / wave | VMEM rd / wave | VMEM wr / wave | SALU / wave |")
    print("|---|---:|---:|---:|---:|---:|---:|---:|---:|---:|---:|---:|")

    def tot(c, pat):
        return sum(v[1] for v in sq[c].values() if pat in v[0])
    for name, pat in PMC_KERNELS.items():
        wc, nw = tot("SQ_WAVE_CYCLES", pat), tot("SQ_WAVES", pat)
        if not wc or not nw:
            continue
        lds = tot("SQ_ACTIVE_INST_LDS", pat)
        print("| %s | %.3f | %.3f | %.3f | %.3f | %.3f | %.3f | %.0f | %.0f | %.1f | %.1f | %.0f |" % (
            name, tot("SQ_ACTIVE_INST_VALU", pat) / wc, lds / wc, tot("SQ_ACTIVE_INST_VMEM", pat) / wc,
            tot("SQ_WAIT_INST_LDS", pat) / wc, tot("SQ_WAIT_ANY", pat) / wc,
            tot("SQ_LDS_BANK_CONFLICT", pat) / lds if lds else 0.0, tot("SQ_INSTS_VALU", pat) / nw,
            tot("SQ_INSTS_LDS", pat) / nw, tot("SQ_INSTS_VMEM_RD", pat) / nw,
            tot("SQ_INSTS_VMEM_WR", pat) / nw, tot("SQ_INSTS_SALU", pat) / nw))


def load_calib(d, name):
    import csv
    out = []
    for f in _csvs(d, "_counter_collection.csv"):
        if name not in f:
            continue
        for r in csv.DictReader(open(f)):
            if r["Counter_Name"] == name and "copyBuffer" in r["Kernel_Name"]:
                out.append((r["Kernel_Name"], float(r["Counter_Value"])))
    return out


def run_autotune(cg, args):
    """Sort / directory tuning before the warmup (never timed).  "stored"
    installs the table the autotuner wrote on an MI355X
    (cablegen_amd/tuned/mi355x.json; the tuner runs if it is missing), "auto"
    runs the autotuner now, "off" keeps the defaults."""
    mode = args.tuning
    t0 = time.perf_counter()
    if mode == "off":
        cg.tune_set("builtin", 0)
        cg.tune_set("cell_code", args.cell_code)
        cg.tune_set("carry_stream", args.carry_stream)
        cg.tune_set("carry_sort", args.carry_sort)
        cg.tune_set("sample_sort", args.sample_sort)
        cg.tune_set("canon_stab", args.canon_stab)
        cg.tune_set("one_sync", args.one_sync)
        if args.sample_bucket:
            cg.tune_set("sample_bucket", args.sample_bucket)
        if args.sample_bmax:
            cg.tune_set("sample_bmax", args.sample_bmax)
        knobs = set_knobs(cg, args)
        return {"source": "built-in defaults: 8-bit digits, 8192-key tiles, window 16, dir delta 0",
                "knobs": knobs,
                "cell_code": args.cell_code, "carry_stream": args.carry_stream,
                "carry_sort": args.carry_sort,
                "sample_sort": args.sample_sort, "sample_bucket": cg.tune_get("sample_bucket"),
                "sample_bmax": cg.tune_get("sample_bmax"),
                "canon_stab": args.canon_stab, "one_sync": args.one_sync}
    if mode == "stored" and os.path.exists(cg.TUNED_TABLE):
        rep = cg.load_tuning()
        src = "stored autotuner table cablegen_amd/tuned/mi355x.json"
    else:
        rep = cg.autotune(1)
        src = "autotuner run now (effort 1)"
    if args.save_tuning:
        with open(args.save_tuning, "w") as f:
            json.dump(rep, f)
    cg.tune_set("cell_code", args.cell_code)
    cg.tune_set("carry_stream", args.carry_stream)
    cg.tune_set("carry_sort", args.carry_sort)
    cg.tune_set("sample_sort", args.sample_sort)
    cg.tune_set("canon_stab", args.canon_stab)
    cg.tune_set("one_sync", args.one_sync)
    if args.sample_bucket:
        cg.tune_set("sample_bucket", args.sample_bucket)
    if args.sample_bmax:
        cg.tune_set("sample_bmax", args.sample_bmax)
    knobs = set_knobs(cg, args)
    dir_delta = rep["dir"]["pick"]
    if args.dir_delta is not None:
        dir_delta = args.dir_delta
        cg.tune_set("dir_delta", dir_delta)
    sel = {}
    for cls, name in enumerate(("n<2^20", "n<2^23", "n>=2^23")):
        sel[name] = {nb: rep["sort"][cls]["pick"][nb] for nb in ("24", "32", "36", "40", "44", "48", "54")}
    return {"source": src, "seconds": time.perf_counter() - t0, "dir_delta": dir_delta,
            "sort_picks_bits_x_kpt/window": sel, "cell_code": args.cell_code,
            "carry_stream": args.carry_stream, "carry_sort": args.carry_sort,
            "sample_sort": args.sample_sort,
            "sample_bucket": cg.tune_get("sample_bucket"), "sample_bmax": cg.tune_get("sample_bmax"),
            "canon_stab": args.canon_stab, "one_sync": args.one_sync, "knobs": knobs}


def set_knobs(cg, args):
    """Applies each --set KNOB=VALUE (cg_tune_set) and returns them."""
    out = {}
    for kv in args.set or []:
        k, v = kv.split("=", 1)
        cg.tune_set(k, int(v))
        out[k] = cg.tune_get(k)
    return out


def census(args):
    """Counts every table of the specified batched draw (seeds 0..63, end =
    start + U[100, 400]) by a bounded device-resident generation: a table
    that passes --census-max boards before its end sum is reported with the
    boards counted so far and the sum reached."""
    import cablegen_amd as cg
    from cablegen_amd.batch_configs import spec_batch, census_sizes, SPEC_LO, SPEC_HI
    cfgs = spec_batch(args.batch_tables)
    t0 = time.perf_counter()
    sizes = census_sizes(cg, cfgs, int(args.census_max))
    rows = []
    for c, (n, done, last) in zip(cfgs, sizes):
        rows.append({"seed": c["seed"], "formation": c["formation"], "initial": "0x%016x" % c["initial"][0],
                     "start": c["start"], "end": c["end"], "boards": n, "complete": done,
                     "last_sum": last})
    fit = [r for r in rows if r["complete"]]
    out = {"metric": "census of the batched draw", "tables": len(rows),
           "draw": "seeds 0..%d, end = start + 2*U{%d..%d}" % (len(rows) - 1, SPEC_LO, SPEC_HI),
           "bound_boards": int(args.census_max), "complete": len(fit),
           "complete_boards": sum(r["boards"] for r in fit),
           "seconds": time.perf_counter() - t0, "rows": rows}
    print(json.dumps(out))


def batch_bench(args):
    """Batching: 64 small tables (cablegen_amd.batch_configs, seeds 0..63)
    generated and solved together vs one at a time on the same GPU.  With
    --batch-spec the draw is the specified one (end = start + U[100, 400])
    and the run takes the tables a bounded count completes under
    --batch-fit boards each (the others cannot be kept resident)."""
    import cablegen_amd as cg
    from cablegen_amd.batch_configs import batch, spec_batch, census_sizes
    r = cg.rules(ignore_f=True)
    excluded = []
    if args.batch_spec:
        allc = spec_batch(args.batch_tables)
        sizes = census_sizes(cg, allc, int(args.batch_fit))
        cfgs = [c for c, (n, done, _) in zip(allc, sizes) if done]
        excluded = [{"seed": c["seed"], "formation": c["formation"], "boards_counted": n,
                     "last_sum": last, "end": c["end"]}
                    for c, (n, done, last) in zip(allc, sizes) if not done]
    else:
        cfgs = batch(args.batch_tables, args.batch_lo, args.batch_hi)
    tuning = run_autotune(cg, args)

    tb = cg.Tablebase(r)   # one device-resident table reused: its arena is kept

    per_table = []

    def one_at_a_time():
        tot = 0
        per_table.clear()
        t0 = time.perf_counter()
        for c in cfgs:
            ta = time.perf_counter()
            tb.generate(c["initial"], c["start"], c["end"])
            tb.solve(c["winstates"])
            n = tb.total_boards()
            per_table.append((n, time.perf_counter() - ta))
            tot += n
        return tot, time.perf_counter() - t0

    B = cg.Batch(r)
    for c in cfgs:
        B.add(c["initial"], c["start"], c["end"], c["winstates"])

    def batched():
        t0 = time.perf_counter()
        B.generate()
        t1 = time.perf_counter()
        B.solve(use_graph=True)
        t2 = time.perf_counter()
        return B.total_boards(), t1 - t0, t2 - t1

    one_at_a_time()
    batched()
    c0 = cg.tune_get("cell_code_sorts")
    r0 = cg.tune_get("sample_repeats")
    n1, t_one = one_at_a_time()
    alone = sorted(per_table, reverse=True)
    c1 = cg.tune_get("cell_code_sorts")
    r1 = cg.tune_get("sample_repeats")
    # the one-at-a-time leg's kernels (an untimed pass with every kernel timed)
    cg.prof_enable(True)
    cg.prof_select(None)
    cg.prof_reset()
    one_at_a_time()
    kern_one = {k: cg.prof_query(k) for k in ("move", "spawn", "sort_split", "sort_part", "sort_local",
                                              "sort_hist", "sort_pass", "unique", "solve")}
    cg.prof_enable(False)
    g0, i0 = cg.tune_get("graph_steps"), cg.tune_get("graph_instantiations")
    nb, tg, ts = batched()
    graph_steps = cg.tune_get("graph_steps") - g0
    graph_inst = cg.tune_get("graph_instantiations") - i0
    c2 = cg.tune_get("cell_code_sorts")
    r2 = cg.tune_get("sample_repeats")
    assert n1 == nb
    # A/B: the same batched launches enqueued one by one instead of one graph
    # per layer step; then one such run with every kernel timed (event
    # records stay out of graphs) for the per-kernel breakdown
    try:
        cg.tune_set("batch_graph", 0)
        _, tg_eager, ts_eager = batched()
        cg.prof_enable(True)
        cg.prof_select(None)
        cg.prof_reset()
        batched()
        kern = {k: cg.prof_query(k) for k in ("move", "spawn", "sort_split", "sort_part", "sort_local",
                                              "sort_hist", "sort_pass", "unique")}
        cg.prof_enable(False)
    finally:
        cg.tune_set("batch_graph", 1)
    sizes = sorted((B.total_boards_table(t) for t in range(len(cfgs))), reverse=True)
    # The largest tables, holding half of the boards, run alone at a rate
    # that batching cannot raise (each already fills the GPU): at that rate
    # for every board the batch would take nb / rate, which bounds its
    # speed-up over the one-at-a-time leg.
    half, acc, t_big = 0, 0, 0.0
    for n, t in alone:
        if acc >= n1 / 2:
            break
        acc += n
        t_big += t
        half += 1
    big_rate = acc / t_big if t_big > 0 else 0.0
    alone_out = {"tables_holding_half_the_boards": half, "their_boards": acc, "their_seconds": t_big,
                 "their_boards_per_s": big_rate,
                 "rest_seconds": t_one - t_big,
                 "speedup_if_every_board_ran_at_their_rate": (t_one / (n1 / big_rate)) if big_rate else None,
                 "per_table_boards_seconds": [[n, round(t, 6)] for n, t in alone]}
    out = {"metric": "batched tablebases: boards/s, batched vs one at a time",
           "tables": len(cfgs), "boards": nb, "one_at_a_time_s": t_one,
           "batched_generate_s": tg, "batched_solve_s": ts,
           "one_at_a_time_boards_per_s": n1 / t_one, "batched_boards_per_s": nb / (tg + ts),
           "speedup": t_one / (tg + ts),
           "batched_step_graphs": graph_steps, "graph_instantiations": graph_inst,
           "batched_eager_generate_s": tg_eager, "batched_eager_solve_s": ts_eager,
           "speedup_eager_steps": t_one / (tg_eager + ts_eager),
           "largest_tables": sizes[:5], "median_table": sizes[len(sizes) // 2],
           "one_at_a_time_tables": alone_out,
           "batched_kernels_ms": {k: v[1] for k, v in kern.items()},
           "one_at_a_time_kernels": {k: [v[0], v[1]] for k, v in kern_one.items()},
           "sample_sort_repeats": {"one_at_a_time": r1 - r0, "batched": r2 - r1},
           "one_sync": cg.tune_get("one_sync"),
           "batched_kernels_source": "one batched run with its launches enqueued one by one, every kernel timed",
           "batched_sorts": kern["sort_hist"][0], "batched_sort_passes": kern["sort_pass"][0],
           "cell_coded_sorts": {"one_at_a_time": c1 - c0, "batched": c2 - c1},
           "tuning": tuning,
           "workload": ("specified draw, seeds 0..%d: end = start + 2*U{50..200}; the %d tables "
                        "whose bounded count completed under %d boards each (the others are listed "
                        "in excluded)" % (args.batch_tables - 1, len(cfgs), int(args.batch_fit))
                        if args.batch_spec else
                        "seeds 0..%d: formation in {2x4, 3x4, DPDGap}, 2-4 tiles of 2^1..2^7, "
                        "goal 2^(max+2) top-left, end = start + 2*U{%d..%d}, ignore_f, no prune"
                        % (len(cfgs) - 1, args.batch_lo, args.batch_hi)),
           "excluded": excluded}
    B.close()
    tb.close()
    print(json.dumps(out))


def dropin_bench(args):
    """C2 through the reference's entry points (api/cablegen.h generate() and
    solve()): layers and tables are written to and read from files in
    args.dropin, as the reference does.  Reports end-to-end boards/s, the
    files' bytes, and how much of the time is not GPU work."""
    import ctypes as C
    import shutil
    import numpy as np
    import cablegen_amd as cg
    d = os.path.abspath(args.dropin)
    os.makedirs(d + "/b", exist_ok=True)
    os.makedirs(d + "/t", exist_ok=True)
    conf = d + "/c2.conf"
    with open(conf, "w") as f:
        f.write("[Cablegen]\nignore_f = true\n[Generate]\nprune = true\nstsl = 58\nltc = 5\n"
                "smallest_large = 6\ndir = %s/b/\nend = %d\n[Solve]\ndir = %s/t/\nend = 0\n"
                % (d, END_GEN, d))
    L = cg.lib()
    L.change_config(conf.encode())
    tuning = run_autotune(cg, args)
    init = np.array(INITIAL, np.uint64)
    win = np.array(WINSTATES, np.uint64)
    sa = cg.SArr(True, init.ctypes.data_as(cg.U64P), len(init))
    wa = cg.SArr(True, win.ctypes.data_as(cg.U64P), len(win))
    start = L.get_sum(INITIAL[0])
    free0 = shutil.disk_usage(d).free
    t0 = time.perf_counter()
    L.generate(start, END_GEN, ("%s/b/%%d.boards" % d).encode(), C.byref(sa))
    t1 = time.perf_counter()
    bbytes = sum(os.path.getsize(os.path.join(d, "b", f)) for f in os.listdir(d + "/b"))
    out = {"metric": "boards/s of generate() + solve() through api/cablegen.h, files included",
           "workload": "DPDGap-512 (the bench's C2) written to / read from %s" % d,
           "generate_s": t1 - t0, "boards_file_bytes": bbytes, "free_disk_bytes_before": free0,
           "tuning": tuning}
    if not args.dropin_generate_only:
        t2 = time.perf_counter()
        L.solve(END_GEN, 0, ("%s/b/%%d.boards" % d).encode(), ("%s/t/%%d.tables" % d).encode(),
                C.byref(wa))
        t3 = time.perf_counter()
        out["solve_s"] = t3 - t2
        out["tables_file_bytes"] = sum(os.path.getsize(os.path.join(d, "t", f))
                                       for f in os.listdir(d + "/t"))
    n = 0
    for f in os.listdir(d + "/b"):
        n += os.path.getsize(os.path.join(d, "b", f)) // 8
    out["boards"] = n
    total_t = out["generate_s"] + out.get("solve_s", 0.0)
    out["value"] = n / total_t
    out["unit"] = "boards/s"
    if not args.dropin_keep:
        shutil.rmtree(d + "/b", ignore_errors=True)
        shutil.rmtree(d + "/t", ignore_errors=True)
    print(json.dumps(out))


def stage_roofline(cg, dom, steps):
    """roofline entry of the dominant stage over its timed launches, per
    stream.  For the sort stage, a sort's algorithmic bytes are 8 B per key
    in and 8 B per unique key out (carried by its final kernel), over the
    time of all of its kernels; the others are single kernels."""
    kinds = SORT_KINDS if dom == "sort" else [dom]
    final = SORT_FINAL if dom == "sort" else [dom]
    lanes = []
    for st in (0, 1):
        ms = b = 0.0
        n = 0
        per = {}
        for k in kinds:
            c, m, by = cg.prof_query_stream(k, st)
            if c <= 0:
                continue
            ms += m
            per[k] = {"launches": c // max(steps, 1), "ms_per_step": m / max(steps, 1)}
            if k in final:
                b += by
                n += c
        lanes.append({"units": n, "ms": ms, "bytes": b, "kernels": per})
    m0 = lanes[0]
    avg_ms = m0["ms"] / max(m0["units"], 1)
    avg_bytes = m0["bytes"] / max(m0["units"], 1)
    achieved = avg_bytes / (avg_ms * 1e6) if avg_ms else 0.0
    s1 = lanes[1]
    traffic, tsrc = stage_traffic(kinds, final)
    out = {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
           "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
           "traffic_source": tsrc,
           "traffic_over_algorithmic": (traffic / avg_bytes) if traffic and avg_bytes else None,
           "units_per_step": m0["units"] // max(steps, 1),
           "avg_unit_us": avg_ms * 1e3, "avg_algorithmic_bytes": avg_bytes,
           "stream": "main (the second stream's units, which overlap the next step's moves, are "
                     "in second_stream)",
           "kernels_main_stream": m0["kernels"],
           "second_stream": {"units": s1["units"] // max(steps, 1),
                             "avg_unit_us": 1e3 * s1["ms"] / max(s1["units"], 1),
                             "avg_algorithmic_bytes": s1["bytes"] / max(s1["units"], 1),
                             "achieved": (s1["bytes"] / s1["ms"] / 1e6) if s1["ms"] else 0.0,
                             "kernels": s1["kernels"]}}
    if dom == "sort":
        out["unit_of_work"] = ("one sort + dedupe (main stream: the move candidates of a layer "
                               "step -> M_s); achieved = (8 N_cand + 8 N_uniq) / time of all of "
                               "the sort's kernels")
    return out


def stage_traffic(kinds, final):
    """HBM bytes per unit of the stage (main stream) from the calibrated PMC
    record in profiles/pmc_traffic.json (C2 workload), or None"""
    tf = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    if not os.path.exists(tf):
        return None, None
    with open(tf) as f:
        tj = json.load(f)
    km = tj.get("kernels_main_stream", {})
    if tj.get("workload") != "c2" or not any(k in km for k in final):
        return None, None
    units = sum(km[k]["launches"] for k in final if k in km)
    tot = sum((v["read_bytes_per_launch"] + v["write_bytes_per_launch"]) * v["launches"]
              for k, v in km.items() if k.split(":")[0] in kinds)
    return (tot / units if units else None,
            "profiles/pmc_traffic.json (rocprofv3 FETCH_SIZE / WRITE_SIZE per dispatch, calibrated "
            "by a 1 GiB copy; main-stream dispatches of the stage's kernels per unit)")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-isolated", action="store_true",
                    help="skip the untimed step that times the sorts with no second-stream overlap")
    ap.add_argument("--adjoint-reps", type=int, default=3, help="sweeps averaged for the adjoint ratio")
    ap.add_argument("--cpu-seconds", type=float, default=20.0,
                    help="CPU baseline: about this much oracle time on sampled layer windows")
    ap.add_argument("--cpu-full", action="store_true",
                    help="CPU baseline: the whole table (generate + solve) instead of a sample")
    ap.add_argument("--micro", action="store_true", help="kernel roofline microbench instead")
    ap.add_argument("--micro-log2", type=int, default=30, help="log2 of the microbench board count")
    ap.add_argument("--batch", action="store_true", help="batched-tables benchmark instead")
    ap.add_argument("--dropin", default=None,
                    help="C2 through generate()/solve() with files in this directory instead")
    ap.add_argument("--dropin-generate-only", action="store_true")
    ap.add_argument("--dropin-keep", action="store_true", help="keep the written files")
    ap.add_argument("--batch-tables", type=int, default=64)
    ap.add_argument("--batch-lo", type=int, default=5, help="end = start + 2*U{lo..hi}")
    ap.add_argument("--batch-hi", type=int, default=20)
    ap.add_argument("--batch-spec", action="store_true",
                    help="the specified draw (end = start + U[100,400]), tables that fit --batch-fit")
    ap.add_argument("--batch-fit", type=float, default=2e8,
                    help="--batch-spec: per-table bound on the boards of a table that is kept")
    ap.add_argument("--census", action="store_true", help="count the specified batched draw's tables")
    ap.add_argument("--census-max", type=float, default=1e9, help="--census: per-table bound (boards)")
    ap.add_argument("--prof-summary", default=None, help="summarise rocprofv3 output in this directory")
    ap.add_argument("--prof-rows", type=int, default=14)
    ap.add_argument("--pmc", action="store_true", help="with --prof-summary: PMC traffic / L2 summary")
    ap.add_argument("--pmc-calib", default=None, help="directory holding the calibration-copy PMC runs")
    ap.add_argument("--pmc-tag", default=None, help="only PMC runs whose path contains this")
    ap.add_argument("--pmc-json", default=None, help="write per-launch traffic here (JSON)")
    ap.add_argument("--pmc-workload", default="", help="workload label stored in --pmc-json")
    ap.add_argument("--tuning", choices=("stored", "auto", "off"), default="stored",
                    help="sort/directory tuning: stored autotuner table, run the autotuner, or defaults")
    ap.add_argument("--save-tuning", default=None, help="write the autotuner report (JSON) here")
    ap.add_argument("--carry-sort", type=int, choices=(0, 1), default=0,
                    help="1: each step's spawn-4 carry is sorted and deduplicated before it is "
                         "moved (on the stream --carry-stream names); 0: moved as written")
    ap.add_argument("--carry-stream", type=int, choices=(0, 1), default=1,
                    help="0: the carry sort on the main stream instead of a second one")
    ap.add_argument("--dir-delta", type=int, choices=range(-4, 4), default=None,
                    help="layer directory size: 2^(floor(log2 n) + delta) buckets (overrides the tuned pick)")
    ap.add_argument("--sample-sort", type=int, choices=(0, 1), default=1,
                    help="1: mid-size sorts take the sample sort (default); 0: LSD radix only")
    ap.add_argument("--sample-bucket", type=int, default=0,
                    help="mean keys per bucket of the sample sort (0: the library's default)")
    ap.add_argument("--set", action="append", metavar="KNOB=VALUE",
                    help="set a library knob (cg_tune_set) after the tuning table; repeatable")
    ap.add_argument("--sample-bmax", type=int, choices=(2048, 8192, 16384), default=0,
                    help="most buckets of a sample sort (0: the library's default)")
    ap.add_argument("--cell-code", type=int, choices=(0, 1), default=1,
                    help="1: layer sorts code pairs of cells when the tile sum bounds them (default)")
    ap.add_argument("--one-sync", type=int, choices=(0, 1), default=1,
                    help="1: layer steps with one host synchronisation (M sort sized on the device; "
                         "default); 0: the candidate count read before the M sort")
    ap.add_argument("--canon-stab", type=int, choices=(0, 1), default=1,
                    help="1: canonical forms through the formation's stabiliser (default); 0: max of 8 images")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this engine runs on one GPU; --gpus {args.gpus} is not supported "
              "(layers would shard by key range across GPUs, which is not built)", file=sys.stderr)
        sys.exit(2)

    if args.prof_summary:
        prof_summary(args)
        return
    if args.micro:
        micro(args)
        return
    if args.batch:
        batch_bench(args)
        return
    if args.census:
        census(args)
        return
    if args.dropin:
        dropin_bench(args)
        return

    import cablegen_amd as cg
    r = cg.rules(**RULES)
    tb = cg.Tablebase(r)
    start = cg.lib().get_sum(INITIAL[0])
    tuning = run_autotune(cg, args)

    def step():
        t0 = time.perf_counter()
        tb.generate(INITIAL, start, END_GEN)
        t1 = time.perf_counter()
        tb.solve(WINSTATES)
        t2 = time.perf_counter()
        return t1 - t0, t2 - t1

    for _ in range(args.warmup):
        step()
    ss0, sr0 = cg.tune_get("sample_sorts"), cg.tune_get("sample_repeats")
    # one untimed step with every kernel timed gives the per-kernel breakdown
    # and names the dominant stage; the timed steps then time only that
    # stage's launches (two events per launch), for the roofline
    cg.prof_enable(True)
    cg.prof_select(None)
    cg.prof_reset()
    cg.tune_set("prof_count", 0)   # the production kernels decide the stage
    step()
    cg.lib().cg_dev_sync()
    cg.tune_set("prof_count", 1)
    kern = {}
    for k in KERNELS:
        n, ms, b = cg.prof_query(k)
        kern[k] = dict(launches=n, ms=ms, bytes=b)
    stage_ms = {k: kern[k]["ms"] for k in ("move", "spawn", "solve")}
    stage_ms["sort"] = sum(kern[k]["ms"] for k in SORT_KINDS)
    dom = max(STAGES, key=lambda k: stage_ms[k])
    # the second stream's units from this step (every launch timed); the
    # timed steps time the main stream's launches only (fewer events in the
    # timed region)
    second = stage_roofline(cg, dom, 1)["second_stream"]
    second["source"] = "the untimed step with every kernel timed"
    # the solve's lookups are counted in one more untimed step (the counting
    # kernel instance); the timed steps time the production kernel and take
    # its algorithmic bytes per launch from there (the same layers, the same
    # lookups)
    solve_counted = (0.0, 0)
    if dom == "solve":
        cg.prof_select("solve")
        cg.prof_reset()
        step()
        cg.lib().cg_dev_sync()
        n_c, _, b_c = cg.prof_query("solve")
        solve_counted = (b_c, n_c)
        kern["solve"]["bytes"] = b_c
        cg.tune_set("prof_count", 0)
    cg.prof_select(dom + "@main")
    cg.prof_reset()
    cg.lib().cg_dev_sync()
    gen_t, sol_t = [], []
    for _ in range(args.steps):
        g, s = step()
        gen_t.append(g)
        sol_t.append(s)
    cg.lib().cg_dev_sync()
    total = tb.total_boards()
    T = (sum(gen_t) + sum(sol_t)) / args.steps
    value = total / T

    roof = stage_roofline(cg, dom, args.steps)
    if dom == "solve":
        cg.tune_set("prof_count", 1)
        avg_bytes = solve_counted[0] / max(solve_counted[1], 1)
        roof["avg_algorithmic_bytes"] = avg_bytes
        roof["achieved"] = avg_bytes / (roof["avg_unit_us"] * 1e3) if roof["avg_unit_us"] else 0.0
        roof["frac"] = roof["achieved"] / HBM_PEAK_GBS
        if roof.get("traffic"):
            roof["traffic_over_algorithmic"] = roof["traffic"] / avg_bytes
        roof["unit_of_work"] = ("one layer's solve launch; algorithmic bytes 16 N + 24 lookups + 8 hits "
                                "(+ 8 N with argmax), the lookups counted in the untimed step, the "
                                "timed steps running the production kernel without counters")
    roof.pop("second_stream", None)
    if second["units"]:
        roof["second_stream"] = second
    else:
        roof["stream"] = "main (no sort runs on the second stream: the carry is moved unsorted)"
    if dom == "sort" and args.carry_sort == 1 and args.carry_stream == 1 and not args.no_isolated:
        # the same stage with nothing beside it: one untimed step with the
        # carry sort on the main stream, so no kernel of another stream
        # shares the GPU with a sort's kernels while they are timed
        cg.prof_reset()
        cg.prof_select(dom)
        cg.tune_set("carry_stream", 0)
        step()
        cg.lib().cg_dev_sync()
        iso = stage_roofline(cg, dom, 1)
        cg.tune_set("carry_stream", 1)
        roof["isolated"] = {"achieved": iso["achieved"], "frac": iso["frac"],
                            "avg_unit_us": iso["avg_unit_us"], "units": iso["units_per_step"],
                            "avg_algorithmic_bytes": iso["avg_algorithmic_bytes"],
                            "source": "one untimed step with the carry sort on the main stream: "
                                      "every sort of the step (move candidates and carries) timed "
                                      "with no other stream's kernels beside it (the timed steps "
                                      "overlap the candidate sorts with the carry sorts)"}
    cg.prof_select(None)
    cg.prof_enable(False)

    out = {
        "metric": METRIC, "value": value, "unit": "boards/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": T * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u64 keys / f64 values", "data": "synthetic (the DPDGap-512 table itself)",
        "config": {"workload": "DPDGap-512 tablebase: initial 0x00008765FFF3FFF2 (sum 492), "
                               "ignore_f, prune stsl=58 ltc=5 smallest_large=6, "
                               "generate 492..1100, solve 1100..0, winstate 0x0000900000000000",
                   "layers": tb.num_layers, "boards": total,
                   "gen_ms": 1e3 * sum(gen_t) / args.steps, "solve_ms": 1e3 * sum(sol_t) / args.steps},
        "roofline": roof,
        "tuning": tuning,
        "kernels": {k: {"launches": v["launches"], "ms_per_step": v["ms"],
                        "GB_s": (v["bytes"] / v["ms"] / 1e6) if v["ms"] else 0.0}
                    for k, v in kern.items()},
        "stages_ms_per_step": stage_ms,
        "sorts": {"sample_sorts_per_step": (cg.tune_get("sample_sorts") - ss0) / (args.steps + 1),
                  "lsd_repeats_per_step": (cg.tune_get("sample_repeats") - sr0) / (args.steps + 1)},
        "kernels_source": "one untimed step after the warm-up with every kernel timed "
                          "(the timed steps time only the roofline stage's kernels)",
    }

    if not args.no_adjoint:
        # mean over several sweeps: plain solve, solve keeping the argmax, adjoint
        tb.solve(WINSTATES)
        cg.lib().cg_dev_sync()
        reps = max(1, args.adjoint_reps)
        ts, ta, tj = [], [], []
        for _ in range(reps):
            t0 = time.perf_counter()
            tb.solve(WINSTATES)
            t1 = time.perf_counter()
            tb.solve(WINSTATES, keep_argmax=True)
            t2 = time.perf_counter()
            P, dP, D = tb.adjoint(INITIAL)
            t3 = time.perf_counter()
            ts.append(t1 - t0)
            ta.append(t2 - t1)
            tj.append(t3 - t2)
        ms, ma, mj = (1e3 * sum(x) / reps for x in (ts, ta, tj))
        out["adjoint"] = {"P_win": P, "dP_dp4": dP, "sweeps": reps, "plain_solve_ms": ms,
                          "solve_with_argmax_ms": ma, "adjoint_sweep_ms": mj,
                          "adjoint_sweep_ms_each": [1e3 * x for x in tj],
                          "adjoint_over_solve": mj / ms,
                          "argmax_solve_plus_adjoint_over_solve": (ma + mj) / ms}

    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(tb, cg, baseline_cores(), args.cpu_seconds, args.cpu_full)
        out["speedup_vs_cpu"] = value / out["cpu_baseline"]["value"]
    tb.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
