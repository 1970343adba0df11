"""The batched workload: many small tablebases of uneven sizes.

table_config(seed) draws one table (numpy Generator seeded with `seed`):
  * formation uniform over {2x4, 3x4 (bottom row 0xF), DPDGap};
  * 2-4 tiles with exponents uniform in 1..7 in distinct random free cells;
  * goal 2^(max exponent + 2) in the top-left cell (free in all three);
  * end tile sum = initial sum + 2 * U{lo..hi}.  The specified batch draws
    the offset from U[100, 400] (SPEC_LO, SPEC_HI = 50, 200).  Without
    pruning, the 3x4 and DPDGap formations grow about 2-4x per 10 of tile
    sum, so most of those tables hold far more boards than one GPU can keep
    resident: `python bench.py --census` counts every table of the draw
    (bounded), and spec_fitting() keeps the tables whose count completed
    under a bound.  The smaller draw U{5..20} (offset 10..40; 64 tables of
    5e3..3.6e7 boards) is kept as a second workload;
  * rules: formation cells as walls (ignore_f), no pruning.
"""
import numpy as np

SPEC_LO, SPEC_HI = 50, 200   # end = start + 2 * U{50..200}, i.e. start + U[100, 400] in steps of 2

FORMATIONS = {
    "2x4": 0x00000000FFFFFFFF,
    "3x4": 0x000000000000FFFF,
    "DPDGap": 0x00000000FFF0FFF0,
}


def tile_sum(b):
    s = 0
    for x in range(16):
        v = (b >> (60 - 4 * x)) & 0xf
        if 0 < v < 0xe:
            s += 1 << v
    return s


def table_config(seed, lo=5, hi=20):
    rng = np.random.default_rng(seed)
    name = ["2x4", "3x4", "DPDGap"][int(rng.integers(0, 3))]
    fmask = FORMATIONS[name]
    free = [x for x in range(16) if not (fmask >> (60 - 4 * x)) & 0xf]
    ntiles = int(rng.integers(2, 5))
    cells = rng.choice(free, size=ntiles, replace=False)
    exps = rng.integers(1, 8, size=ntiles)
    b = fmask
    for c, e in zip(cells, exps):
        b |= int(e) << (60 - 4 * int(c))
    start = tile_sum(b)
    end = start + 2 * int(rng.integers(lo, hi + 1))
    goal = (int(exps.max()) + 2) << 60
    return dict(seed=seed, formation=name, initial=[b], winstates=[goal], start=start, end=end,
                rules=dict(ignore_f=True))


def batch(n=64, lo=5, hi=20):
    return [table_config(s, lo, hi) for s in range(n)]


def spec_batch(n=64):
    """the specified draw: seeds 0..n-1, end = start + 2 * U{50..200}"""
    return batch(n, SPEC_LO, SPEC_HI)


def census_sizes(cg, cfgs, max_boards):
    """per table: (total boards, complete) from a bounded device-resident
    generation; complete is False when the table passed max_boards before
    its end sum (the total then counts the layers generated so far)"""
    out = []
    for c in cfgs:
        tb = cg.Tablebase(cg.rules(**c["rules"]))
        try:
            last = tb.generate_bounded(c["initial"], c["start"], c["end"], max_boards)
            out.append((tb.total_boards(), last >= c["end"], last))
        except cg.CablegenError:
            # device memory ran out inside a step: the layers before it are counted
            out.append((tb.total_boards(), False, c["start"] + 2 * (tb.num_layers - 1)))
        finally:
            tb.close()
    return out
