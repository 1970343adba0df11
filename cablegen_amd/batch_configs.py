"""The batched workload: many small tablebases of uneven sizes.

table_config(seed) draws one table (numpy Generator seeded with `seed`):
  * formation uniform over {2x4, 3x4 (bottom row 0xF), DPDGap};
  * 2-4 tiles with exponents uniform in 1..7 in distinct random free cells;
  * goal 2^(max exponent + 2) in the top-left cell (free in all three);
  * end tile sum = initial sum + 2 * U{lo..hi}.  Without pruning these
    formations grow about 2-4x per 10 of tile sum (3x4 fastest), so an offset
    of 100..400 makes every table a full-GPU job of 1e8-1e10 boards.  The
    default U{5..20} (offset 10..40) gives 64 tables of 5e3..3.6e7 boards
    (1.6e8 in all, median 1.5e5); U{10..40} gives 4e3..4.9e8 (1.7e9 in all,
    five tables hold two thirds of the boards);
  * rules: formation cells as walls (ignore_f), no pruning.
"""
import numpy as np

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
