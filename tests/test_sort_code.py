"""The sort's cell-pair key code, host view (no GPU work).

A generation's sorts may code each pair of varying cells (values 0..B-1) as
a * B + b.  The radix passes then see fewer bits, and the unique kernel
decodes the codes back to keys.  The layers stay identical only if the code
is a bijection on the keys that keeps their numeric order.  These tests check
exactly that: the code is built from the library's own host functions, and
the checks are order and round trips, not a second copy of the formula.
"""
import ctypes as C
import itertools

import numpy as np
import pytest


@pytest.fixture(scope="module")
def cg():
    import cablegen_amd
    cablegen_amd.lib()
    return cablegen_amd


def bounded(v):
    """a cell_max word bounding every cell by v"""
    return v * 0x1111111111111111


def probe(cg, keys, vary, cell_max):
    keys = np.ascontiguousarray(keys, dtype=np.uint64)
    codes = np.zeros(max(len(keys), 1), np.uint64)
    back = np.zeros_like(codes)
    nb = cg.lib().cg_sort_code_probe(keys.ctypes.data, len(keys), C.c_uint64(vary),
                                     C.c_uint64(cell_max), codes.ctypes.data, back.ctypes.data)
    return nb, codes[: len(keys)], back[: len(keys)]


def keys_with_cells(rng, n, const, cells, vmax):
    """n keys equal to `const` except in `cells` (0 = the most significant
    nibble), which hold values 0..vmax"""
    k = np.full(n, const, np.uint64)
    for c in cells:
        sh = np.uint64(60 - 4 * c)
        k &= ~(np.uint64(15) << sh)
        k |= rng.integers(0, vmax + 1, n).astype(np.uint64) << sh
    return k


def vary_of(k):
    o = np.bitwise_or.reduce(k)
    a = np.bitwise_and.reduce(k)
    return int(o & ~a)


def check_order_and_round_trip(keys, codes, back, nbits):
    assert np.array_equal(back, keys)
    assert nbits < 64 and np.all(codes < (np.uint64(1) << np.uint64(nbits)))
    order = np.argsort(keys, kind="stable")
    ks, cs = keys[order], codes[order]
    # order kept; equal codes exactly where keys are equal
    assert np.all(cs[1:] >= cs[:-1])
    assert np.array_equal(ks[1:] == ks[:-1], cs[1:] == cs[:-1])


# DPDGap layers: 0xf cells at cells 0,1,2,4,5,6, ten varying cells
DPD_CONST = 0xFFF0FFF000000000
DPD_CELLS = [3, 7, 8, 9, 10, 11, 12, 13, 14, 15]


@pytest.mark.parametrize("vmax,bits", [(8, 35), (9, 35), (10, 35), (7, 30), (5, 30), (3, 20),
                                       (11, 40), (14, 40)])
def test_dpdgap_cells(cg, vmax, bits):
    """Ten varying cells: a pair of cells of values <= vmax takes the
    smallest field holding (vmax + 1)^2 values: 7 bits up to 10 (the DPDGap
    layers of tile sums 492..1100), 6 bits up to 7, 8 bits from 11."""
    rng = np.random.default_rng(vmax)
    k = keys_with_cells(rng, 20000, DPD_CONST, DPD_CELLS, vmax)
    k[0] = DPD_CONST | int(sum(vmax << (60 - 4 * c) for c in DPD_CELLS))   # every cell at the bound
    k[1] = DPD_CONST
    vary = vary_of(k)
    nb, codes, back = probe(cg, k, vary, bounded(vmax))
    assert nb == bits
    check_order_and_round_trip(k, codes, back, nb)


def test_exhaustive_small(cg):
    """Every key of three varying cells of values 0..2 (an odd cell count):
    codes strictly increase in key order, and each decodes to its key."""
    const = 0x1234000000000000
    cells = [5, 9, 14]
    keys = []
    for vals in itertools.product(range(3), repeat=3):
        k = const
        for c, v in zip(cells, vals):
            k |= v << (60 - 4 * c)
        keys.append(k)
    k = np.array(sorted(keys), np.uint64)
    nb, codes, back = probe(cg, k, vary_of(k), bounded(2))
    assert nb == 4 + 2   # one pair of 9 values, one cell of 3
    assert np.all(codes[1:] > codes[:-1])
    assert np.array_equal(back, k)


@pytest.mark.parametrize("ncell", [1, 2, 5, 16])
def test_cell_counts(cg, ncell):
    """One to sixteen varying cells, values up to 6; the constant cells come
    back from the keys themselves."""
    rng = np.random.default_rng(ncell)
    cells = sorted(rng.choice(16, ncell, replace=False).tolist())
    const = int(rng.integers(0, 1 << 62)) | (1 << 63)
    k = keys_with_cells(rng, 5000, const, cells, 6)
    vary = vary_of(k)
    nb, codes, back = probe(cg, k, vary, bounded(6))
    nvary = sum(1 for c in range(16) if (vary >> (60 - 4 * c)) & 15)
    assert nb == (nvary // 2) * 6 + (nvary % 2) * 3
    check_order_and_round_trip(k, codes, back, nb)


def test_partially_varying_cells(cg):
    """A cell counts as varying when any of its bits varies; its constant
    bits are coded with it and come back."""
    k = np.array([0x0000000000000070, 0x0000000000000071, 0x0000000000000060,
                  0x0000000000000861], np.uint64)
    nb, codes, back = probe(cg, k, vary_of(k), bounded(8))
    assert nb == 7 + 4   # cells 13, 14, 15 vary: a pair of 81 values, a cell of 9
    check_order_and_round_trip(k, codes, back, nb)


def test_per_cell_bases(cg):
    """Composite keys as the batch builds them: a table id in bits 48..55
    (ids 0..39: the high nibble holds 0..2, the low one 0..15) above ten free
    cells of values 0..8.  The id's two cells form one 6-bit field (3 * 16
    values), each pair of free cells a 7-bit field (81 values)."""
    rng = np.random.default_rng(5)
    n = 30000
    tid = rng.integers(0, 40, n).astype(np.uint64)
    k = keys_with_cells(rng, n, 0, list(range(6, 16)), 8) | (tid << np.uint64(48))
    cell_max = (2 << 52) | (15 << 48) | (bounded(8) >> 16)
    nb, codes, back = probe(cg, k, vary_of(k), cell_max)
    assert nb == 6 + 5 * 7
    check_order_and_round_trip(k, codes, back, nb)


def test_unbounded_and_empty(cg):
    """Cells bounded by 15 take 8 bits a pair (no gain, so sorts keep bit
    packing); no varying bits, no code."""
    k = np.array([0x12, 0x34, 0x56], np.uint64)
    nb, codes, back = probe(cg, k, vary_of(k), ~0 & (2**64 - 1))
    assert nb == 8
    check_order_and_round_trip(k, codes, back, nb)
    assert probe(cg, k, 0, bounded(5))[0] == -1
