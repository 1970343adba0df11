"""The canonical form through a formation's stabiliser, host view (no GPU
work).  For boards whose 0xF cells are a formation's and whose other cells
are below 0xF, only the images that put the formation at its canonical
place can be the largest, when the formation decides every other image (see
cg_canon_fast in cablegen_amd/csrc/board.h).  The kernels take that max
for boards whose fmask matches and the full max of 8 images otherwise; the
checks here run the library's own host code (cg_canon_sel_probe) against
the oracle's canonicalize_b restatement (oracle/oracle.c orc_canon,
following src/board.c:314-337) on random boards of every orientation, and
pin how many images each formation keeps."""
import numpy as np
import pytest

from tests import oracle_lib as O

FORMATIONS = {
    "dpdgap (canonical place)": 0xFFF0FFF000000000,
    "dpdgap (bottom right)": 0x00000000FFF0FFF0,
    "2x4 top": 0xFFFFFFFF00000000,
    "2x4 bottom": 0x00000000FFFFFFFF,
    "bottom row": 0x000000000000FFFF,
    "one corner": 0xF000000000000000,
    "two corners": 0xF00F000000000000,
    "L": 0xFF00F00000000000,
    "column": 0xF000F000F000F000,
    "none": 0,
}
# images kept for boards as the layer holds them (orientation 0) when the
# formation is at its canonical place; None: the formation does not decide
KEPT = {"dpdgap (canonical place)": 1, "2x4 top": 2, "one corner": 2, "none": 8,
        "two corners": None}


@pytest.fixture(scope="module")
def cg():
    import cablegen_amd
    cablegen_amd.lib()
    return cablegen_amd


def rot_cw(b):
    return O.orc().orc_rotate_cw(int(b))


def fast(cg, e, ff, boards, o):
    boards = np.ascontiguousarray(boards, dtype=np.uint64)
    out = np.zeros(max(len(boards), 1), np.uint64)
    sel = np.zeros(2, np.uint32)
    assert cg.lib().cg_canon_sel_probe(e, ff, boards.ctypes.data, len(boards), o, out.ctypes.data,
                                       sel.ctypes.data) == 0
    return out[: len(boards)], sel


def boards_with(rng, form, n):
    """n boards with 0xF exactly at form's 0xF cells, other cells 0..0xE"""
    cells = [x for x in range(16) if (form >> (60 - 4 * x)) & 0xF == 0xF]
    v = rng.integers(0, 15, size=(n, 16), dtype=np.uint64)
    v[:, cells] = 0xF
    sh = np.array([60 - 4 * x for x in range(16)], dtype=np.uint64)
    return np.bitwise_or.reduce(v << sh, axis=1)


@pytest.mark.parametrize("name", list(FORMATIONS))
@pytest.mark.parametrize("ff", [0, 1])
def test_fast_canon_equals_full(cg, name, ff):
    e = FORMATIONS[name]
    rng = np.random.default_rng(abs(hash(name)) % 2**32 + ff)
    for o in (0, 1):
        form = rot_cw(e) if o else e
        b = np.concatenate([boards_with(rng, form, 3000),
                            boards_with(rng, 0, 1000),               # no 0xF cells
                            rng.integers(0, 1 << 63, size=1000, dtype=np.uint64) * 2 + 1])
        got, sel = fast(cg, e, ff, b, o)
        exp = np.array([O.canon(int(x)) for x in b], dtype=np.uint64)
        assert np.array_equal(got, exp), (name, ff, o)
        if ff:
            assert list(sel) == [0xff, 0xff]


@pytest.mark.parametrize("name", list(KEPT))
def test_images_kept(cg, name):
    """how many images each orientation keeps: one for DPDGap, two for a
    formation with a mirror symmetry, all eight without a formation, and none
    left out when the formation does not decide (two corners of a row: the
    cells between them are free)"""
    e = FORMATIONS[name]
    _, sel = fast(cg, e, 0, np.zeros(0, np.uint64), 0)
    k = KEPT[name]
    if k is None:
        assert list(sel) == [0xff, 0xff]
    else:
        assert bin(int(sel[0])).count("1") == k and bin(int(sel[1])).count("1") == k


def test_dpdgap_layer_orientations(cg):
    """DPDGap boards as a layer holds them keep the identity image, and boards
    turned clockwise by an up or down move keep only the image that turns
    them back (flip o mirror o cw = counter-clockwise)"""
    _, sel = fast(cg, FORMATIONS["dpdgap (canonical place)"], 0, np.zeros(0, np.uint64), 0)
    assert int(sel[0]) == 0x01 and int(sel[1]) == 0x80
