/*
 * board.h -- nibble bitboard primitives shared by the host C ABI and the
 * gfx950 kernels (compiled as __host__ __device__ under hipcc and as plain C
 * by the host compiler).
 *
 * Encoding (api/cablegen.h:9-12): tile x in [0,16) is the nibble at bits
 * [60-4x, 63-4x]; row r is the 16-bit group at bits [48-16r, 63-16r]; within
 * a row value column 0 is the most significant nibble.  0 = empty, 1..13 =
 * 2^v, 0xE = unmergeable non-formation tile, 0xF = formation cell.
 *
 * The implementations are bit-parallel (no per-tile loops on the hot path):
 * rotations and mirrors are mask/shift networks, the lock test is a nibble
 * "all ones" mask, and canonicalisation is the unsigned max of the 8 images.
 */
#ifndef CABLEGEN_BOARD_H
#define CABLEGEN_BOARD_H
#include <stdint.h>

#if defined(__HIPCC__)
#define CG_HD __host__ __device__ __forceinline__
#else
#define CG_HD static inline
#endif

#define CG_NIB_LO 0x1111111111111111ULL

/* mask with 0x1 in every nibble equal to 0xF */
CG_HD uint64_t cg_fmask(uint64_t b)
{
    return b & (b >> 1) & (b >> 2) & (b >> 3) & CG_NIB_LO;
}

/* mask with 0x1 in every nibble equal to 0 */
CG_HD uint64_t cg_zmask(uint64_t b)
{
    uint64_t t = b | (b >> 1) | (b >> 2) | (b >> 3);
    return ~t & CG_NIB_LO;
}

CG_HD int cg_popc64(uint64_t x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __popcll(x);
#else
    return __builtin_popcountll(x);
#endif
}

CG_HD int cg_empties(uint64_t b) { return cg_popc64(cg_zmask(b)); }

/* The D4 permutations work on the two 32-bit halves (rows 0-1 and rows
 * 2-3), where they are byte permutes, 16-bit rotates and masked nibble
 * shifts -- a few single-cycle ALU operations each on gfx950. */
CG_HD uint64_t cg_join(uint32_t hi, uint32_t lo) { return ((uint64_t)hi << 32) | lo; }

/* rotate_clockwise (board.c:239-248): out(r,c) = in(3-c, r).  Step 1 moves
 * bytes between the halves; step 2 moves nibbles inside each half. */
CG_HD uint32_t cg_rot_half(uint32_t x)
{
    uint32_t r16 = (x >> 16) | (x << 16);
    return (r16 & 0xf0f00f0fu) | ((x >> 4) & 0x0f0f0000u) | ((x << 4) & 0x0000f0f0u);
}

CG_HD uint64_t cg_rot_cw(uint64_t b)
{
    uint32_t hi = (uint32_t)(b >> 32), lo = (uint32_t)b;
    uint32_t h1 = ((hi >> 8) & 0x00ff00ffu) | (lo & 0xff00ff00u);
    uint32_t l1 = (hi & 0x00ff00ffu) | ((lo << 8) & 0xff00ff00u);
    return cg_join(cg_rot_half(h1), cg_rot_half(l1));
}

/* row-order reversal (flip, board.c:376-381): swap the halves and the rows
 * inside each */
CG_HD uint64_t cg_flip_rows(uint64_t b)
{
    uint32_t hi = (uint32_t)(b >> 32), lo = (uint32_t)b;
    return cg_join((lo >> 16) | (lo << 16), (hi >> 16) | (hi << 16));
}

/* nibble reversal inside every 16-bit row (mirror left/right) */
CG_HD uint32_t cg_mirror_half(uint32_t x)
{
    x = ((x >> 8) & 0x00ff00ffu) | ((x << 8) & 0xff00ff00u);
    return ((x >> 4) & 0x0f0f0f0fu) | ((x << 4) & 0xf0f0f0f0u);
}

CG_HD uint64_t cg_mirror_rows(uint64_t b)
{
    return cg_join(cg_mirror_half((uint32_t)(b >> 32)), cg_mirror_half((uint32_t)b));
}

CG_HD uint64_t cg_umax(uint64_t a, uint64_t b) { return a > b ? a : b; }

/* Canonical form: unsigned max over the 8 D4 images (board.c:314-337).
 * With y = cw(b): images are b, y, rot180 b, rot180 y and their row flips;
 * rot180 = flip o mirror. */
CG_HD uint64_t cg_canon(uint64_t b)
{
    uint64_t y = cg_rot_cw(b);
    uint64_t mb = cg_mirror_rows(b), my = cg_mirror_rows(y);
    uint64_t m = cg_umax(b, y);
    m = cg_umax(m, cg_umax(mb, my));               /* f a^2 b, f a^3 b */
    m = cg_umax(m, cg_umax(cg_flip_rows(b), cg_flip_rows(y)));
    m = cg_umax(m, cg_umax(cg_flip_rows(mb), cg_flip_rows(my)));  /* a^2, a^3 */
    return m;
}

/* Image i of b in cg_canon's order: bit 0 turns it clockwise, then bit 1
 * mirrors the rows, then bit 2 reverses the row order (0 b, 1 cw b, 2
 * mirror b, 3 mirror cw b, 4 flip b, 5 flip cw b, 6 flip mirror b, 7 flip
 * mirror cw b). */
CG_HD uint64_t cg_image(uint64_t b, int i)
{
    uint64_t x = (i & 1) ? cg_rot_cw(b) : b;
    if (i & 2) x = cg_mirror_rows(x);
    if (i & 4) x = cg_flip_rows(x);
    return x;
}

/* The max over the images whose bit is set in sel (0xff: cg_canon). */
CG_HD uint64_t cg_canon_sel(uint64_t b, unsigned sel)
{
    if (sel == 0xffu) return cg_canon(b);
    const uint64_t y = (sel & 0xaau) ? cg_rot_cw(b) : 0;
    const uint64_t mb = (sel & 0x44u) ? cg_mirror_rows(b) : 0;
    const uint64_t my = (sel & 0x88u) ? cg_mirror_rows(y) : 0;
    uint64_t m = 0;
    if (sel & 0x01u) m = cg_umax(m, b);
    if (sel & 0x02u) m = cg_umax(m, y);
    if (sel & 0x04u) m = cg_umax(m, mb);
    if (sel & 0x08u) m = cg_umax(m, my);
    if (sel & 0x10u) m = cg_umax(m, cg_flip_rows(b));
    if (sel & 0x20u) m = cg_umax(m, cg_flip_rows(y));
    if (sel & 0x40u) m = cg_umax(m, cg_flip_rows(mb));
    if (sel & 0x80u) m = cg_umax(m, cg_flip_rows(my));
    return m;
}

/* Canonical form through a formation's stabiliser.  Let F be the 0xF cells
 * of a board whose other cells are all below 0xF, and A the canonical place
 * of F (the largest of its images, 0xF being the largest nibble).  If every
 * image P != A of F, read from the top nibble, first differs from A at a
 * cell where A is 0xF after cells that are all 0xF in A, then an image of
 * the board with its formation at P is below one with its formation at A:
 * equal nibbles up to that cell, where one holds 0xF and the other a free
 * cell.  The canonical form is then the max over the images that map F to
 * A only -- one image for DPDGap, two for the 2x4 formations.
 *   f[o]:   the fmask of the boards of orientation o (0: as the layer holds
 *           them; 1: turned clockwise, as an up or down move leaves them)
 *   sel[o]: the images of those boards that put the formation at A
 * A board whose fmask is not f[o] takes the full max of 8 images, so the
 * result is cg_canon's for every board. */
typedef struct {
    uint64_t f[2];
    uint32_t sel[2];
} cg_canon_sel_t;

CG_HD uint64_t cg_canon_fast(uint64_t b, cg_canon_sel_t cs, int o)
{
    return cg_fmask(b) == cs.f[o] ? cg_canon_sel(b, cs.sel[o]) : cg_canon(b);
}

/* every board takes the full max of 8 images */
CG_HD cg_canon_sel_t cg_canon_all(void)
{
    cg_canon_sel_t cs;
    cs.f[0] = cs.f[1] = ~0ULL;
    cs.sel[0] = cs.sel[1] = 0xffu;
    return cs;
}

/* the selection for boards whose 0xF cells are those of e, in e's
 * orientation; with free_formation (formation cells move) or a formation
 * whose images are not decided by it, f never matches and every board takes
 * the full max */
CG_HD cg_canon_sel_t cg_make_canon_sel(uint64_t e, int ff)
{
    cg_canon_sel_t cs = cg_canon_all();
    if (ff) return cs;
    const uint64_t fe = cg_fmask(e) * 0xfULL;   /* e's formation cells as 0xF, the rest 0 */
    const uint64_t A = cg_canon(fe);
    for (int i = 0; i < 8; i++) {
        const uint64_t P = cg_image(A, i);
        for (int x = 0; x < 16 && P != A; x++) {
            const unsigned a = (unsigned)(A >> (60 - 4 * x)) & 0xfu, p = (unsigned)(P >> (60 - 4 * x)) & 0xfu;
            if (a != 0xfu) return cs;   /* a free cell before the first difference */
            if (p != 0xfu) break;      /* decided: A holds 0xF, P a free cell */
        }
    }
    const uint64_t fr = cg_rot_cw(fe);
    uint32_t s0 = 0, s1 = 0;
    for (int i = 0; i < 8; i++) {
        if (cg_image(fe, i) == A) s0 |= 1u << i;
        if (cg_image(fr, i) == A) s1 |= 1u << i;
    }
    cs.f[0] = cg_fmask(fe);
    cs.f[1] = cg_fmask(fr);
    cs.sel[0] = s0;
    cs.sel[1] = s1;
    return cs;
}

/* One horizontal board move using a left-move row table `L` (65536 u16).
 * dir 0 = left, 1 = right.  Returns 1 iff valid (movedir_hori
 * board.c:151-169): some row changed and every row is locked -- the lock rule
 * (board.c:29-38, 117) allows a row move unless a 0xF cell was vacated, or
 * always with free_formation. */
CG_HD int cg_move_hori_tab(const uint16_t *L, uint64_t *b, int dir, int ff)
{
    /* a right move is a left move of the mirrored board, mirrored back; the
     * changed and lock tests of the four rows are tests on the whole board */
    const uint64_t pre = *b;
    const uint64_t src = dir ? cg_mirror_rows(pre) : pre;
    uint64_t out = 0;
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
    for (int i = 0; i < 4; i++) out |= (uint64_t)L[(src >> (16 * i)) & 0xffffu] << (16 * i);
    if (dir) out = cg_mirror_rows(out);
    if (out != pre && (ff || (cg_fmask(pre) & ~cg_fmask(out)) == 0)) { *b = out; return 1; }
    return 0;
}

/* movedir_unstable (board.c:195-214): up = cw then right, down = cw then
 * left; the moved board stays rotated.  On failure *b is unspecified. */
CG_HD int cg_move_tab(const uint16_t *L, uint64_t *b, int dir, int ff)
{
    if (dir >= 2) *b = cg_rot_cw(*b);
    return cg_move_hori_tab(L, b, (dir == 2) ? 1 : (dir == 3 ? 0 : dir), ff);
}

/* prune_board (generate.c:98-131) as a multiset test: a value v with
 * smallest_large <= v < 0xE is large; v == smallest_large may occur 5 times,
 * other large values twice; the small-tile sum counts every v <
 * smallest_large (an empty cell adds 1); prune if it exceeds stsl + 64 or if
 * there are more than ltc large tiles. */
CG_HD int cg_prune(uint64_t b, long stsl, long ltc, long sl)
{
    /* H: a 4-bit count of every value v at bits [4v, 4v+3].  A count reaches
     * 16 (and carries into the next counter) only on a board whose 16 cells
     * are equal, which is handled at the end. */
    uint64_t H = 0;
    /* 4 * value of the even / odd cells, one per byte (the shift amounts) */
    const uint64_t ev = (b & 0x0F0F0F0F0F0F0F0FULL) << 2, od = (b >> 2) & 0x3C3C3C3C3C3C3C3CULL;
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
    for (int k = 0; k < 8; k++) {
        H += 1ULL << ((ev >> (8 * k)) & 63);
        H += 1ULL << ((od >> (8 * k)) & 63);
    }
    /* value classes: small v < slc, large slc <= v < 0xE; slc = sl clamped to [0, 14] */
    const long slc = sl < 0 ? 0 : (sl > 14 ? 14 : sl);
    const uint64_t small_m = (1ULL << (4 * slc)) - 1ULL;
    const uint64_t large_m = ((1ULL << 56) - 1ULL) & ~small_m;
    const uint64_t sl_bit = (sl >= 0 && sl < 14) ? 1ULL << (4 * sl) : 0;
    /* bit 0 of each counter: count >= 3, count >= 6 */
    const uint64_t ge3 = ((H >> 3) | (H >> 2) | ((H >> 1) & H)) & CG_NIB_LO;
    const uint64_t ge6 = ((H >> 3) | ((H >> 2) & (H >> 1))) & CG_NIB_LO;
    int bad = ((ge3 & large_m & ~sl_bit) | (ge6 & sl_bit)) != 0;
    uint64_t t = H & large_m;
    t = (t & 0x0F0F0F0F0F0F0F0FULL) + ((t >> 4) & 0x0F0F0F0F0F0F0F0FULL);
    uint32_t t32 = (uint32_t)t + (uint32_t)(t >> 32);
    t32 += t32 >> 16;
    int large = (int)((t32 + (t32 >> 8)) & 0xff);
    /* small-tile sum: sum over v < slc of count(v) * 2^v, folding the counter
     * digits pairwise (fields of 8, 16, 32 bits, weights 2, 4, 16, 256) */
    uint64_t hs = H & small_m;
    hs = (hs & 0x0F0F0F0F0F0F0F0FULL) + 2 * ((hs >> 4) & 0x0F0F0F0F0F0F0F0FULL);
    hs = (hs & 0x00FF00FF00FF00FFULL) + 4 * ((hs >> 8) & 0x00FF00FF00FF00FFULL);
    hs = (hs & 0x0000FFFF0000FFFFULL) + 16 * ((hs >> 16) & 0x0000FFFF0000FFFFULL);
    int sts = (int)((uint32_t)hs + ((uint32_t)(hs >> 32) << 8));
    /* a board of 16 equal cells u */
    const uint64_t u = b & 0xf;
    if (b == u * CG_NIB_LO) {
        const int is_large = (long)u >= sl && u < 0xe;
        large = is_large ? 16 : 0;
        bad = is_large;
        sts = (!is_large && u < 0xe) ? 16 << u : 0;
    }
    return bad || sts > stsl + 64 || large > ltc;
}

/* cmpbrd (solve.c:226-231) with the mask precomputed: (b & mask) == w */
CG_HD uint64_t cg_win_mask(uint64_t w)
{
    uint64_t m = (w >> 2) | w;
    m = ((m >> 1) | m) & CG_NIB_LO;
    return m * 0xf;
}

/* board holds the nibble v (nonzero v) */
CG_HD int cg_has_nibble(uint64_t b, unsigned v)
{
    return cg_fmask(~(b ^ (CG_NIB_LO * (uint64_t)(v & 0xf)))) != 0;
}

/* tile sum (board.c:340-347): sum of 2^v over v not in {0, 0xE, 0xF} */
CG_HD int cg_tile_sum(uint64_t b)
{
    int s = 0;
    for (int x = 0; x < 16; x++) {
        int v = (int)(b >> (60 - 4 * x)) & 0xf;
        if (v != 0 && v < 0xe) s += 1 << v;
    }
    return s;
}

CG_HD uint64_t cg_set_tile(uint64_t b, int x, unsigned v)
{
    int sh = 60 - 4 * x;
    return (b & ~(0xfULL << sh)) | ((uint64_t)v << sh);
}

#endif
