/*
 * Host check of the product's bit-parallel board primitives (board.h, the
 * code the kernels run) against the oracle: prune_board (generate.c:98-131)
 * over its parameter range, the four moves with the lock rule
 * (board.c:151-214), and the canonical form (board.c:314-337).  Built and
 * run by tests/test_board_host.py; exits non-zero on the first mismatch.
 */
#include <stdio.h>
#include <stdlib.h>
#include "board.h"
#include "oracle.h"

static uint64_t g_s = 88172645463325252ULL;
static uint64_t rnd(void)
{
    g_s ^= g_s << 13;
    g_s ^= g_s >> 7;
    g_s ^= g_s << 17;
    return g_s;
}

/* boards of several shapes: uniform nibbles, sparse small tiles, clustered
 * values, formations (0xF), and boards of 16 (or 15) equal cells */
static uint64_t board(long it)
{
    uint64_t b = 0;
    int mode = (int)(it % 4), kmax = 1 + (int)(rnd() % 15);
    for (int x = 0; x < 16; x++) {
        uint64_t r = rnd(), v;
        if (mode == 0) v = r & 0xf;
        else if (mode == 1) v = (r % 3 == 0) ? 0 : r % (uint64_t)(kmax + 1);
        else if (mode == 2) v = (r >> 8) % 4 + (uint64_t)(it % 11);
        else v = (r % 5 == 0) ? 0xf : r % 8;
        if (v > 15) v = 15;
        b |= v << (4 * x);
    }
    if (it % 97 == 0) b = (rnd() & 0xf) * CG_NIB_LO;
    if (it % 89 == 0) b = ((rnd() & 0xf) * CG_NIB_LO) ^ ((rnd() & 0xf) << (4 * (rnd() % 16)));
    return b;
}

int main(int argc, char **argv)
{
    long n = argc > 1 ? atol(argv[1]) : 1000000;
    static const long sls[] = { -1, 0, 1, 2, 5, 6, 7, 9, 13, 14, 15, 20 };
    static const long stsls[] = { -100, -64, 0, 10, 58, 200, 100000 };
    static const long ltcs[] = { -1, 0, 1, 3, 5, 8, 16, 17 };
    static uint16_t mlut[2][65536];
    static uint8_t llut[2][65536];
    long pruned = 0, moved = 0;
    for (int ff = 0; ff < 2; ff++)
        for (int ig = 0; ig < 2; ig++) {
            orc_cfg cfg = { 0 };
            cfg.free_formation = ff;
            cfg.ignore_f = ig;
            cfg.cores = 1;
            orc_build_lut(ff, ig, mlut, llut);
            for (long it = 0; it < n / 4; it++) {
                const uint64_t b = board(it);
                for (int d = 0; d < 4; d++) {
                    uint64_t p = b, q = b;
                    const int a = cg_move_tab(mlut[0], &p, d, ff);
                    const int c = orc_move(&cfg, &q, d);
                    if (a != c || (a && p != q)) {
                        printf("move mismatch b=%016llx d=%d ff=%d ig=%d\n", (unsigned long long)b, d, ff,
                               ig);
                        return 1;
                    }
                    moved += a;
                }
            }
        }
    for (long it = 0; it < n; it++) {
        const uint64_t b = board(it);
        const long sl = sls[rnd() % 12], stsl = stsls[rnd() % 7], ltc = ltcs[rnd() % 8];
        const int a = orc_prune(b, stsl, ltc, sl), c = cg_prune(b, stsl, ltc, sl);
        if (a != c) {
            printf("prune mismatch b=%016llx sl=%ld stsl=%ld ltc=%ld\n", (unsigned long long)b, sl, stsl,
                   ltc);
            return 1;
        }
        pruned += a;
        if (cg_canon(b) != orc_canon(b)) {
            printf("canon mismatch b=%016llx\n", (unsigned long long)b);
            return 1;
        }
    }
    for (int u = 0; u < 16; u++)
        for (int i = 0; i < 12; i++)
            for (int j = 0; j < 7; j++)
                for (int k = 0; k < 8; k++) {
                    const uint64_t b = (uint64_t)u * CG_NIB_LO;
                    if (orc_prune(b, stsls[j], ltcs[k], sls[i]) != cg_prune(b, stsls[j], ltcs[k], sls[i])) {
                        printf("prune mismatch on a uniform board u=%d\n", u);
                        return 1;
                    }
                }
    printf("ok %ld boards, %ld pruned, %ld legal moves\n", n, pruned, moved);
    return 0;
}
