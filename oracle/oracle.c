/*
 * oracle.c -- CPU restatement of cablegen's layered generate/solve path.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker and the timed CPU baseline
 * ("port").  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it.  The product library never links or calls it.
 *
 * It follows the reference's algorithm class deliberately (so that its timing
 * is a fair stand-in for the reference CPU path): row-LUT moves, canonical
 * form as the max over the 8 D4 images computed by rotate/flip, pthreads with
 * a static contiguous partition of each layer, a per-thread comparison sort,
 * a serial k-way merge-dedupe, and binary-search lookups with the reference's
 * linear tail.  Compiled like the reference's bench build: -O2, no
 * -march=native, no -ffast-math.
 *
 * Deviation (documented): prune uses smallest_large from the configuration,
 * the intended semantics of generate.c:154-175; the reference's CPU path
 * reads an uninitialised field there (generate.c:31, :215-230).
 */
#define _GNU_SOURCE
#include "oracle.h"
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#include <math.h>

#define TILE(b, x)   ((unsigned)(((b) >> (60 - 4 * (x))) & 0xfu))
#define SETT(b, x, v) ((b) = ((b) & ~(0xfULL << (60 - 4 * (x)))) | ((uint64_t)(v) << (60 - 4 * (x))))

/* ------------------------------------------------------------------------ */
/* Row rules and LUTs: board.c:29-132                                        */
/* ------------------------------------------------------------------------ */

/* Slide-merge-slide of the cells p[0..n) toward p[0] (board.c:62-98).  With
 * ignore_f the trailing run of 0xf cells is excluded (board.c:65-74). */
static void row_move_cells(uint8_t *t, int n, int ignore_f)
{
    if (ignore_f)
        while (n > 0 && t[n - 1] == 0xf) n--;
    uint8_t s[4];
    int k = 0;
    for (int i = 0; i < n; i++)
        if (t[i]) s[k++] = t[i];
    uint8_t m[4];
    int j = 0;
    for (int i = 0; i < k;) {
        /* merge rule board.c:81-90: equal, nonzero, and below 0xE */
        if (i + 1 < k && s[i] == s[i + 1] && s[i] < 0xe) {
            m[j++] = (uint8_t)(s[i] + 1);
            i += 2;
        } else {
            m[j++] = s[i];
            i += 1;
        }
    }
    for (int i = 0; i < n; i++) t[i] = i < j ? m[i] : 0;
}

/* generate_lut: board.c:100-132.  Left moves toward column 0 (cell order
 * 0,1,2,3); right moves use the cell order 3,2,1,0.  lock = no 0xf cell was
 * vacated (board.c:29-38), or free_formation. */
void orc_build_lut(int free_formation, int ignore_f,
                   uint16_t move_lut[2][65536], uint8_t lock_lut[2][65536])
{
    for (unsigned r = 0; r < 65536; r++) {
        uint8_t c[4] = { (uint8_t)(r >> 12 & 0xf), (uint8_t)(r >> 8 & 0xf),
                         (uint8_t)(r >> 4 & 0xf), (uint8_t)(r & 0xf) };
        for (int d = 0; d < 2; d++) {
            uint8_t t[4];
            for (int i = 0; i < 4; i++) t[i] = d == 0 ? c[i] : c[3 - i];
            row_move_cells(t, 4, ignore_f);
            uint8_t o[4];
            for (int i = 0; i < 4; i++) o[d == 0 ? i : 3 - i] = t[i];
            move_lut[d][r] = (uint16_t)(o[0] << 12 | o[1] << 8 | o[2] << 4 | o[3]);
            int vacated = 0;
            for (int i = 0; i < 4; i++)
                if (c[i] == 0xf && o[i] != 0xf) vacated = 1;
            lock_lut[d][r] = (uint8_t)(free_formation || !vacated);
        }
    }
    /* explicit entries, board.c:126-129 */
    move_lut[0][0x0000] = move_lut[1][0x0000] = 0x0000;
    move_lut[0][0xffff] = move_lut[1][0xffff] = 0xffff;
}

typedef struct {
    uint16_t move[2][65536];
    uint8_t lock[2][65536];
} lut_t;

static lut_t *g_luts[4];
static pthread_mutex_t g_lut_mu = PTHREAD_MUTEX_INITIALIZER;

static const lut_t *get_lut(const orc_cfg *cfg)
{
    int idx = (cfg->free_formation ? 1 : 0) | (cfg->ignore_f ? 2 : 0);
    pthread_mutex_lock(&g_lut_mu);
    if (!g_luts[idx]) {
        lut_t *l = malloc(sizeof(lut_t));
        orc_build_lut(cfg->free_formation, cfg->ignore_f, l->move, l->lock);
        g_luts[idx] = l;
    }
    pthread_mutex_unlock(&g_lut_mu);
    return g_luts[idx];
}

/* ------------------------------------------------------------------------ */
/* D4 symmetries: board.c:239-268, 270-338, 349-393                          */
/* ------------------------------------------------------------------------ */

/* rotate_clockwise (board.c:239-248): out(r,c) = in(3-c, r). */
uint64_t orc_rotate_cw(uint64_t b)
{
    uint64_t o = 0;
    for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++)
            SETT(o, 4 * r + c, TILE(b, 4 * (3 - c) + r));
    return o;
}

/* flip (board.c:376-381): row-order reversal; the per-tile loop after the
 * early return is dead code. */
uint64_t orc_flip(uint64_t b)
{
    uint64_t o = 0;
    for (int r = 0; r < 4; r++)
        o |= ((b >> (16 * r)) & 0xffffULL) << (16 * (3 - r));
    return o;
}

/* canonicalize_b scalar path (board.c:314-337): unsigned max over
 * {e, a, a^2, a^3, f, f a, f a^2, f a^3}. */
uint64_t orc_canon(uint64_t b)
{
    uint64_t best = b, x = b;
    uint64_t rots[3];
    for (int i = 0; i < 3; i++) {
        x = orc_rotate_cw(x);
        rots[i] = x;
        if (x > best) best = x;
    }
    uint64_t f = orc_flip(b);
    if (f > best) best = f;
    for (int i = 0; i < 3; i++) {
        f = orc_flip(rots[i]);
        if (f > best) best = f;
    }
    return best;
}

/* get_all_rots (board.c:349-374): {e, a, a^2, a^3, b, b·a, b·a^2, b·a^3} where
 * "b·a" means flip first, then rotate (board.c:361-373). */
void orc_all_rots(uint64_t b, uint64_t out[8])
{
    uint64_t x = b;
    out[0] = b;
    for (int i = 1; i < 4; i++) { x = orc_rotate_cw(x); out[i] = x; }
    x = orc_flip(b);
    out[4] = x;
    for (int i = 5; i < 8; i++) { x = orc_rotate_cw(x); out[i] = x; }
}

/* get_sum (board.c:340-347) */
int orc_get_sum(uint64_t b)
{
    int s = 0;
    for (int i = 0; i < 16; i++) {
        unsigned v = TILE(b, i);
        if (v != 0 && v != 0xe && v != 0xf) s += 1 << v;
    }
    return s;
}

/* movedir_hori (board.c:151-169): per row, apply the LUT; invalid if any row
 * is not locked; valid iff some row changed. */
static int move_hori(const lut_t *L, uint64_t *b, int d)
{
    uint64_t pre = *b, out = *b;
    int changed = 0;
    for (int i = 0; i < 4; i++) {
        unsigned row = (unsigned)((pre >> (16 * i)) & 0xffff);
        unsigned nr = L->move[d][row];
        if (nr != row) {
            changed = 1;
            out = (out & ~(0xffffULL << (16 * i))) | ((uint64_t)nr << (16 * i));
        }
        if (!L->lock[d][row]) {
            *b = pre;
            return 0;
        }
    }
    *b = out;
    return changed;
}

/* movedir_unstable (board.c:195-214): up = rotate_cw then right, down =
 * rotate_cw then left; the result stays rotated. */
static int move_unstable(const lut_t *L, uint64_t *b, int dir)
{
    switch (dir) {
    case 0: return move_hori(L, b, 0);
    case 1: return move_hori(L, b, 1);
    case 2: *b = orc_rotate_cw(*b); return move_hori(L, b, 1);
    default: *b = orc_rotate_cw(*b); return move_hori(L, b, 0);
    }
}

int orc_move(const orc_cfg *cfg, uint64_t *b, int dir)
{
    return move_unstable(get_lut(cfg), b, dir);
}

/* prune_board (generate.c:98-131).  Large tiles are smallest_large <= v < 0xE;
 * a value equal to smallest_large may appear 5 times (three counted by c64,
 * then the two bitsets), any other large value twice; the small-tile sum
 * counts every v < smallest_large (an empty cell adds 1 << 0); prune if that
 * sum exceeds stsl + 64 or there are more than ltc large tiles. */
int orc_prune(uint64_t b, long stsl, long ltc, long smallest_large)
{
    int large = 0, c_sl = 0, sts = 0;
    unsigned seen1 = 0, seen2 = 0;
    for (int i = 0; i < 16; i++) {
        long v = TILE(b, i);
        if (v >= smallest_large && v < 0xe) {
            large++;
            if (v == smallest_large && c_sl < 3) {
                c_sl++;
            } else if (!(seen1 >> v & 1u)) {
                seen1 |= 1u << v;
            } else if (!(seen2 >> v & 1u)) {
                seen2 |= 1u << v;
            } else {
                return 1;
            }
        } else if (v < 0xe) {
            sts += 1 << v;
        }
    }
    if (sts > stsl + 64) return 1;
    if (large > ltc) return 1;
    return 0;
}

/* cmpbrd (solve.c:226-231): every nonzero nibble of w must match b. */
int orc_cmpbrd(uint64_t b, uint64_t w)
{
    uint64_t m = 0;
    for (int i = 0; i < 16; i++)
        if (TILE(w, i)) m |= 0xfULL << (60 - 4 * i);
    return (b & m) == w;
}

/* lookup (solve.c:108-164): bisection over [bottom, top) until the range is
 * narrower than 50, then a linear scan; a miss or an empty table gives 0. */
double orc_lookup(uint64_t key, const uint64_t *keys, const double *vals, size_t n,
                  int canonicalize)
{
    if (n == 0) return 0.0;
    if (canonicalize) key = orc_canon(key);
    size_t top = n, bottom = 0, mid = n / 2;
    while (keys[mid] != key) {
        if (top - bottom < 50) {
            for (size_t i = bottom; i < top; i++)
                if (keys[i] == key) return vals[i];
            return 0.0;
        }
        if (keys[mid] < key) bottom = mid;
        else top = mid;
        mid = (top + bottom) / 2;
    }
    return vals[mid];
}

/* exact index of key in a sorted unique array, or -1 */
static long find_index(uint64_t key, const uint64_t *keys, size_t n)
{
    size_t lo = 0, hi = n;
    while (lo < hi) {
        size_t mid = lo + (hi - lo) / 2;
        if (keys[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    return (lo < n && keys[lo] == key) ? (long)lo : -1;
}

/* ------------------------------------------------------------------------ */
/* Sort and merge: array.c:246-291, fluxsort.h:112-121 (ascending unsigned)  */
/* ------------------------------------------------------------------------ */

static void insertion_sort(uint64_t *a, size_t n)
{
    for (size_t i = 1; i < n; i++) {
        uint64_t v = a[i];
        size_t j = i;
        while (j > 0 && a[j - 1] > v) { a[j] = a[j - 1]; j--; }
        a[j] = v;
    }
}

static void sift_down(uint64_t *a, size_t n, size_t i)
{
    for (;;) {
        size_t c = 2 * i + 1;
        if (c >= n) return;
        if (c + 1 < n && a[c + 1] > a[c]) c++;
        if (a[i] >= a[c]) return;
        uint64_t t = a[i]; a[i] = a[c]; a[c] = t;
        i = c;
    }
}

static void heap_sort(uint64_t *a, size_t n)
{
    for (size_t i = n / 2; i-- > 0;) sift_down(a, n, i);
    for (size_t i = n; i-- > 1;) {
        uint64_t t = a[0]; a[0] = a[i]; a[i] = t;
        sift_down(a, i, 0);
    }
}

static void intro_sort(uint64_t *a, size_t n, int depth)
{
    while (n > 24) {
        if (depth-- == 0) { heap_sort(a, n); return; }
        uint64_t x = a[0], y = a[n / 2], z = a[n - 1];
        uint64_t p = x < y ? (y < z ? y : (x < z ? z : x)) : (x < z ? x : (y < z ? z : y));
        size_t i = 0, j = n - 1;
        for (;;) {
            while (a[i] < p) i++;
            while (a[j] > p) j--;
            if (i >= j) break;
            uint64_t t = a[i]; a[i] = a[j]; a[j] = t;
            i++; j--;
        }
        /* [0, j] <= p <= [j+1, n) */
        if (j + 1 < n - j - 1) { intro_sort(a, j + 1, depth); a += j + 1; n -= j + 1; }
        else { intro_sort(a + j + 1, n - j - 1, depth); n = j + 1; }
    }
    insertion_sort(a, n);
}

void orc_sort_u64(uint64_t *a, size_t n)
{
    int depth = 0;
    for (size_t m = n; m > 1; m >>= 1) depth += 2;
    intro_sort(a, n, depth);
}

/* deduplicate_threads (array.c:246-287): serial k-way merge of sorted arrays,
 * taking the minimum head each step and dropping repeats of the last output. */
uint64_t *orc_kway_dedupe(uint64_t *const *arrs, const size_t *sizes, size_t k, size_t *out_n)
{
    size_t total = 0;
    for (size_t i = 0; i < k; i++) total += sizes[i];
    uint64_t *res = malloc((total ? total : 1) * sizeof(uint64_t));
    size_t *off = calloc(k ? k : 1, sizeof(size_t));
    size_t n = 0;
    for (;;) {
        size_t best = k;
        uint64_t mn = 0;
        for (size_t i = 0; i < k; i++) {
            if (off[i] >= sizes[i]) continue;
            uint64_t v = arrs[i][off[i]];
            if (best == k || v < mn) { mn = v; best = i; }
        }
        if (best == k) break;
        off[best]++;
        if (n && res[n - 1] == mn) continue;
        res[n++] = mn;
    }
    free(off);
    *out_n = n;
    return res;
}

/* ------------------------------------------------------------------------ */
/* Generation: generate.c:133-199 (thread bodies), 201-301 (fan-out),        */
/* 409-550 (layer), 551-592 (recurrence)                                     */
/* ------------------------------------------------------------------------ */

typedef struct { uint64_t *p; size_t n, cap; } darr;

static void darr_init(darr *d, size_t cap)
{
    d->cap = cap ? cap : 16;
    d->n = 0;
    d->p = malloc(d->cap * sizeof(uint64_t));
}

static inline void darr_push(darr *d, uint64_t v)
{
    if (d->n == d->cap) {               /* x3 growth, array.c:17-50 */
        d->cap *= 3;
        d->p = realloc(d->p, d->cap * sizeof(uint64_t));
    }
    d->p[d->n++] = v;
}

enum { OP_MOVE, OP_SPAWN };

typedef struct {
    const orc_cfg *cfg;
    const lut_t *lut;
    const uint64_t *in;
    size_t start, end;
    int op;
    darr out, out2, out4;
    pthread_t th;
} gen_task;

static void *gen_worker(void *arg)
{
    gen_task *t = arg;
    const orc_cfg *cfg = t->cfg;
    if (t->op == OP_MOVE) {
        for (size_t i = t->start; i < t->end; i++) {
            uint64_t old = t->in[i];
            if (cfg->prune && orc_prune(old, cfg->stsl, cfg->ltc, cfg->smallest_large))
                continue;
            for (int d = 0; d < 4; d++) {
                uint64_t b = old;
                if (!move_unstable(t->lut, &b, d)) continue;
                if (cfg->prune && orc_prune(b, cfg->stsl, cfg->ltc, cfg->smallest_large))
                    continue;
                darr_push(&t->out, orc_canon(b));
            }
        }
        orc_sort_u64(t->out.p, t->out.n);
    } else {
        for (size_t i = t->start; i < t->end; i++) {
            uint64_t old = t->in[i];
            for (int x = 0; x < 16; x++) {
                if (TILE(old, x)) continue;
                uint64_t b = old;
                SETT(b, x, 1);
                darr_push(&t->out2, orc_canon(b));
                b = old;
                SETT(b, x, 2);
                darr_push(&t->out4, orc_canon(b));
            }
        }
        orc_sort_u64(t->out2.p, t->out2.n);
        orc_sort_u64(t->out4.p, t->out4.n);
    }
    return NULL;
}

static void run_tasks(const orc_cfg *cfg, const uint64_t *in, size_t n, int op, gen_task *tasks)
{
    int cores = cfg->cores > 0 ? cfg->cores : 1;
    size_t block = n / (size_t)cores;   /* static partition, generate.c:232-238 */
    const lut_t *L = get_lut(cfg);
    for (int i = 0; i < cores; i++) {
        gen_task *t = &tasks[i];
        t->cfg = cfg;
        t->lut = L;
        t->in = in;
        t->op = op;
        t->start = (size_t)i * block;
        t->end = (i + 1 == cores) ? n : (size_t)(i + 1) * block;
        if (op == OP_MOVE) {
            darr_init(&t->out, 3 * n / (size_t)cores);
        } else {
            darr_init(&t->out2, 4 * n / (size_t)cores);
            darr_init(&t->out4, 4 * n / (size_t)cores);
        }
    }
    if (cores == 1) {
        gen_worker(&tasks[0]);
        return;
    }
    for (int i = 0; i < cores; i++) pthread_create(&tasks[i].th, NULL, gen_worker, &tasks[i]);
    for (int i = 0; i < cores; i++) pthread_join(tasks[i].th, NULL);
}

int orc_gen_layer(const orc_cfg *cfg, const uint64_t *P, size_t nP,
                  const uint64_t *C, size_t nC,
                  uint64_t **M, size_t *nM, uint64_t **P2, size_t *nP2,
                  uint64_t **C2, size_t *nC2)
{
    int cores = cfg->cores > 0 ? cfg->cores : 1;
    gen_task *tasks = calloc((size_t)cores, sizeof(gen_task));
    uint64_t **arrs = malloc(sizeof(uint64_t *) * (size_t)(cores + 1));
    size_t *sz = malloc(sizeof(size_t) * (size_t)(cores + 1));

    /* move phase + merge (generate.c:471-516, replace_n :288-301) */
    run_tasks(cfg, P, nP, OP_MOVE, tasks);
    for (int i = 0; i < cores; i++) { arrs[i] = tasks[i].out.p; sz[i] = tasks[i].out.n; }
    uint64_t *m = orc_kway_dedupe(arrs, sz, (size_t)cores, nM);
    for (int i = 0; i < cores; i++) free(tasks[i].out.p);
    *M = m;

    /* spawn phase + merges with the carry (generate.c:518-549) */
    run_tasks(cfg, m, *nM, OP_SPAWN, tasks);
    for (int i = 0; i < cores; i++) { arrs[i] = tasks[i].out2.p; sz[i] = tasks[i].out2.n; }
    arrs[cores] = (uint64_t *)C; sz[cores] = nC;
    *P2 = orc_kway_dedupe(arrs, sz, (size_t)cores + 1, nP2);
    for (int i = 0; i < cores; i++) { arrs[i] = tasks[i].out4.p; sz[i] = tasks[i].out4.n; }
    *C2 = orc_kway_dedupe(arrs, sz, (size_t)cores, nC2);
    for (int i = 0; i < cores; i++) { free(tasks[i].out2.p); free(tasks[i].out4.p); }
    free(tasks);
    free(arrs);
    free(sz);
    return 0;
}

int orc_spawn_sets(const orc_cfg *cfg, const uint64_t *M, size_t nM,
                   uint64_t **S2, size_t *n2, uint64_t **S4, size_t *n4)
{
    int cores = cfg->cores > 0 ? cfg->cores : 1;
    gen_task *tasks = calloc((size_t)cores, sizeof(gen_task));
    uint64_t **arrs = malloc(sizeof(uint64_t *) * (size_t)cores);
    size_t *sz = malloc(sizeof(size_t) * (size_t)cores);
    run_tasks(cfg, M, nM, OP_SPAWN, tasks);
    for (int i = 0; i < cores; i++) { arrs[i] = tasks[i].out2.p; sz[i] = tasks[i].out2.n; }
    *S2 = orc_kway_dedupe(arrs, sz, (size_t)cores, n2);
    for (int i = 0; i < cores; i++) { arrs[i] = tasks[i].out4.p; sz[i] = tasks[i].out4.n; }
    *S4 = orc_kway_dedupe(arrs, sz, (size_t)cores, n4);
    for (int i = 0; i < cores; i++) { free(tasks[i].out2.p); free(tasks[i].out4.p); }
    free(tasks);
    free(arrs);
    free(sz);
    return 0;
}

int orc_generate(const orc_cfg *cfg, const uint64_t *initial, size_t n_init,
                 int start, int end, uint64_t ***layers, size_t **sizes, int *nlayers)
{
    int nl = end >= start ? (end - start) / 2 + 1 : 0;
    uint64_t **L = calloc((size_t)(nl ? nl : 1), sizeof(uint64_t *));
    size_t *S = calloc((size_t)(nl ? nl : 1), sizeof(size_t));
    uint64_t *P = malloc((n_init ? n_init : 1) * sizeof(uint64_t));
    memcpy(P, initial, n_init * sizeof(uint64_t));
    size_t nP = n_init;
    uint64_t *C = NULL;
    size_t nC = 0;
    int first = 0;
    if (cfg->premove && nl > 0) {
        /* layer 0 = the canonical set of the initial boards; the recurrence
         * continues from its spawns */
        for (size_t i = 0; i < nP; i++) P[i] = orc_canon(P[i]);
        size_t nM = 0;
        uint64_t *arr[1] = { P };
        size_t sz[1] = { nP };
        orc_sort_u64(P, nP);
        uint64_t *M = orc_kway_dedupe(arr, sz, 1, &nM);
        free(P);
        L[0] = M;
        S[0] = nM;
        orc_spawn_sets(cfg, M, nM, &P, &nP, &C, &nC);
        first = 1;
    }
    for (int i = first; i < nl; i++) {
        uint64_t *M, *P2, *C2;
        size_t nM, nP2, nC2;
        orc_gen_layer(cfg, P, nP, C, nC, &M, &nM, &P2, &nP2, &C2, &nC2);
        free(P);
        free(C);
        L[i] = M;
        S[i] = nM;
        P = P2; nP = nP2;
        C = C2; nC = nC2;
    }
    free(P);
    free(C);
    *layers = L;
    *sizes = S;
    *nlayers = nl;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Solve: solve.c:172-340                                                    */
/* ------------------------------------------------------------------------ */

size_t orc_expand_winstates(const uint64_t *ws, size_t n, uint64_t *out)
{
    for (size_t i = 0; i < n; i++) orc_all_rots(ws[i], out + 8 * i);
    return 8 * n;
}

/* satisfied (solve.c:233-244): nox (without score) wins when the board holds
 * the nibble nox (checkx, generate.c:89-96); score mode is never satisfied;
 * otherwise any winstate image matches under cmpbrd. */
static int satisfied(const orc_cfg *cfg, uint64_t b, const uint64_t *W, size_t nW)
{
    if (cfg->nox && !cfg->score) {
        for (int i = 0; i < 16; i++)
            if ((int)TILE(b, i) == (cfg->nox & 0xff)) return 1;
        return 0;
    }
    if (cfg->score) return 0;
    for (size_t i = 0; i < nW; i++)
        if (orc_cmpbrd(b, W[i])) return 1;
    return 0;
}

typedef struct {
    const orc_cfg *cfg;
    const lut_t *lut;
    const uint64_t *M;
    const uint64_t *K2; const double *V2; size_t n2;
    const uint64_t *K4; const double *V4; size_t n4;
    const uint64_t *W; size_t nW;
    double p4;
    double *V;
    size_t start, end;
    pthread_t th;
} solve_task;

/* maxmove (solve.c:246-256): max over the 4 moves of the looked-up value;
 * an invalid move or a miss contributes 0. */
static double maxmove(const lut_t *L, uint64_t b, const uint64_t *K, const double *V, size_t n)
{
    double p[4] = { 0, 0, 0, 0 };
    for (int d = 0; d < 4; d++) {
        uint64_t t = b;
        if (move_unstable(L, &t, d)) p[d] = orc_lookup(t, K, V, n, 1);
    }
    return fmax(fmax(p[0], p[1]), fmax(p[2], p[3]));
}

/* expectimax (solve.c:257-281): cells in ascending order; the spawn average
 * is (w2*A2/E) + (w4*A4/E) with w2 = 1 - p4 (0.9 for p4 = 0.1, exactly). */
static double expectimax(const solve_task *t, uint64_t b)
{
    int E = 0;
    double a2 = 0, a4 = 0;
    for (int x = 0; x < 16; x++) {
        if (TILE(b, x)) continue;
        E++;
        uint64_t c = b;
        SETT(c, x, 1);
        a2 += maxmove(t->lut, c, t->K2, t->V2, t->n2);
        SETT(c, x, 2);
        a4 += maxmove(t->lut, c, t->K4, t->V4, t->n4);
    }
    double res = 0;
    if (E) res = ((1.0 - t->p4) * a2 / E) + (t->p4 * a4 / E);
    if (t->cfg->score && res == 0) res = orc_get_sum(b);
    return res;
}

static void *solve_worker(void *arg)
{
    solve_task *t = arg;
    for (size_t i = t->start; i < t->end; i++) {
        uint64_t b = t->M[i];
        t->V[i] = satisfied(t->cfg, b, t->W, t->nW) ? 1.0 : expectimax(t, b);
    }
    return NULL;
}

int orc_solve_layer(const orc_cfg *cfg, const uint64_t *M, size_t nM,
                    const uint64_t *K2, const double *V2, size_t n2,
                    const uint64_t *K4, const double *V4, size_t n4,
                    const uint64_t *W, size_t nW, double p4, double *V)
{
    int cores = cfg->cores > 0 ? cfg->cores : 1;
    solve_task *tasks = calloc((size_t)cores, sizeof(solve_task));
    size_t block = nM / (size_t)cores;   /* solve.c:322-328 */
    const lut_t *L = get_lut(cfg);
    for (int i = 0; i < cores; i++) {
        solve_task *t = &tasks[i];
        *t = (solve_task){ cfg, L, M, K2, V2, n2, K4, V4, n4, W, nW, p4, V, 0, 0, 0 };
        t->start = (size_t)i * block;
        t->end = (i + 1 == cores) ? nM : (size_t)(i + 1) * block;
    }
    if (cores == 1) {
        solve_worker(&tasks[0]);
    } else {
        for (int i = 0; i < cores; i++) pthread_create(&tasks[i].th, NULL, solve_worker, &tasks[i]);
        for (int i = 0; i < cores; i++) pthread_join(tasks[i].th, NULL);
    }
    free(tasks);
    return 0;
}

int orc_solve(const orc_cfg *cfg, uint64_t **layers, const size_t *sizes, int nlayers,
              int start_gen, int end_gen, int end_solve,
              const uint64_t *winstates, size_t n_win, double p4, double **values)
{
    uint64_t *W = malloc((8 * n_win + 1) * sizeof(uint64_t));
    size_t nW = orc_expand_winstates(winstates, n_win, W);
    for (int i = 0; i < nlayers; i++) values[i] = NULL;
    const uint64_t *K2 = NULL, *K4 = NULL;
    const double *V2 = NULL, *V4 = NULL;
    size_t n2 = 0, n4 = 0;
    /* solve.c:203: for (s = start; s >= end && s; s -= 2), unsigned */
    for (long s = end_gen; s >= end_solve && s > 0; s -= 2) {
        long idx = (s - start_gen) / 2;
        const uint64_t *M = NULL;
        size_t nM = 0;
        double *V = NULL;
        if (s >= start_gen && (s - start_gen) % 2 == 0 && idx < nlayers) {
            M = layers[idx];
            nM = sizes[idx];
            V = malloc((nM ? nM : 1) * sizeof(double));
            orc_solve_layer(cfg, M, nM, K2, V2, n2, K4, V4, n4, W, nW, p4, V);
            values[idx] = V;
        }
        K4 = K2; V4 = V2; n4 = n2;
        K2 = M; V2 = V; n2 = nM;
    }
    free(W);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Adjoint (no reference counterpart).                                       */
/* ------------------------------------------------------------------------ */

/* best move under the CLI's tie rule (frontends/cli/main.c:174-189): the first
 * strict maximum over left, right, up, down starting from 0. */
static double best_move(const lut_t *L, uint64_t b, const uint64_t *K, const double *V, size_t n,
                        uint64_t *succ, int *found)
{
    double maxp = 0;
    *found = 0;
    for (int d = 0; d < 4; d++) {
        uint64_t t = b;
        if (!move_unstable(L, &t, d)) continue;
        double p = orc_lookup(t, K, V, n, 1);
        if (p > maxp) { maxp = p; *succ = orc_canon(t); *found = 1; }
    }
    return maxp;
}

int orc_adjoint(const orc_cfg *cfg, uint64_t **layers, const size_t *sizes, int nlayers,
                int start_gen, double **values, const uint64_t *W, size_t nW,
                const uint64_t *initial, size_t n_init, double p4,
                double *P_out, double *dP_out, double *D)
{
    const lut_t *L = get_lut(cfg);
    double **occ = calloc((size_t)nlayers + 2, sizeof(double *));
    for (int i = 0; i < nlayers; i++) {
        occ[i] = calloc(sizes[i] ? sizes[i] : 1, sizeof(double));
        D[i] = 0;
    }
    static const double empty_v = 0;
    double P = 0;
    if (nlayers > 0 && values[0]) {
        for (size_t i = 0; i < n_init; i++) {
            uint64_t succ = 0;
            int found;
            double v;
            if (cfg->premove) {   /* the initial board is a layer-0 position */
                succ = orc_canon(initial[i]);
                long j = find_index(succ, layers[0], sizes[0]);
                found = j >= 0;
                v = found ? values[0][j] : 0.0;
            } else {
                v = best_move(L, initial[i], layers[0], values[0], sizes[0], &succ, &found);
            }
            P += v / (double)n_init;
            if (found) {
                long j = find_index(succ, layers[0], sizes[0]);
                if (j >= 0) occ[0][j] += 1.0 / (double)n_init;
            }
        }
    }
    double w2 = 1.0 - p4, w4 = p4;
    double dP = 0;
    for (int li = 0; li < nlayers; li++) {
        if (!values[li]) continue;
        const uint64_t *M = layers[li];
        for (size_t i = 0; i < sizes[li]; i++) {
            double o = occ[li][i];
            if (o == 0) continue;
            uint64_t b = M[i];
            if (satisfied(cfg, b, W, nW)) continue;
            int E = 0;
            for (int x = 0; x < 16; x++) E += TILE(b, x) == 0;
            if (!E) continue;
            double a2 = 0, a4 = 0;
            for (int x = 0; x < 16; x++) {
                if (TILE(b, x)) continue;
                for (int k = 0; k < 2; k++) {
                    int tl = li + 1 + k;   /* layer s+2 or s+4 */
                    const uint64_t *K = (tl < nlayers && values[tl]) ? layers[tl] : NULL;
                    const double *V = K ? values[tl] : &empty_v;
                    size_t n = K ? sizes[tl] : 0;
                    uint64_t c = b;
                    SETT(c, x, (uint64_t)(k + 1));
                    uint64_t succ = 0;
                    int found;
                    double m = best_move(L, c, K, V, n, &succ, &found);
                    if (k == 0) a2 += m; else a4 += m;
                    if (found) {
                        long j = find_index(succ, K, n);
                        if (j >= 0) occ[tl][j] += o * (k == 0 ? w2 : w4) / E;
                    }
                }
            }
            D[li] += o * (a4 - a2) / E;
        }
        dP += D[li];
    }
    for (int i = 0; i < nlayers; i++) free(occ[i]);
    free(occ);
    *P_out = P;
    *dP_out = dP;
    return 0;
}

uint64_t orc_hash_u64(const uint64_t *a, size_t n)
{
    /* FNV-1a over 64-bit words (not bytes): h = (h ^ w) * prime */
    uint64_t h = 1469598103934665603ULL;
    for (size_t i = 0; i < n; i++) {
        h ^= a[i];
        h *= 1099511628211ULL;
    }
    return h;
}

void orc_free(void *p) { free(p); }
