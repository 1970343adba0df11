/*
 * host.c -- the drop-in C ABI (include/cablegen.h) of the MI355X engine.
 *
 * generate() and solve() drive the gfx950 engine through the thin internal
 * ABI in engine.h; there is no CPU fallback: a device failure is reported
 * and the process exits, as the reference does on allocation failure
 * (src/array.c:178-191).  The remaining entry points are host utilities with
 * the reference's semantics: settings singleton and INI keys
 * (src/settings.c:34-138), growable arrays (src/array.c:17-136), the
 * %d.boards / %d.tables formats (src/generate.c:74-87, src/solve.c:29-106),
 * host-side movedir/lookup for table consumers (src/board.c:171-193,
 * src/solve.c:108-164).
 */
#define _GNU_SOURCE
#include <ctype.h>
#include <errno.h>
#include <limits.h>
#include <pthread.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <time.h>

#include "../../include/cablegen.h"
#include "../../include/cablegen_gpu.h"
#include "board.h"
#include "engine.h"

#define CG_VERSION "cablegen-mi355x 0.1"

/* ------------------------------------------------------------------------ */
/* logging (warnings and errors on stderr)                                   */
/* ------------------------------------------------------------------------ */

static void warnf(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    fputs("[WARN] ", stderr);
    vfprintf(stderr, fmt, ap);
    fputc('\n', stderr);
    va_end(ap);
}

static void die(const char *what)
{
    fprintf(stderr, "[ERROR] %s: %s\n", what, cge_last_error());
    exit(EXIT_FAILURE);
}

/* ------------------------------------------------------------------------ */
/* row table (left moves)                                                    */
/* ------------------------------------------------------------------------ */

/* Left move of one 16-bit row (board.c:62-98): with ignore_f the trailing run
 * of 0xF cells is left in place; the remaining cells slide toward column 0,
 * equal neighbours below 0xE merge once (left to right), and slide again. */
static unsigned row_left(unsigned r, int ignore_f)
{
    unsigned c[4] = { r >> 12 & 0xf, r >> 8 & 0xf, r >> 4 & 0xf, r & 0xf };
    int n = 4;
    if (ignore_f)
        while (n > 0 && c[n - 1] == 0xf) n--;
    unsigned t[4], k = 0;
    for (int i = 0; i < n; i++)
        if (c[i]) t[k++] = c[i];
    unsigned o[4] = { 0, 0, 0, 0 }, j = 0;
    for (unsigned i = 0; i < k; i++) {
        if (i + 1 < k && t[i] == t[i + 1] && t[i] < 0xe) {
            o[j++] = t[i] + 1;
            i++;
        } else {
            o[j++] = t[i];
        }
    }
    for (int i = n; i < 4; i++) o[i] = c[i];
    return o[0] << 12 | o[1] << 8 | o[2] << 4 | o[3];
}

void cge_build_row_table(int ignore_f, uint16_t *tab)
{
    for (unsigned r = 0; r < 65536; r++) tab[r] = (uint16_t)row_left(r, ignore_f);
}

static uint16_t g_tab[65536];
static int g_tab_ready = 0, g_tab_ignore_f = -1;

/* ------------------------------------------------------------------------ */
/* settings (src/settings.c:34-138)                                          */
/* ------------------------------------------------------------------------ */

typedef struct {
    min_settings_t min;
    bool delete_boards;
    bool mask;
    bool compress;
    long long max_prealloc;
    bool use_gpu;
    bool premove;
    long long stsl;
    long long ltc;
    long long smallest_large;
    bool prune;
    bool score;
} settings_t;

static settings_t g_set = {
    .min = { .free_formation = false, .ignore_f = false, .cores = 1, .nox = 0,
             .bdir = "./boards/", .initial = "./initial", .end_gen = 1200,
             .tdir = "./tables/", .winstates = "./winstates", .end_solve = 0,
             .delete_boards = false },
    .delete_boards = false, .mask = false, .compress = false, .max_prealloc = 0,
    .use_gpu = true, .premove = false, .stsl = 200, .ltc = 5, .smallest_large = 6,
    .prune = false, .score = false,
};
static bool g_set_read = false;
static bool g_custom_path = false;
static char g_cfg_path[4096];

/* INI lookup: section and key compare case-insensitively; ';' and '#'
 * start comments; returns a malloc'd trimmed value or NULL. */
static char *ini_get_value(const char *path, const char *section, const char *key)
{
    FILE *f = fopen(path, "r");
    if (!f) return NULL;
    char line[4096], cur[256] = "";
    char *res = NULL;
    while (fgets(line, sizeof line, f)) {
        char *p = line;
        while (isspace((unsigned char)*p)) p++;
        if (*p == ';' || *p == '#' || *p == 0) continue;
        if (*p == '[') {
            char *e = strchr(p, ']');
            if (!e) continue;
            *e = 0;
            snprintf(cur, sizeof cur, "%s", p + 1);
            continue;
        }
        char *eq = strchr(p, '=');
        if (!eq) continue;
        *eq = 0;
        char *k = p, *v = eq + 1;
        char *ke = k + strlen(k);
        while (ke > k && isspace((unsigned char)ke[-1])) *--ke = 0;
        while (isspace((unsigned char)*v)) v++;
        char *ve = v + strlen(v);
        while (ve > v && isspace((unsigned char)ve[-1])) *--ve = 0;
        if (!strcasecmp(cur, section) && !strcasecmp(k, key)) {
            free(res);
            res = strdup(v);
        }
    }
    fclose(f);
    return res;
}

static const char *config_path(void)
{
    if (g_custom_path) return g_cfg_path;
    const char *x = getenv("XDG_CONFIG_HOME");
    if (x) snprintf(g_cfg_path, sizeof g_cfg_path, "%s/cablegen.conf", x);
    else {
        const char *h = getenv("HOME");
        if (!h) return NULL;
        snprintf(g_cfg_path, sizeof g_cfg_path, "%s/.config/cablegen.conf", h);
    }
    return g_cfg_path;
}

static void get_bool(const char *sec, const char *key, bool *out)
{
    const char *p = config_path();
    if (!p) return;
    char *v = ini_get_value(p, sec, key);
    if (!v) return;
    for (char *c = v; *c; c++) *c = (char)tolower((unsigned char)*c);
    *out = !strcmp(v, "true");
    free(v);
}

static void get_int(const char *sec, const char *key, long long *out)
{
    const char *p = config_path();
    if (!p) return;
    char *v = ini_get_value(p, sec, key);
    if (!v) return;
    *out = atoll(v);
    free(v);
}

static void get_str(const char *sec, const char *key, char **out)
{
    const char *p = config_path();
    if (!p) return;
    char *v = ini_get_value(p, sec, key);
    if (v) *out = v;
}

static settings_t *get_settings(void)
{
    if (g_set_read) return &g_set;
    g_set_read = true;
    get_bool("Cablegen", "free_formation", &g_set.min.free_formation);
    get_bool("Cablegen", "ignore_f", &g_set.min.ignore_f);
    get_int("Cablegen", "cores", &g_set.min.cores);
    get_int("Cablegen", "nox", &g_set.min.nox);
    get_bool("Cablegen", "mask", &g_set.mask);
    get_bool("Cablegen", "compress", &g_set.compress);
    get_int("Cablegen", "max_prealloc", &g_set.max_prealloc);
    get_bool("Cablegen", "use_gpu", &g_set.use_gpu);
    get_bool("Generate", "premove", &g_set.premove);
    get_str("Generate", "dir", &g_set.min.bdir);
    get_str("Generate", "initial", &g_set.min.initial);
    get_int("Generate", "end", &g_set.min.end_gen);
    get_int("Generate", "stsl", &g_set.stsl);
    get_int("Generate", "ltc", &g_set.ltc);
    get_int("Generate", "smallest_large", &g_set.smallest_large);
    get_bool("Generate", "prune", &g_set.prune);
    get_str("Solve", "dir", &g_set.min.tdir);
    get_str("Solve", "winstates", &g_set.min.winstates);
    get_int("Solve", "end", &g_set.min.end_solve);
    get_bool("Solve", "score", &g_set.score);
    get_bool("Solve", "delete_boards", &g_set.delete_boards);
    return &g_set;
}

min_settings_t *get_settings_min(void) { return (min_settings_t *)get_settings(); }

void change_config(char *cfg)
{
    if (!cfg || strlen(cfg) >= sizeof g_cfg_path) {
        warnf("Invalid config file location!");
        return;
    }
    g_set_read = false;
    g_custom_path = true;
    memcpy(g_cfg_path, cfg, strlen(cfg) + 1);
}

void init_settings(void)
{
    char buf[4096];
    if (!getcwd(buf, sizeof buf - 16)) return;
    get_settings();
    strcat(buf, "/cablegen.conf");
    change_config(buf);
    get_settings();
}

void cg_rules_from_settings(cg_rules *r)
{
    settings_t *s = get_settings();
    r->free_formation = s->min.free_formation;
    r->ignore_f = s->min.ignore_f;
    r->prune = s->prune;
    r->stsl = (long)s->stsl;
    r->ltc = (long)s->ltc;
    r->smallest_large = (long)s->smallest_large;
    /* the reference passes nox as a char (solve.c:172-175) */
    r->nox = (int)(signed char)s->min.nox;
    r->score = s->score;
    r->premove = s->premove;
}

static cge_rules engine_rules(void)
{
    cg_rules r;
    cg_rules_from_settings(&r);
    cge_rules e = { r.free_formation, r.ignore_f, r.prune, r.stsl, r.ltc, r.smallest_large,
                    r.nox, r.score, r.premove };
    return e;
}

/* ------------------------------------------------------------------------ */
/* arrays (src/array.c)                                                      */
/* ------------------------------------------------------------------------ */

void *malloc_errcheck(size_t size)
{
    void *p = malloc(size ? size : 1);
    if (!p) {
        fprintf(stderr, "[ERROR] Alloc failed!\n");
        exit(EXIT_FAILURE);
    }
    return p;
}

bool push_back(dynamic_arr_info *a, uint64_t v)
{
    bool grew = false;
    if (!a->valid || !a->sp) {
        warnf("Invalid array, refusing to push");
        a->valid = false;
        return false;
    }
    if (a->sp == a->bp + a->size) {
        size_t used = (size_t)(a->sp - a->bp);
        size_t nsz = a->size ? a->size * 3 : 1;   /* x3 growth */
        uint64_t *nb = realloc(a->bp, nsz * sizeof(uint64_t));
        if (!nb) {
            a->valid = false;
            return false;
        }
        a->bp = nb;
        a->size = nsz;
        a->sp = nb + used;
        grew = true;
    }
    *a->sp++ = v;
    return grew;
}

dynamic_arr_info init_darr(bool zero, size_t size)
{
    dynamic_arr_info d;
    d.valid = true;
    d.bp = zero ? calloc(size, sizeof(uint64_t)) : malloc(sizeof(uint64_t) * size);
    if (!d.bp) d.valid = false;
    d.sp = d.bp;
    d.size = size;
    return d;
}

static_arr_info init_sarr(bool zero, size_t size)
{
    static_arr_info s;
    s.valid = true;
    s.bp = zero ? calloc(size, sizeof(uint64_t)) : malloc(sizeof(uint64_t) * size);
    if (!s.bp) s.valid = false;
    s.size = size;
    return s;
}

static_arr_info shrink_darr(dynamic_arr_info *a)
{
    size_t n = (size_t)(a->sp - a->bp);
    if (!a->valid) return (static_arr_info){ .valid = false, .bp = a->bp, .size = n };
    a->valid = false;
    if (n == 0) {
        free(a->bp);
        a->bp = a->sp = NULL;
        a->size = 0;
        return (static_arr_info){ .valid = true, .bp = NULL, .size = 0 };
    }
    uint64_t *nb = realloc(a->bp, n * sizeof(uint64_t));
    if (!nb) return (static_arr_info){ .valid = false, .bp = a->bp, .size = n };
    return (static_arr_info){ .valid = true, .bp = nb, .size = n };
}

dynamic_arr_info concat(dynamic_arr_info *a1, dynamic_arr_info *a2)
{
    if (!a1->valid || !a2->valid) {
        warnf("Refusing to concatenate invalid arrays!");
        return *a1;
    }
    a1->valid = a2->valid = false;
    size_t n1 = (size_t)(a1->sp - a1->bp), n2 = (size_t)(a2->sp - a2->bp);
    dynamic_arr_info r = { .valid = true, .bp = NULL, .sp = NULL, .size = n1 + n2 };
    if (a1->size >= r.size) {
        r.bp = a1->bp;
    } else {
        r.bp = malloc_errcheck(r.size * sizeof(uint64_t));
        if (n1) memcpy(r.bp, a1->bp, n1 * sizeof(uint64_t));
        free(a1->bp);
    }
    if (n2) memcpy(r.bp + n1, a2->bp, n2 * sizeof(uint64_t));
    free(a2->bp);
    r.sp = r.bp + r.size;
    return r;
}

void destroy_darr(dynamic_arr_info *a)
{
    free(a->bp);
    a->valid = false;
}

void destroy_sarr(static_arr_info *a)
{
    free(a->bp);
    a->valid = false;
}

/* ------------------------------------------------------------------------ */
/* strings (src/format.c), exported for the reference CLI                    */
/* ------------------------------------------------------------------------ */

typedef struct { long res; int status; } atoi_res;
typedef struct { long res; int status; } atoi_res_ll;

char *format_str(const char *fmt, ...)
{
    va_list a, b;
    va_start(a, fmt);
    va_copy(b, a);
    int n = vsnprintf(NULL, 0, fmt, a);
    va_end(a);
    char *buf = malloc_errcheck((size_t)(n < 0 ? 0 : n) + 1);
    vsnprintf(buf, (size_t)(n < 0 ? 0 : n) + 1, fmt, b);
    va_end(b);
    return buf;
}

static int digit_val(char c)
{
    if (c >= '0' && c <= '9') return c - '0';
    if (isalpha((unsigned char)c)) return tolower((unsigned char)c) - 'a' + 10;
    return -1;
}

bool isdigit_radix(const char input, const unsigned int radix)
{
    int v = digit_val(input);
    return v >= 0 && (unsigned)v < radix;
}

static int parse_status(const char *input, long long v, int radix, long long vmin, long long vmax)
{
    if (v == vmax || v == vmin) return ERANGE;
    if (errno) {
        int e = errno;
        errno = 0;
        return e;
    }
    if (!v) {
        for (const char *c = input; *c; c++) {
            if (isspace((unsigned char)*c)) continue;
            if (!isdigit_radix(*c, (unsigned)radix)) return EINVAL;
        }
    }
    return 0;
}

atoi_res safe_atoi(const char *input)
{
    atoi_res r = { 0, 0 };
    errno = 0;
    r.res = strtol(input, NULL, 10);
    r.status = parse_status(input, r.res, 10, LONG_MIN, LONG_MAX);
    return r;
}

atoi_res_ll safe_atoi_ll(const char *input, const int radix)
{
    atoi_res_ll r = { 0, 0 };
    if (radix < 2 || radix > 36) {
        r.status = EINVAL;
        return r;
    }
    errno = 0;
    r.res = strtoll(input, NULL, radix);
    r.status = parse_status(input, r.res, radix, LLONG_MIN, LLONG_MAX);
    return r;
}

/* ------------------------------------------------------------------------ */
/* board utilities                                                           */
/* ------------------------------------------------------------------------ */

void generate_lut(void)
{
    int ig = get_settings()->min.ignore_f ? 1 : 0;
    if (!g_tab_ready || g_tab_ignore_f != ig) {
        cge_build_row_table(ig, g_tab);
        g_tab_ready = 1;
        g_tab_ignore_f = ig;
    }
}

/* movedir (board.c:171-193): the board is returned in its own orientation */
bool movedir(uint64_t *board, dir d)
{
    generate_lut();
    int ff = get_settings()->min.free_formation;
    uint64_t b = *board;
    int ok;
    if (d == up || d == down) {
        b = cg_rot_cw(b);
        ok = cg_move_hori_tab(g_tab, &b, d == up, ff);
        if (ok) {
            /* rotate back: three more clockwise turns */
            b = cg_rot_cw(cg_rot_cw(cg_rot_cw(b)));
            *board = b;
        }
    } else {
        ok = cg_move_hori_tab(g_tab, &b, d == right, ff);
        if (ok) *board = b;
    }
    return ok != 0;
}

int get_sum(uint64_t board) { return cg_tile_sum(board); }

double pun_uint64(uint64_t num)
{
    double d;
    memcpy(&d, &num, sizeof d);
    return d;
}

/* lookup (solve.c:108-164): exact match in a sorted table, 0.0 on a miss or
 * an empty table */
double lookup(uint64_t key, table *t, bool canonicalize)
{
    if (t->key.size == 0) return 0.0;
    if (canonicalize) key = cg_canon(key);
    size_t lo = 0, hi = t->key.size;
    while (lo < hi) {
        size_t mid = lo + (hi - lo) / 2;
        if (t->key.bp[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    if (lo < t->key.size && t->key.bp[lo] == key) return pun_uint64(t->value.bp[lo]);
    return 0.0;
}

void output_board(uint64_t board)
{
    static const char *names[16] = { "   ", "2  ", "4  ", "8  ", "16 ", "32 ", "64 ", "128",
                                     "256", "512", "1k ", "2k ", "4k ", "8k ", "m  ", "x  " };
    for (int r = 0; r < 4; r++) {
        for (int c = 0; c < 4; c++) printf("%s ", names[(board >> (60 - 4 * (4 * r + c))) & 0xf]);
        printf("\n");
    }
}

char *get_version(void) { return CG_VERSION; }

/* ------------------------------------------------------------------------ */
/* files                                                                     */
/* ------------------------------------------------------------------------ */

void write_boards(const static_arr_info n, const char *fmt, const int layer)
{
    char *name = format_str(fmt, layer);
    FILE *f = fopen(name, "wb");
    if (!f) {
        warnf("Couldn't write to %s!", name);
        free(name);
        return;
    }
    if (n.size) fwrite(n.bp, sizeof(uint64_t), n.size, f);
    fclose(f);
    free(name);
}

static long file_size(FILE *f)
{
    if (fseek(f, 0L, SEEK_END)) return -1;
    long sz = ftell(f);
    rewind(f);
    return sz;
}

static_arr_info read_boards(const char *path)
{
    FILE *f = fopen(path, "rb");
    if (!f) {
        warnf("Couldn't read %s!", path);
        return (static_arr_info){ .valid = false };
    }
    long sz = file_size(f);
    if (sz < 0) {
        fclose(f);
        return (static_arr_info){ .valid = false };
    }
    if (sz % 8) warnf("sz %%8 != 0, this is probably not a real table!");
    uint64_t *data = malloc_errcheck((size_t)sz);
    if (fread(data, 1, (size_t)sz, f) != (size_t)sz) {
        fclose(f);
        free(data);
        warnf("Couldn't read %s!", path);
        return (static_arr_info){ .valid = false };
    }
    fclose(f);
    return (static_arr_info){ .valid = true, .bp = data, .size = (size_t)sz / 8 };
}

void read_boards2(static_arr_info *b, const char *path) { *b = read_boards(path); }

static void write_table_file(const uint64_t *k, const double *v, size_t n, const char *path)
{
    FILE *f = fopen(path, "wb");
    if (!f) {
        warnf("Couldn't write to %s!", path);
        return;
    }
    if (n) {
        fwrite(k, sizeof(uint64_t), n, f);
        fwrite(v, sizeof(double), n, f);
    }
    fclose(f);
}

void read_table(table *t, const char *path)
{
    t->key = (static_arr_info){ .valid = false };
    t->value = (static_arr_info){ .valid = false };
    FILE *f = fopen(path, "rb");
    if (!f) {
        warnf("Couldn't read %s!", path);
        t->key = init_sarr(0, 0);
        t->value = init_sarr(0, 0);
        t->key.valid = t->value.valid = false;
        return;
    }
    long sz = file_size(f);
    if (sz < 0) {
        fclose(f);
        t->key = init_sarr(0, 0);
        t->value = init_sarr(0, 0);
        t->key.valid = t->value.valid = false;
        return;
    }
    if (sz % 16) warnf("sz %%16 != 0, this is probably not a real table!");
    size_t n = (size_t)sz / 16;
    t->key = init_sarr(0, n);
    t->value = init_sarr(0, n);
    if (fread(t->key.bp, sizeof(uint64_t), n, f) != n ||
        fread(t->value.bp, sizeof(double), n, f) != n) {
        warnf("Error reading file!");
        t->key.valid = t->value.valid = false;
    }
    fclose(f);
}

/* ------------------------------------------------------------------------ */
/* generate() and solve() on the GPU, file I/O overlapped with the GPU work  */
/* ------------------------------------------------------------------------ */

/* Layers move between the GPU and files through a ring of slots in pinned
 * host memory: a slot is filled by a device copy (or a file read) and
 * emptied by a writer thread, so several files are written at once while the
 * GPU works on the next layers.  Buffers are (re)allocated on the calling
 * thread only, after the slot's previous writer has finished. */
#define IO_SLOTS 6

typedef struct {
    pthread_t th;
    int busy;
    uint64_t *k;     /* keys (pinned) */
    double *v;       /* values (pinned; solve only) */
    size_t cap, n;
    char *path;
    int ok;
} io_slot_t;

static void *board_writer(void *arg)
{
    io_slot_t *w = arg;
    FILE *f = fopen(w->path, "wb");
    if (!f) {
        warnf("Couldn't write to %s!", w->path);
        return NULL;
    }
    if (w->n) fwrite(w->k, sizeof(uint64_t), w->n, f);
    fclose(f);
    return NULL;
}

static void *table_writer(void *arg)
{
    io_slot_t *w = arg;
    write_table_file(w->k, w->v, w->n, w->path);
    return NULL;
}

/* reads a %d.boards file into the slot (capacity ensured by the caller) */
static void *board_reader(void *arg)
{
    io_slot_t *w = arg;
    w->ok = 0;
    FILE *f = fopen(w->path, "rb");
    if (!f) {
        warnf("Couldn't read %s!", w->path);
        return NULL;
    }
    if (w->n && fread(w->k, sizeof(uint64_t), w->n, f) != w->n) {
        fclose(f);
        warnf("Couldn't read %s!", w->path);
        return NULL;
    }
    fclose(f);
    w->ok = 1;
    if (get_settings()->delete_boards) remove(w->path);
    return NULL;
}

static void slot_join(io_slot_t *w)
{
    if (w->busy) {
        pthread_join(w->th, NULL);
        w->busy = 0;
    }
}

static void slot_run(io_slot_t *w, void *(*fn)(void *))
{
    if (pthread_create(&w->th, NULL, fn, w)) {
        fn(w);
    } else {
        w->busy = 1;
    }
}

/* capacity for n keys (and values when with_vals) */
static void slot_reserve(io_slot_t *w, size_t n, int with_vals, const char *what)
{
    if (n <= w->cap && w->k) return;
    cge_host_free(w->k);
    cge_host_free(w->v);
    w->cap = n + n / 4 + 16;
    w->k = cge_host_alloc(w->cap * sizeof(uint64_t));
    w->v = with_vals ? cge_host_alloc(w->cap * sizeof(double)) : NULL;
    if (!w->k || (with_vals && !w->v)) die(what);
}

static void slots_release(io_slot_t *w)
{
    for (int i = 0; i < IO_SLOTS; i++) {
        slot_join(&w[i]);
        cge_host_free(w[i].k);
        cge_host_free(w[i].v);
        free(w[i].path);
    }
}

void generate(const int start, const int end, const char *fmt, const static_arr_info *initial)
{
    cge_rules r = engine_rules();
    if (cge_gen_begin(&r, initial->bp, initial->size)) die("generate");
    io_slot_t w[IO_SLOTS];
    memset(w, 0, sizeof w);
    int k = 0;
    for (int s = start; s <= end; s += 2) {
        size_t n = 0;
        if (cge_gen_step(&n)) die("generate");
        io_slot_t *x = &w[k];
        slot_join(x);
        slot_reserve(x, n, 0, "generate");
        if (cge_gen_fetch(x->k, n)) die("generate");
        x->n = n;
        free(x->path);
        x->path = format_str(fmt, s);
        slot_run(x, board_writer);
        k = (k + 1) % IO_SLOTS;
    }
    slots_release(w);
    cge_gen_end();
}

static long path_size(const char *path)
{
    struct stat st;
    if (stat(path, &st)) return -1;
    return (long)st.st_size;
}

/* starts reading layer file `path` into slot x; a missing file reads as an
 * empty layer (x->ok = 0) */
static void start_read(io_slot_t *x, const char *path)
{
    slot_join(x);
    free(x->path);
    x->path = strdup(path);
    long sz = path_size(path);
    if (sz < 0) {
        warnf("Couldn't read %s!", path);
        x->n = 0;
        x->ok = 0;
        return;
    }
    if (sz % 8) warnf("sz %%8 != 0, this is probably not a real table!");
    x->n = (size_t)sz / 8;
    slot_reserve(x, x->n, 1, "solve");
    slot_run(x, board_reader);
}

void solve(unsigned start, unsigned end, char *posfmt, char *tablefmt, const static_arr_info *winstates)
{
    cge_rules r = engine_rules();
    size_t nw = winstates->size;
    uint64_t *img = malloc_errcheck((8 * nw + 1) * sizeof(uint64_t));
    for (size_t i = 0; i < nw; i++) {   /* get_all_rots, board.c:349-374 */
        uint64_t b = winstates->bp[i], x = b;
        img[8 * i] = b;
        for (int j = 1; j < 4; j++) img[8 * i + j] = x = cg_rot_cw(x);
        img[8 * i + 4] = x = cg_flip_rows(b);
        for (int j = 5; j < 8; j++) img[8 * i + j] = x = cg_rot_cw(x);
    }
    if (cge_solve_begin(&r, img, 8 * nw, 0.1)) die("solve");
    free(img);
    /* the layers, top down, as the reference walks them */
    size_t nl = 0;
    for (unsigned s = start; s >= end && s; s -= 2) {
        nl++;
        if (s < 2) break;
    }
    io_slot_t w[IO_SLOTS];
    memset(w, 0, sizeof w);
    char name[4096];
    if (nl) {
        snprintf(name, sizeof name, posfmt, start);
        start_read(&w[0], name);
    }
    for (size_t i = 0; i < nl; i++) {
        const unsigned s = start - 2 * (unsigned)i;
        io_slot_t *x = &w[i % IO_SLOTS];
        /* x: read of layer s (started one layer earlier) */
        if (x->busy) slot_join(x);
        if (!x->ok) x->n = 0;
        if (i + 1 < nl) {
            /* prefetch the next layer's file while this one is on the GPU */
            snprintf(name, sizeof name, posfmt, s - 2);
            start_read(&w[(i + 1) % IO_SLOTS], name);
        }
        slot_reserve(x, x->n, 1, "solve");
        if (cge_solve_step(x->k, x->n, x->v)) die("solve");
        snprintf(name, sizeof name, tablefmt, s);
        free(x->path);
        x->path = strdup(name);
        slot_run(x, table_writer);
    }
    slots_release(w);
    cge_solve_end();
}

/* ------------------------------------------------------------------------ */
/* self-test (src/test.c:366-385 runs the reference's equivalents)           */
/* ------------------------------------------------------------------------ */

bool test(void)
{
    bool ok = true;
    /* lookup known answer: key i -> 99 - i (test.c:22-52) */
    table t = { init_sarr(0, 100), init_sarr(0, 100) };
    for (int i = 0; i < 100; i++) {
        double v = 99 - i;
        t.key.bp[i] = (uint64_t)i;
        memcpy(&t.value.bp[i], &v, 8);
    }
    for (int i = 0; i < 100; i++) ok &= lookup((uint64_t)i, &t, false) == 99.0 - i;
    ok &= lookup(1000, &t, false) == 0.0;
    destroy_sarr(&t.key);
    destroy_sarr(&t.value);
    /* symmetry invariants (test.c:189-237) on full 64-bit boards */
    uint64_t x = 0x9e3779b97f4a7c15ULL;
    for (int i = 0; i < 10000; i++) {
        x ^= x << 13;
        x ^= x >> 7;
        x ^= x << 17;
        uint64_t b = x, c = cg_canon(b), y = b;
        for (int k = 0; k < 4; k++) {
            ok &= cg_canon(y) == c && cg_canon(cg_flip_rows(y)) == c;
            ok &= cg_tile_sum(y) == cg_tile_sum(b);
            y = cg_rot_cw(y);
        }
        ok &= y == b;
    }
    /* cmpbrd subset rule (test.c:335-364) */
    ok &= (0x1234000000000000ULL & cg_win_mask(0x1004000000000000ULL)) == 0x1004000000000000ULL;
    ok &= (0x1234000000000000ULL & cg_win_mask(0x2000000000000000ULL)) != 0x2000000000000000ULL;
    return ok;
}
