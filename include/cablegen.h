/*
 * cablegen.h -- drop-in C ABI of the MI355X engine.
 *
 * Declares exactly the types, macros and functions of the reference's public
 * header (api/cablegen.h:1-80) with identical names, signatures and struct
 * layouts, so that a program built against the reference (its CLI,
 * frontends/cli/main.c, or gen_conf) links against libcablegen.so unchanged.
 * generate() and solve() run on the GPU; the remaining entry points are host
 * utilities with the reference's semantics.  Extensions (device-resident
 * pipeline, adjoint, batching, device primitives) are in cablegen_gpu.h.
 */
#ifndef CABLEGEN_H
#define CABLEGEN_H
#include <stdbool.h>
#include <stdint.h>
#include <stddef.h>

/* bit helpers and tile access, api/cablegen.h:6-12: tile x is the nibble at
 * bits [60-4x, 63-4x] */
#define SETBIT(x, y) (x |= (1 << y))
#define CLEARBIT(x, y) (x &= ~(1 << y))
#define GETBIT(x, y) (x & (1 << y))
#define BIT_MASK (0xf000000000000000)
#define OFFSET(x) (x * 4)
#define GET_TILE(b, x) (uint8_t)(((b << OFFSET(x)) & BIT_MASK) >> OFFSET(15))
#define SET_TILE(b, x, v) ( b = (~(~b | (BIT_MASK >> OFFSET(x))) | ((BIT_MASK & ((uint64_t)v << OFFSET(15))) >> OFFSET(x)) ))

#ifndef REMOVE_STRUCT_DECL
/* api/cablegen.h:15-20 (32 bytes) */
typedef struct {
    bool valid;
    uint64_t *bp;
    uint64_t *sp;
    size_t size;
} dynamic_arr_info;

/* api/cablegen.h:22-26 (24 bytes) */
typedef struct {
    bool valid;
    uint64_t *bp;
    size_t size;
} static_arr_info;

/* api/cablegen.h:28-31 */
typedef struct {
    static_arr_info key;
    static_arr_info value;
} table;

/* api/cablegen.h:33-38 */
typedef enum {
    left,
    right,
    up,
    down
} dir;
#endif

/* replaces movedir, src/board.c:171-193 (up/down rotate back) */
bool movedir(uint64_t *, dir);
/* replaces generate_lut, src/board.c:100-132 (host row tables) */
void generate_lut(void);
/* replaces read_boards, src/generate.c:616-648 */
static_arr_info read_boards(const char *dir);
/* replaces read_boards2, src/generate.c:650-652 */
void read_boards2(static_arr_info *b, const char *dir);
/* replaces change_config, src/settings.c:110-120 */
void change_config(char *cfg);
/* replaces lookup, src/solve.c:108-164 (exact match, 0 on a miss) */
double lookup(uint64_t key, table *t, bool canonicalize);

/* api/cablegen.h:47-59; first member of the settings singleton */
typedef struct {
    bool free_formation;
    bool ignore_f;
    long long cores;
    long long nox;
    char *bdir;
    char *initial;
    long long end_gen;
    char *tdir;
    char *winstates;
    long long end_solve;
    bool delete_boards;
} min_settings_t;

/* replaces get_settings_min, src/settings.c:30-32 */
min_settings_t *get_settings_min(void);
/* replaces init_settings, src/settings.c:121-138 */
void init_settings(void);
/* replaces generate, src/generate.c:551-615: layers s = start..end step 2 on the GPU */
void generate(const int start, const int end, const char *fmt, const static_arr_info *initial);
/* replaces solve, src/solve.c:172-224: layers s = start down to end on the GPU */
void solve(unsigned start, unsigned end, char *posfmt, char *tablefmt, const static_arr_info *winstates);
/* replaces malloc_errcheck, src/array.c:178-191 */
void *malloc_errcheck(size_t size);
/* replace the array helpers of src/array.c:17-136, 200-207 */
dynamic_arr_info init_darr(bool zero, size_t size);
static_arr_info init_sarr(bool zero, size_t size);
static_arr_info shrink_darr(dynamic_arr_info *info);
dynamic_arr_info concat(dynamic_arr_info *arr1, dynamic_arr_info *arr2);
void destroy_darr(dynamic_arr_info *arr);
void destroy_sarr(static_arr_info *arr);
bool push_back(dynamic_arr_info *, uint64_t);
/* replaces write_boards, src/generate.c:74-87 */
void write_boards(const static_arr_info n, const char *fmt, const int layer);
/* replaces read_table, src/solve.c:47-106 */
void read_table(table *t, const char *path);
/* replaces pun_uint64, src/generate.c:654-656 */
double pun_uint64(uint64_t);
/* replaces test, src/test.c:366-385 (self-test) */
bool test(void);
/* replaces get_sum, src/board.c:340-347 */
int get_sum(uint64_t board);
/* replaces output_board, src/board.c:395-425 */
void output_board(uint64_t);
/* replaces get_version, src/test.c:18-20 */
char *get_version(void);
#endif
