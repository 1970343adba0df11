/*
 * cablegen_gpu.h -- extension entry points of the MI355X engine.
 *
 * These have no counterpart in the reference's public header
 * (api/cablegen.h); the drop-in entry points are in cablegen.h.  They expose
 * the device-resident pipeline (layers kept in HBM between generate, solve
 * and the adjoint), the per-stage device primitives used by the kernel
 * microbenchmarks and parity tests, the batched multi-table run and the
 * per-kernel timers.
 *
 * Conventions: plain C types, host or device pointers as stated, sizes in
 * elements.  Every function returns 0 on success and a negative value on
 * failure; cg_last_error() gives the message.  Calls are blocking unless
 * stated and, like the reference's API, not reentrant.
 */
#ifndef CABLEGEN_GPU_H
#define CABLEGEN_GPU_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* The rules that shape layers and values: the settings the reference reads
 * on this path (inc/settings.h:15-31, src/settings.c:35-59). */
typedef struct {
    int free_formation;     /* [Cablegen] free_formation            */
    int ignore_f;           /* [Cablegen] ignore_f                  */
    int prune;              /* [Generate] prune                     */
    long stsl;              /* [Generate] stsl                      */
    long ltc;               /* [Generate] ltc                       */
    long smallest_large;    /* [Generate] smallest_large            */
    int nox;                /* [Cablegen] nox                       */
    int score;              /* [Solve] score                        */
    int premove;            /* [Generate] premove: the initial boards are
                             * positions after the move, so the first layer
                             * is their canonical set (no move step); the
                             * win probability of the adjoint is then the
                             * mean of their values */
} cg_rules;

const char *cg_last_error(void);

/* ---------------- device-resident tablebase ---------------- */

typedef struct cg_tablebase cg_tablebase;

cg_tablebase *cg_tb_create(const cg_rules *rules);
void cg_tb_free(cg_tablebase *tb);

/* Generates M_s for s = start, start+2, ..., end from the (pre-move)
 * initial boards (host memory), keeping every layer in HBM.  Same recurrence
 * as generate() (src/generate.c:551-592). */
int cg_tb_generate(cg_tablebase *tb, const uint64_t *initial, size_t n, int start, int end);
/* As cg_tb_generate, but stops after the first layer that brings the total
 * above max_boards (0: no bound), so a table too large to keep resident is
 * counted up to that point instead of exhausting HBM; *last_sum (may be NULL)
 * receives the tile sum of the last layer generated. */
int cg_tb_generate_bounded(cg_tablebase *tb, const uint64_t *initial, size_t n, int start, int end,
                           size_t max_boards, int *last_sum);

/* Solves every stored layer with s >= end_solve from the top down, as
 * solve() does (src/solve.c:172-224).  winstates are the unexpanded boards
 * (host); their 8 images are added here.  p4 is the 4-spawn probability
 * (0.1 reproduces the reference).  keep_argmax stores the 2-bit argmax move
 * per (board, empty cell, spawn) needed by cg_tb_adjoint. */
int cg_tb_solve(cg_tablebase *tb, const uint64_t *winstates, size_t n_win, int end_solve,
                double p4, int keep_argmax);

/* Exact derivative of the initial boards' mean win probability with respect
 * to p4, by an upward occupation-measure sweep over the optimal policy
 * (first strict maximum over left, right, up, down, as the CLI's best()
 * picks moves, frontends/cli/main.c:174-189).  Needs cg_tb_solve(...,
 * keep_argmax=1) with the same p4.  D receives cg_tb_num_layers() per-layer
 * sensitivities (may be NULL). */
int cg_tb_adjoint(cg_tablebase *tb, const uint64_t *initial, size_t n, double *P, double *dP,
                  double *D);

int cg_tb_num_layers(const cg_tablebase *tb);
int cg_tb_start(const cg_tablebase *tb);
size_t cg_tb_layer_size(const cg_tablebase *tb, int idx);
/* total boards over all layers */
size_t cg_tb_total_boards(const cg_tablebase *tb);
/* copies layer idx (tile sum start + 2 idx) to host; either pointer may be NULL */
int cg_tb_copy_layer(const cg_tablebase *tb, int idx, uint64_t *keys, double *values);
/* device pointers of layer idx (values NULL before solve) */
int cg_tb_layer_device(const cg_tablebase *tb, int idx, const uint64_t **keys,
                       const double **values);
/* writes %d.boards / %d.tables files in the reference's formats
 * (src/generate.c:74-87, src/solve.c:29-45); either fmt may be NULL */
int cg_tb_write_files(const cg_tablebase *tb, const char *boards_fmt, const char *tables_fmt);

/* ---------------- batched tables ---------------- */

/* Many small tables generated and solved together: every layer step runs one
 * set of kernels for all tables (keys carry the table id), and the solve's
 * downward sweep is one kernel per step, captured in a hipGraph.  Tables
 * share the rules (free_formation must be false) and may differ in
 * formation (at most 12 free cells), initial boards, tile-sum range and
 * winstates; each is identical to the same table run alone. */
typedef struct cg_batch cg_batch;
cg_batch *cg_batch_create(const cg_rules *rules);
void cg_batch_free(cg_batch *b);
/* adds a table (initial boards and unexpanded winstates in host memory);
 * returns its index or -1 */
int cg_batch_add(cg_batch *b, const uint64_t *initial, size_t n, int start, int end,
                 const uint64_t *winstates, size_t n_win);
int cg_batch_generate(cg_batch *b);
/* use_graph: replay the captured sweep (captured on first use per p4) */
int cg_batch_solve(cg_batch *b, double p4, int use_graph);
int cg_batch_num_tables(const cg_batch *b);
int cg_batch_num_layers(const cg_batch *b, int table);
size_t cg_batch_layer_size(const cg_batch *b, int table, int layer);
size_t cg_batch_total_boards(const cg_batch *b);
int cg_batch_copy_layer(const cg_batch *b, int table, int layer, uint64_t *keys, double *values);

/* ---------------- device primitives (device pointers) ---------------- */

/* move + lock + prune + canonicalise of n boards into out (capacity >= 4n);
 * *n_out receives the number of candidates (unsorted). */
int cg_dev_move(const cg_rules *rules, const uint64_t *d_in, size_t n, uint64_t *d_out,
                size_t *n_out);
/* spawn expansion: out2/out4 (capacity >= 16n) receive canon(b[x<-1]) and
 * canon(b[x<-2]) for every empty cell; *n_out the count. */
int cg_dev_spawn(const uint64_t *d_in, size_t n, uint64_t *d_out2, uint64_t *d_out4,
                 size_t *n_out);
/* sorts d_keys[0..n) and writes the unique keys to d_out (capacity n);
 * d_keys is clobbered.  *n_unique receives the count. */
int cg_dev_sort_unique(uint64_t *d_keys, size_t n, uint64_t *d_out, size_t *n_unique);
/* as cg_dev_sort_unique, with the largest value of every 4-bit cell stated
 * in cell_max (the nibble at the cell's position; all ones: no bound): the
 * sort may then code pairs of varying cells as a * base + b.  A key above
 * its bound makes the call fail. */
int cg_dev_sort_unique_bounded(uint64_t *d_keys, size_t n, uint64_t *d_out, size_t *n_unique,
                               uint64_t cell_max);
/* Host-side view of the sort's cell-pair code (no GPU work): for the cells
 * with a bit in vary, bounded by cell_max as above, codes[i] is the code of
 * keys[i] and back[i] the key decoded from it.  Returns the code's width in
 * bits, or -1 when vary is 0. */
int cg_sort_code_probe(const uint64_t *keys, size_t n, uint64_t vary, uint64_t cell_max,
                       uint64_t *codes, uint64_t *back);
/* Host-side view of the canonical form through a formation's stabiliser (no
 * GPU work): with the selection made for boards whose 0xF cells are those of
 * e (free_formation ff), out[i] is the canonical form of b[i] taken as a
 * board of orientation o (0 as stored, 1 turned clockwise) -- always equal
 * to the full max of 8 images; sel[0..1] receive the images each orientation
 * keeps (0xff: all, the formation does not decide). */
int cg_canon_sel_probe(uint64_t e, int ff, const uint64_t *b, size_t n, int o, uint64_t *out,
                       uint32_t *sel);
/* Host-side view of a layer's index plan (no GPU work): the directory
 * depth in bits for a sorted unique layer of n keys with the given OR / AND.
 * Fails (-1, with a message) for n >= 2^32, which the 32-bit directory
 * cannot index -- the same check every layer build makes. */
int cg_layer_index_probe(size_t n, uint64_t key_or, uint64_t key_and, int *bits);
/* exact-match lookup of nq queries into a sorted unique layer (keys, values);
 * misses give 0 (lookup, src/solve.c:108-164). */
int cg_dev_lookup(const uint64_t *d_keys, const double *d_vals, size_t n, const uint64_t *d_q,
                  size_t nq, double *d_out, int canonicalize);
/* cg_dev_lookup without allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * directory and its build buffers live in one fixed device array of about
 * 12.6 MB per GPU (held until the process ends), with a directory of at most
 * 2^20 buckets: a layer above ~2^18 keys gets a coarser directory than
 * cg_dev_lookup's (longer searches inside a bucket) and the same results.
 * n >= 2^32 fails with a message, as in cg_dev_lookup.  Runs on the null
 * stream and returns after it has synchronised (one 24-byte read-back of
 * the layer's OR / AND).  Calls that overlap on one GPU share the array and
 * race. */
int cg_dev_lookup_noalloc(const uint64_t *d_keys, const double *d_vals, size_t n, const uint64_t *d_q,
                          size_t nq, double *d_out, int canonicalize);
/* solve one layer against two solved successor layers (device pointers);
 * win_images are the already-expanded winstates (host). */
int cg_dev_solve_layer(const cg_rules *rules, const uint64_t *d_M, size_t n, double *d_V,
                       const uint64_t *d_K2, const double *d_V2, size_t n2,
                       const uint64_t *d_K4, const double *d_V4, size_t n4,
                       const uint64_t *win_images, size_t n_images, double p4);
/* n random boards: cells in fmask are 0xF; every other cell is empty with
 * probability p_empty, else exponent k in 1..kmax with P(k) ~ 2^-k
 * (uniform = 0), or uniform in 0..kmax (uniform = 1); drawn from a
 * counter-based splitmix64 stream keyed by (seed, board index). */
int cg_dev_random_boards(uint64_t *d_out, size_t n, uint64_t seed, uint64_t fmask, double p_empty,
                         int kmax, int uniform);
/* in place: d_keys[i] = canonical form (max of the 8 symmetry images) */
int cg_dev_canon(uint64_t *d_keys, size_t n);
/* synchronises the library's stream */
int cg_dev_sync(void);

/* ---------------- tuning knobs and the autotuner ---------------- */

/* Knobs (value 0 returns a forced field to the tuned table):
 *   "sort_bits"  radix digit width of every sort: 8, 9 or 10 bits (a pass
 *                block has one thread per digit);
 *   "sort_kpt"   keys per thread of the passes: 32 or 16 (8 and 9 bits, one
 *                thread per digit), 8 (8 and 10 bits) or 7 (9 bits) with 1024
 *                threads; a tile is threads * kpt keys;
 *   "sort_lbw"   look-back window of the passes: 16 or 32 tiles;
 *   "dir_delta"  prefix-bucket directory depth: bits = floor(log2 n) + delta
 *                (-4..3), clamped to the layer's varying bits and 30;
 *   "spawn_estimate"  0: spawn buffers sized for about twice the previous
 *                layer, a step that overflows them is repeated with exact
 *                buffers; 1: sized for half the previous layer, so that
 *                growing layers take the repeat (for tests);
 *   "carry_sort" 0: a layer step's spawn-4 boards are moved as written
 *                (duplicates included) two steps later; 1: they are sorted
 *                and deduplicated first (the carry sort);
 *   "solve_waves"  the solve kernel's registers fitted to 6 (default) or
 *                5 waves per SIMD; 0: the compiler's choice (4 waves);
 *   "prof_count"  1 (default): a solve timed by the profiler also counts its
 *                lookups (the algorithmic bytes) in a kernel instance that
 *                keeps the counters; 0: timed without counting (the
 *                production kernel);
 *   "spawn_blocks"  blocks of a spawn launch whose 4-spawns are moved
 *                unsorted (64..65536, default 1024);
 *   "carry_stream"  with carry_sort 1 (and in batches): 1: the carry sort
 *                runs on a second stream beside the next step's moves; 0: on
 *                the main stream after the spawns;
 *   "cell_code"  1: a generation's sorts code pairs of varying cells as
 *                a * base + b when the layer's tile sum bounds the cell
 *                values and that saves a radix pass; 0: bit packing only;
 *   "canon_stab" 1: the generation's move and spawn kernels take the
 *                canonical form through the formation's stabiliser (the
 *                images that put the formation at its canonical place; the
 *                full max of 8 for any other board); 0: always the max of 8;
 *   "sample_sort"  1: sorts of up to 16384 buckets of 7000 keys (1.15e8)
 *                partition into LDS-sized buckets by splitters from a
 *                regular sample and deduplicate and sort each bucket in LDS
 *                (a sort of at most 8192 keys is one block); 0: every sort
 *                takes the LSD radix path;
 *   "sample_bucket"  mean keys per bucket up to 2048 buckets (1000..16384);
 *   "sample_cap" keys a bucket may hold (0: 131072, its 8 regions); a bucket
 *                above it makes the library repeat that sort on the LSD path
 *                (lowered only to exercise the repeat);
 *   "sample_bmax"  most buckets of a sort: 16384 (sorts up to 1.15e8 keys),
 *                8192 (up to 5.7e7) or 2048 (up to 1.4e7); larger sorts
 *                take the LSD path;
 *   "sample_probes"  longest hash-set probe of a bucket that one block can
 *                also sort whole (0: 64); past it the bucket is sorted by a
 *                bitonic network, equal runs dropped;
 *   "part_blocks"  sample sort partition blocks: at most this many, each
 *                striding over the 8192-key tiles (default 512); 0: one
 *                block per tile;
 *   "part_prefetch"  1: a partition block loads its next tile's keys while
 *                the current tile reserves its places (default); 0: at the
 *                start of each tile;
 *   "bucket_bitonic"  1: buckets that fit one block are sorted whole by the
 *                bitonic network first; 0: through the LDS hash set first;
 *   "distinct_bitonic"  a bucket's distinct values are sorted by a bitonic
 *                network when there are at most this many (0: never; at
 *                most 16384) and they span at least "distinct_bitonic_bits"
 *                bits (0..64), else by 8-bit LSD passes;
 *   "prune_in_sort"  with prune, 1: the moved boards of a layer step are
 *                tested by the sort of its candidates, once per distinct
 *                board; 0: each moved board is tested in the move kernel;
 *   "dir_deferred"  1: the layer directories of a generation are built
 *                after its last step; 0: after each step;
 *   "one_sync"   1: a layer step reads its counters from the device once
 *                (the candidate sort sized on the device); 0: twice;
 *   "batch_graph"  1: each batched layer step is captured and launched as
 *                one graph; 0: its launches are enqueued one by one;
 *   "builtin"    built-in defaults: 8-bit digits, 8192-key tiles, window 16
 *                for every entry, no forced field, delta 0, spawn_estimate 0,
 *                cell_code 1, carry_stream 1, canon_stab 1, sample_sort 1
 *                with 5000 keys per bucket, capacity 131072, 16384 buckets,
 *                probes 64, part_blocks 512, part_prefetch 1, bucket_bitonic 0, distinct_bitonic 2048 with 9
 *                bits, prune_in_sort 1, dir_deferred 1, carry_sort 0, spawn_blocks
 *                1024, solve_waves 6, prof_count 1, one_sync 1,
 *                batch_graph 1;
 *   "reset"      the state after initialisation: the built-in defaults, then
 *                the table loaded at initialisation if there was one.
 * Returns 0, or -1 for an unknown knob or value.  cg_tune_get also reads
 * "cell_code_sorts": the number of sorts that used the cell-pair code,
 * "sample_sorts": the number of sample sorts, "sample_repeats": the number
 * of sample sorts repeated on the LSD path after an oversize bucket,
 * "graph_steps": the number of batched layer steps launched as a graph, and
 * "graph_instantiations": the number of graphs instantiated for them (a
 * step whose graph has the shape of the previous one updates it in place). */
int cg_tune_set(const char *knob, int value);
int cg_tune_get(const char *knob);
/* the sorts read their configuration from a table indexed by size class
 * (0: n < 2^20, 1: n < 2^23, 2: larger) and number of varying key bits
 * (0..64); these set one entry and report the configuration a sort of n
 * keys with nbits varying bits would use */
int cg_tune_sort_entry(int size_class, int nbits, int bits, int kpt, int lbw);
int cg_tune_sort_query(size_t n, int nbits, int *bits, int *kpt, int *lbw);
/* Autotuner: times every sort configuration at every pass count on
 * board-like keys for each size class and fills the table with the fastest
 * (a configuration replaces the default only when at least 3% faster), then
 * times directory depths on a lookup workload and sets "dir_delta".  effort
 * 0 takes about a second, 1 uses larger samples.  report (may be NULL)
 * receives the measurements and picks as JSON.  On failure the defaults are
 * restored. */
int cg_autotune(int effort, char *report, size_t cap);
/* Installs the picks of a saved autotuner report (JSON as cg_autotune
 * writes it).  The library does this at initialisation with tuned/
 * mi355x.json next to libcablegen.so, or with the file the environment
 * variable CABLEGEN_TUNING names (an empty value keeps the built-in
 * defaults).  Returns 0, or -1 for an unreadable or malformed file. */
int cg_tune_load(const char *path);

/* ---------------- kernel timers ---------------- */

/* When enabled, the library brackets the launches of its main kernels with
 * HIP events on the stream they run on and accumulates the time and the
 * algorithmic bytes of each launch. */
void cg_prof_enable(int on);
void cg_prof_reset(void);
/* kernel: "move", "spawn", "sort_hist", "sort_pass", "unique", "sort_split",
 * "sort_part", "sort_local", "solve", "adjoint", "lookup".  Returns the
 * number of launches. */
long cg_prof_query(const char *kernel, double *total_ms, double *total_bytes);
/* times only the launches of one kernel (names as for cg_prof_query; NULL or
 * "" times every kernel; "sort" times every kernel of the sort stage; a name
 * followed by "@main" times those launches on the main stream only): each
 * timed launch records two events */
int cg_prof_select(const char *kernel);
/* the same for the launches of one stream: 0 the library's main stream, 1
 * the second stream, which runs the carry sort beside the next step's moves */
long cg_prof_query_stream(const char *kernel, int stream, double *total_ms, double *total_bytes);

/* ---------------- settings used by the drop-in API ---------------- */

/* Copies the library's current settings singleton into rules. */
void cg_rules_from_settings(cg_rules *rules);

#ifdef __cplusplus
}
#endif
#endif
