/*
 * kernels.h -- launchers for the gfx950 kernels (kernels.hip).  Host-callable
 * C++; every launcher enqueues on `st` and does not synchronise unless noted.
 */
#ifndef CABLEGEN_KERNELS_H
#define CABLEGEN_KERNELS_H
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>
#include "board.h"

namespace cg {

struct MoveRules {
    int ff;          /* free_formation */
    int prune;
    /* with prune: 1 tests every input board in the move kernel; 0 when the
     * inputs are spawns that the spawn kernel already tested */
    int prune_inputs;
    /* with prune: 1 tests every moved board in the move kernel; 0 leaves
     * the moved boards to the sort of the candidates (KeyPrune), which tests
     * each distinct board once */
    int prune_moves;
    long stsl, ltc, sl;
    cg_canon_sel_t cs;   /* canonical form through the formation's stabiliser */
};

/* prune_board (generate.c:98-131) applied by a sort whose keys are boards:
 * distinct keys that fail the test are left out of its output and count.
 * The test depends only on the multiset of cell values, so it gives the
 * same set on canonical forms and after deduplication. */
struct KeyPrune {
    int on;
    long stsl, ltc, sl;
};

/* Winstate test as used by the solve: `nox` (without score) makes "holds the
 * nibble nox" a win; score is never satisfied; otherwise any image i with
 * (b & mask[i]) == val[i] wins. */
struct WinRules {
    const uint64_t *mask;
    const uint64_t *val;
    int n;
    int nox;
    int score;
};

/* Order-preserving packing of the varying key bits: up to 4 runs of bits,
 * most significant first (more fragmented masks use one contiguous range). */
struct KeyXform {
    int nseg;
    int nbits;
    int lo[4];
    int len[4];
    uint64_t mask[4];
};
KeyXform make_xform(uint64_t vary);

/* The sort's key code.  cells == 0: the bit packing x.  cells == 1 (the
 * cell-pair code): the ncell varying cells (nibbles at bit cell_lo[i], most
 * significant first; cell i holds values below cbase[i]) are taken two at a
 * time, a pair (a, b) as a * cbase(b) + b in a field of pbits[p] bits (npair =
 * ncell / 2; an odd last cell in lbits bits), so two cells of values 0..10
 * take 7 bits instead of 8.  Both codes keep the keys' order: field order is
 * cell order, and a * cbase(b) + b orders pairs lexicographically. */
struct SortCode {
    KeyXform x;
    int cells, npair, ncell, lbits;
    int nbits;              /* bits of the code */
    uint64_t vmask;         /* the key bits the code covers */
    uint8_t cell_lo[16];
    uint8_t cbase[16];
    uint8_t pbits[8];
    uint32_t magic[8];      /* ceil(2^16 / cbase of the pair's second cell):
                               f / base = (f * magic) >> 16 for f < 256 */
};
/* bit packing of vary */
SortCode make_bit_code(uint64_t vary);
/* the cell-pair code of vary's cells, where the nibble of cell_max at each
 * cell's position is the cell's largest value (bit packing when vary is 0) */
SortCode make_pair_code(uint64_t vary, uint64_t cell_max);
/* cell_max with every cell bounded by v (0..15) */
constexpr uint64_t cells_bounded_by(int v) { return (uint64_t)(v & 15) * 0x1111111111111111ULL; }
constexpr uint64_t CELLS_UNBOUNDED = ~0ULL;

/* A sorted unique layer with its prefix-bucket index.  Keys of the layer
 * agree on the bits in cmask (value cval); the remaining bits are packed by
 * x and the top `bits` packed bits select a bucket: dir[b] is the index of
 * the first key whose bucket is >= b (dir has 2^bits + 1 entries).  With
 * bits ~ log2(n) a bucket holds one or two keys on average, so a lookup is
 * one directory read and a short search. */
struct LayerView {
    const uint64_t *keys;
    const double *vals;
    size_t n;
    const uint32_t *dir;
    int bits;
    int shift;        /* bucket = xform(q) >> shift */
    uint64_t cmask;
    uint64_t cval;
    KeyXform x;
};

/* Directory entries, and the positions the solve and the adjoint derive from
 * them, are 32-bit: a layer (or a batch step's concatenated layers) of
 * LAYER_INDEX_MAX boards or more cannot be indexed, and the launchers refuse
 * it (kernels_note_error) instead of wrapping. */
constexpr size_t LAYER_INDEX_MAX = (size_t)1 << 32;

/* index of a layer: directory sizing from n and the OR / AND of its keys */
struct LayerIndexPlan {
    KeyXform x;
    int bits, shift;
    uint64_t cmask, cval;
    size_t dir_entries;   /* 2^bits + 1 */
};
LayerIndexPlan plan_index(size_t n, uint64_t key_or, uint64_t key_and);
/* builds dir for a sorted unique layer; scratch holds dir_scratch_bytes(entries) */
size_t dir_scratch_bytes(size_t entries);
void launch_build_dir(hipStream_t st, const uint64_t *keys, size_t n, const LayerIndexPlan &p,
                      uint32_t *dir, void *scratch);

/* kernel-timer brackets provided by the engine (kind: 2 = radix histogram,
 * 3 = radix pass, 4 = unique, 8 = sample-sort splitters, 9 = its partition,
 * 10 = its bucket sort + dedupe), recorded on stream st; a negative handle
 * means timing is off */
int prof_hook_begin(int kind, hipStream_t st);
void prof_hook_end(int h, double bytes, hipStream_t st);

/* first device error seen by a launcher's runtime calls since the last call */
hipError_t kernels_take_error();

/* scratch sizes for a sort of n keys */
size_t sort_scratch_bytes(size_t n);

/* move + lock + prune + canonicalise; appends valid canonical successors to
 * out (capacity 4n), adds their number to stats[0] and ORs / ANDs them into
 * stats[1] / stats[2].  With dn, the number of boards is min(n, *dn) read on
 * the device (n bounds the grid); likewise for launch_spawn.  With dedupe, a
 * candidate produced twice within a 1024-board chunk is appended once (the
 * set of candidates is unchanged). */
void launch_move(hipStream_t st, const uint64_t *in, size_t n, const unsigned long long *dn,
                 uint64_t *out, unsigned long long *stats, const uint16_t *lut, MoveRules r,
                 int dedupe = 0);

/* sum of empty cells over a layer into *count (device counter) */
void launch_count_empties(hipStream_t st, const uint64_t *in, size_t n, unsigned long long *count);

/* spawn expansion: for each board and empty cell x (ascending), canon(b[x<-1])
 * into out2 and canon(b[x<-2]) into out4 at the same positions; stats[0]
 * receives the number of spawns, stats[3..4] the OR / AND of the out4 keys
 * (the carry sort's key bits; the out2 keys are only moved, never sorted, so
 * stats[1..2] are left as they are).  Only positions below cap are written: a
 * count above cap means the buffers were too small and the launch must be
 * repeated.  With kp.on, a 2-spawn that fails the prune test is written to
 * out2 as 0, a board with no tile, which has no legal move: the 2-spawns
 * are only moved, so they need no prune test of their own.  unsorted4: the
 * 4-spawns are moved unsorted too, so with kp.on a failing one is written
 * to out4 as 0 as well (a sort of them drops such boards itself, KeyPrune,
 * where zeros would pile up in one bucket), and their OR / AND is not
 * collected.  The test depends only on the multiset of cells, so it is
 * taken once per board and value, not once per empty cell. */
void launch_spawn(hipStream_t st, const uint64_t *in, size_t n, const unsigned long long *dn,
                  uint64_t *out2, uint64_t *out4, size_t cap, unsigned long long *stats,
                  cg_canon_sel_t cs = cg_canon_all(), KeyPrune kp = KeyPrune(), int unsorted4 = 0);
/* blocks of a spawn launch whose 4-spawns are moved unsorted (no OR / AND
 * atomics at the end of each block; default 1024; 512 otherwise) */
void spawn_set_blocks(int b);
int spawn_blocks();


/* Sorts keys[0..n) ascending and writes the unique keys to dst (capacity n),
 * entirely on the stream (no host synchronisation): onesweep LSD radix over
 * the bits set in `vary` (OR & ~AND of the keys; bits outside it must be
 * constant), then a chained-scan compaction.  The unique count lands in
 * *d_nuniq.  tmp holds n keys; scratch holds sort_scratch_bytes(n).  Sorts of
 * different contexts (ctx < SORT_CONTEXTS) may run concurrently on different
 * streams, each with its own tmp and scratch.  A
 * chained-scan wait that exceeds its budget sets *d_err (never cleared
 * here); the caller checks it after synchronising.  keys and tmp are
 * clobbered.  *prof_unique receives the timer handle of the compaction.
 * cell_max states, per 4-bit cell (the nibble of cell_max at the cell's
 * position), the largest value the keys hold there; the sort uses the
 * cell-pair code for the cells with a bit in `vary` when that needs fewer
 * passes.  A key that breaks the bound ORs SORT_ERR_CODE into *d_err. */
constexpr int SORT_CONTEXTS = 2;
constexpr unsigned SORT_ERR_CODE = 4u;
/* a bucket of the sample sort held more keys than one block sorts in LDS:
 * the output is incomplete and the sort must be repeated on the LSD path */
constexpr unsigned SORT_ERR_OVERSIZE = 8u;
/* part (capacity part_cap keys, may be NULL) holds the buckets of the
 * sample sort, which sorts of up to about 10^7 keys take when part has
 * sort_part_keys(n) keys (sorts of at most 16384 keys need none) and
 * allow_sample is set; it only reads keys, and a bucket above its capacity
 * ORs SORT_ERR_OVERSIZE into *d_err: the caller then repeats the sort with
 * allow_sample = false (the LSD path, which takes any size). */
/* sample: SORT_LSD (the LSD path), SORT_SAMPLE (the sample sort for the
 * sizes it takes, up to sort_sample_limit(SORT_SAMPLE) keys, with buckets
 * for keys that repeat as move candidates do) or SORT_SAMPLE_SPARSE (the
 * same with smaller buckets, for keys with few repeats, such as a step's
 * spawn-4 boards) */
constexpr int SORT_LSD = 0, SORT_SAMPLE = 1, SORT_SAMPLE_SPARSE = 2;
uint64_t *sort_unique(hipStream_t st, uint64_t *keys, uint64_t *tmp, size_t n, uint64_t vary,
                      void *scratch, uint64_t *dst, unsigned long long *d_nuniq, unsigned *d_err,
                      int *prof_unique, int ctx = 0, uint64_t cell_max = CELLS_UNBOUNDED,
                      uint64_t *part = nullptr, size_t part_cap = 0, int sample = SORT_SAMPLE,
                      KeyPrune kp = KeyPrune());
/* keys of the part buffer a sort of n keys needs (0: none) */
size_t sort_part_keys(size_t n, int sample = SORT_SAMPLE);
/* The sample sort of *dn keys, a count known only on the device (nmax bounds
 * it and sizes the grids and buffers: part holds sort_part_keys(nmax) keys,
 * scratch sort_scratch_bytes): the same result as sort_unique, with no host
 * synchronisation and no host decision that depends on the count.  Returns
 * -1 (nothing enqueued) when nmax is above sort_sample_dev_limit() or part
 * is too small.  A bucket above its capacity ORs SORT_ERR_OVERSIZE into *d_err;
 * the caller then repeats the sort with sort_unique(allow_sample = false). */
int sort_unique_dev(hipStream_t st, const uint64_t *keys, const unsigned long long *dn, size_t nmax,
                    void *scratch, uint64_t *dst, unsigned long long *d_nuniq, unsigned *d_err,
                    int *prof_unique, int ctx, uint64_t *part, size_t part_cap, KeyPrune kp = KeyPrune());
/* the largest sort the sample sort takes in a mode (0 when it is switched
 * off, or for SORT_LSD) */
size_t sort_sample_limit(int mode);
/* the largest bound sort_unique_dev takes (2048 buckets) */
size_t sort_sample_dev_limit();
/* Makes room for `entries` look-back words in context ctx and for
 * `launches` more launches before the epoch counter wraps, so that the
 * launches that follow clear nothing (no memset, e.g. inside a graph
 * capture). */
int sort_reserve(hipStream_t st, int ctx, size_t entries, unsigned launches);
/* the same for a sort_unique of n keys (part assumed large enough) */
int sort_prepare(hipStream_t st, int ctx, size_t n, uint64_t vary, uint64_t cell_max, int sample);
/* sample sort: on / off, mean keys per bucket, keys a bucket may hold
 * (lowered only to exercise the repeat; 0 restores the default, 131072: the
 * keys of a bucket's regions, whose distinct values must fit the LDS) */
void sort_set_sample(int on, int bucket_avg, int cap);
/* most buckets of a sort (2048 or 8192) and the longest hash-set probe of a
 * bucket the direct sort can take (0: the default) */
void sort_set_sample_limits(int bmax, int probes);
int sort_sample_bmax();
/* buckets that fit one block sorted whole by a bitonic network first (1),
 * or through the LDS hash set and an LSD sort of the distinct values, with
 * the network only when the set fills (0, the default) */
void sort_set_bucket_bitonic(int on);
int sort_bucket_bitonic();
void sort_set_part_blocks(int nb);
int sort_part_blocks();
void sort_set_part_prefetch(int on);
int sort_part_prefetch();
/* a bucket's distinct values sorted by a bitonic network instead of 8-bit
 * LSD passes when there are at most dmax of them (0: never) and they span
 * at least dnb bits */
void sort_set_distinct_bitonic(int dmax, int dnb);
int sort_distinct_bitonic_max();
int sort_distinct_bitonic_bits();
int sort_sample_probes();
int sort_sample_on();
int sort_sample_bucket();
int sort_sample_cap();
unsigned long long sort_sample_count();

/* Sort configurations: digit width (8, 9 or 10 bits; the pass block has one
 * thread per digit), keys per thread (tile = 2^bits * kpt keys) and look-back
 * window (16 or 32 tiles).  sort_config(i), i < sort_num_configs(), lists
 * the combinations that have kernels. */
struct SortCfg {
    int bits, kpt, lbw;
};
constexpr int SORT_SIZE_CLASSES = 3;   /* n < 2^20, n < 2^23, larger */
int sort_num_configs();
SortCfg sort_config(int i);
int sort_size_class(size_t n);
/* per (size class, varying key bits) table used by sort_unique */
int sort_table_set(int cls, int nbits, SortCfg c);
SortCfg sort_table_get(int cls, int nbits);
void sort_table_reset();
/* fields that are nonzero override the table for every sort */
int sort_force(int bits, int kpt, int lbw);
SortCfg sort_forced();
SortCfg sort_choose(size_t n, int nbits);
/* frees the look-back status array */
void sort_release();
/* whether bounded sorts may use the cell-pair code (default on), and the
 * number of sorts that used it */
void sort_set_cell_code(int on);
int sort_cell_code();
unsigned long long sort_code_count();

/* prefix-bucket directory depth: bits = floor(log2 n) + delta, clamped */
void set_dir_delta(int d);
int get_dir_delta();

/* records a launch error for kernels_take_error() */
void kernels_note_error(hipError_t e);

void launch_fill_u64(hipStream_t st, uint64_t *p, size_t n, uint64_t v);
/* hpub[lo..lo+n) = cnt[lo..lo+n), hpub[err_at + j] = err[j] for j = 0, 1
 * (hpub: mapped host memory; lo + n <= err_at, n <= 62) */
void launch_publish(hipStream_t st, const unsigned long long *cnt, int lo, int n, const unsigned *err,
                    unsigned long long *hpub, int err_at);
/* dst[i] = src[i] & mask */
void launch_copy_and_u64(hipStream_t st, uint64_t *dst, const uint64_t *src, size_t n, uint64_t mask);
/* stats[0] = 0 (count), stats[1,3] = 0 (OR), stats[2,4] = ~0 (AND); the
 * same for stats2 when given (two producers' statistics, one launch) */
void launch_init_stats(hipStream_t st, unsigned long long *stats, unsigned long long *stats2 = nullptr);
/* OR / AND of keys into stats[1] / stats[2] */
void launch_or_and(hipStream_t st, const uint64_t *k, size_t n, unsigned long long *stats);


/* solve one layer: V (and optionally the 2-bit argmax word per board) */
/* the solve kernel's registers fitted to 5 (the default) or 6 waves per
 * SIMD (0: the compiler's choice, 4 waves) */
void solve_set_waves(int w);
int solve_waves();
void launch_solve(hipStream_t st, const uint64_t *M, size_t n, double *V, uint64_t *argmax,
                  LayerView t2, LayerView t4, WinRules w, double p4, const uint16_t *lut, int ff,
                  unsigned long long *stats);

/* adjoint sweep of one layer (occupation push-forward and D_s) */
void launch_adjoint(hipStream_t st, const uint64_t *M, size_t n, const double *occ,
                    const uint64_t *argmax, LayerView t2, double *occ2, LayerView t4, double *occ4,
                    WinRules w, double p4, const uint16_t *lut, int ff, double *D);

/* initial boards: best move over layer t0; adds 1/n to occ0 of the argmax
 * successor and the mean win probability to *P */
void launch_adjoint_init(hipStream_t st, const uint64_t *init, size_t n, LayerView t0, double *occ0,
                         const uint16_t *lut, int ff, double *P, int premove = 0);

/* random boards under a formation mask (see k_random_boards) */
void launch_random_boards(hipStream_t st, uint64_t *out, size_t n, uint64_t seed, uint64_t fmask,
                          double p_empty, int kmax, int uniform);
/* a[i] = canonical form of a[i] */
void launch_canon(hipStream_t st, uint64_t *a, size_t n);

/* batched exact-match lookups (microbench / API): out[i] = value or 0 */
void launch_lookup(hipStream_t st, LayerView t, const uint64_t *q, size_t nq, double *out,
                   int canonicalize);

/* ---- batched tables ---- */
constexpr int BATCH_TID_SHIFT = 48;

/* per table: formation nibbles (0xF) of its canonical boards, free tiles in
 * ascending order, number of layer steps */
struct BatchTable {
    uint64_t fnib;
    uint8_t pos[16];
    int nfree;
    int nsteps;
    int runs;        /* free-cell bits form at most 4 runs: fx packs them (pext-like) */
    KeyXform fx;
};

struct BatchIndexPlan {
    KeyXform x;
    int shift;
    size_t dir_base;   /* first bucket of this table in the concatenated directory */
};

struct BSolveDesc {
    const uint64_t *M;
    size_t n;
    double *V;
    uint64_t *arg;
    LayerView t2, t4;
    WinRules win;
    size_t tile0;
};

/* composite-key move / spawn; err |= 2 if a board's formation is not where
 * the table's canonical form puts it */
void launch_bmove(hipStream_t st, const uint64_t *in, size_t n, const unsigned long long *dn,
                  uint64_t *out, unsigned long long *stats, const uint16_t *lut, MoveRules r,
                  const BatchTable *tabs, int step, unsigned *err);
void launch_bspawn(hipStream_t st, const uint64_t *in, size_t n, const unsigned long long *dn,
                   uint64_t *out2, uint64_t *out4, size_t cap, unsigned long long *stats,
                   const BatchTable *tabs, int step, unsigned *err);
/* composite -> boards, with per-table count / OR / AND (tstat[3 t + 0..2],
 * AND entries preset to all ones) */
/* per-table statistics (count, OR, AND) set to 0, 0, all ones */
void launch_tstat_init(hipStream_t st, unsigned long long *t, int ntab);
void launch_bdecode(hipStream_t st, const uint64_t *in, size_t n, const unsigned long long *dn,
                    uint64_t *out, const BatchTable *tabs, unsigned long long *tstat);
/* directories of every table of a step in one concatenated array; entries
 * are indices into the step's concatenated keys */
void launch_bdir(hipStream_t st, const uint64_t *keys, size_t n, const size_t *seg, int ntab,
                 const BatchIndexPlan *plan, size_t total_entries, uint32_t *dir, void *scratch);
void launch_bsolve(hipStream_t st, const BSolveDesc *desc, int ntab, size_t total_tiles, double p4,
                   const uint16_t *lut, int ff);

} // namespace cg
#endif
