/*
 * oracle.h -- CPU restatement of cablegen's generate/solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker and the timed
 * CPU baseline.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product library (libcablegen.so) never
 * links, loads or calls it.
 *
 * Each function restates one piece of the reference (clepdn/cablegen) and
 * cites the file:line it follows.  Board encoding (api/cablegen.h:9-12):
 * tile x in [0,16) is the nibble at bits [60-4x, 63-4x]; row r holds tiles
 * 4r..4r+3; inside a 16-bit row value column 0 is the most significant nibble.
 */
#ifndef CABLEGEN_ORACLE_H
#define CABLEGEN_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int free_formation;      /* settings.c:37  [Cablegen] free_formation */
    int ignore_f;            /* settings.c:38  [Cablegen] ignore_f       */
    int cores;               /* settings.c:39  [Cablegen] cores          */
    int nox;                 /* settings.c:40  [Cablegen] nox            */
    int prune;               /* settings.c:52  [Generate] prune          */
    long stsl;               /* settings.c:49  [Generate] stsl           */
    long ltc;                /* settings.c:50  [Generate] ltc            */
    long smallest_large;     /* settings.c:51  [Generate] smallest_large */
    int score;               /* settings.c:57  [Solve] score             */
    int premove;             /* settings.c:45,72 [Generate] premove: its
                              * reference block (generate.c:571-579) moves
                              * the initial boards into a buffer nothing
                              * reads; defined here as "the initial boards
                              * are positions after the move": the first
                              * layer is their canonical set, the adjoint's
                              * P the mean of their values */
} orc_cfg;

/* --- board primitives (board.c) --- */
void     orc_build_lut(int free_formation, int ignore_f,
                       uint16_t move_lut[2][65536], uint8_t lock_lut[2][65536]);
uint64_t orc_rotate_cw(uint64_t b);
uint64_t orc_flip(uint64_t b);
uint64_t orc_canon(uint64_t b);
void     orc_all_rots(uint64_t b, uint64_t out[8]);
int      orc_get_sum(uint64_t b);
/* movedir_unstable semantics: returns validity, *b is the (possibly rotated)
 * moved board on success.  Uses the LUT built for cfg. */
int      orc_move(const orc_cfg *cfg, uint64_t *b, int dir);
int      orc_prune(uint64_t b, long stsl, long ltc, long smallest_large);
int      orc_cmpbrd(uint64_t b, uint64_t w);
double   orc_lookup(uint64_t key, const uint64_t *keys, const double *vals, size_t n,
                    int canonicalize);

/* --- sort / merge (array.c, fluxsort.h) --- */
void     orc_sort_u64(uint64_t *a, size_t n);
/* k-way merge-dedupe of sorted arrays; returns a malloc'd sorted unique array */
uint64_t *orc_kway_dedupe(uint64_t *const *arrs, const size_t *sizes, size_t k, size_t *out_n);

/* --- generation (generate.c) --- */
/* One layer step: from pre-move layer P (any order) and the carried spawn-4
 * set C (sorted unique), produce M (sorted unique post-move layer), P2 (the
 * next pre-move layer) and C2 (the new carry).  All outputs are malloc'd. */
int orc_gen_layer(const orc_cfg *cfg, const uint64_t *P, size_t nP,
                  const uint64_t *C, size_t nC,
                  uint64_t **M, size_t *nM, uint64_t **P2, size_t *nP2,
                  uint64_t **C2, size_t *nC2);
/* The spawn sets of a post-move layer: S2 = uniq{canon(m[x<-1])},
 * S4 = uniq{canon(m[x<-2])} over empty cells (generate.c:177-199), with the
 * same threads/sort/merge as a generation step.  Outputs are malloc'd. */
int orc_spawn_sets(const orc_cfg *cfg, const uint64_t *M, size_t nM,
                   uint64_t **S2, size_t *n2, uint64_t **S4, size_t *n4);
/* Full generation s = start..end step 2; stores every M_s.  layers[i] is the
 * layer at tile sum start + 2i; nlayers = (end-start)/2 + 1. */
int orc_generate(const orc_cfg *cfg, const uint64_t *initial, size_t n_init,
                 int start, int end, uint64_t ***layers, size_t **sizes, int *nlayers);

/* --- solve (solve.c) --- */
/* Solve one layer: V[i] for M[i], given the solved layers s+2 (T2) and s+4
 * (T4).  W holds the winstates already expanded to their 8 images.
 * p4 generalises the 0.9/0.1 spawn split; p4 = 0.1 reproduces it exactly. */
int orc_solve_layer(const orc_cfg *cfg, const uint64_t *M, size_t nM,
                    const uint64_t *K2, const double *V2, size_t n2,
                    const uint64_t *K4, const double *V4, size_t n4,
                    const uint64_t *W, size_t nW, double p4, double *V);
/* Full solve over stored layers (layers[i] at start_gen + 2i), from s = end_gen
 * down to end_solve (s != 0).  values[i] receives a malloc'd array for every
 * layer in range (NULL where the layer is outside the solved range). */
int orc_solve(const orc_cfg *cfg, uint64_t **layers, const size_t *sizes, int nlayers,
              int start_gen, int end_gen, int end_solve,
              const uint64_t *winstates, size_t n_win, double p4, double **values);
size_t orc_expand_winstates(const uint64_t *ws, size_t n, uint64_t *out /* 8n */);

/* --- adjoint (no reference counterpart) --- */
/* Occupation-measure sweep from the initial boards (mean over them) giving
 * dP/dp4 and per-layer sensitivities D[i] for layer start_gen + 2i.
 * *P_out receives the mean win probability of the initial boards. */
int orc_adjoint(const orc_cfg *cfg, uint64_t **layers, const size_t *sizes, int nlayers,
                int start_gen, double **values, const uint64_t *W, size_t nW,
                const uint64_t *initial, size_t n_init, double p4,
                double *P_out, double *dP_out, double *D);

/* FNV-1a 64 over the u64 words of an array (layer manifests). */
uint64_t orc_hash_u64(const uint64_t *a, size_t n);
void     orc_free(void *p);

#ifdef __cplusplus
}
#endif
#endif
