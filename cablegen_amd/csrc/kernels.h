/*
 * kernels.h -- launchers for the gfx950 kernels (kernels.hip).  Host-callable
 * C++; every launcher enqueues on `st` and does not synchronise unless noted.
 */
#ifndef CABLEGEN_KERNELS_H
#define CABLEGEN_KERNELS_H
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

namespace cg {

struct MoveRules {
    int ff;          /* free_formation */
    int prune;
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

/* A solved (or generated) layer with its lookup index. */
struct LayerView {
    const uint64_t *keys;
    const double *vals;
    size_t n;
    const uint64_t *split;  /* ns splitters: split[i] = keys[i*n/ns] */
    int ns;
};

/* kernel-timer brackets provided by the engine (kind: 2 = radix histogram,
 * 3 = radix scatter); a negative handle means timing is off */
int prof_hook_begin(int kind);
void prof_hook_end(int h, double bytes);

/* first device error seen by a launcher's runtime calls since the last call */
hipError_t kernels_take_error();

/* scratch sizes for a sort of n keys */
size_t sort_scratch_bytes(size_t n);

/* move + lock + prune + canonicalise; appends valid canonical successors to
 * out (capacity 4n) and adds their number to *count (device counter). */
void launch_move(hipStream_t st, const uint64_t *in, size_t n, uint64_t *out,
                 unsigned long long *count, const uint16_t *lut, MoveRules r);

/* sum of empty cells over a layer into *count (device counter) */
void launch_count_empties(hipStream_t st, const uint64_t *in, size_t n, unsigned long long *count);

/* spawn expansion: for each board and empty cell x (ascending), canon(b[x<-1])
 * into out2 and canon(b[x<-2]) into out4 at the same positions; *count
 * receives the number written. */
void launch_spawn(hipStream_t st, const uint64_t *in, size_t n, uint64_t *out2, uint64_t *out4,
                  unsigned long long *count);

/* Sort + unique in two steps.  sort_unique_begin sorts keys[0..n) ascending
 * (LSD radix over the varying bits only, ping-ponging with tmp), counts the
 * unique keys and synchronises st; sort_unique_finish writes the unique keys
 * to dst (capacity >= n_unique).  `scratch` must hold sort_scratch_bytes(n)
 * and must not be reused before finish has been enqueued. */
struct SortState {
    const uint64_t *sorted;
    size_t n;
    size_t n_unique;
    const uint64_t *chunk_off;
};
SortState sort_unique_begin(hipStream_t st, uint64_t *keys, uint64_t *tmp, size_t n, void *scratch);
void sort_unique_finish(hipStream_t st, const SortState &ss, uint64_t *dst);

/* builds ns splitters for a sorted layer */
void launch_build_split(hipStream_t st, const uint64_t *keys, size_t n, uint64_t *split, int ns);

/* solve one layer: V (and optionally the 2-bit argmax word per board) */
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
                         const uint16_t *lut, int ff, double *P);

/* batched exact-match lookups (microbench / API): out[i] = value or 0 */
void launch_lookup(hipStream_t st, LayerView t, const uint64_t *q, size_t nq, double *out,
                   int canonicalize);

} // namespace cg
#endif
