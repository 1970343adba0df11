/*
 * engine.h -- thin internal C ABI between the host C library (api.c) and the
 * gfx950 HIP engine (engine.hip, kernels.hip).  Plain pointers and sizes.
 * Functions return 0 on success and a negative code on a device error (the
 * message is available from cge_last_error()).
 */
#ifndef CABLEGEN_ENGINE_H
#define CABLEGEN_ENGINE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Rules that shape the layers and the values.  Mirrors the settings the
 * reference consumes on this path (settings.h:15-31). */
typedef struct {
    int free_formation;
    int ignore_f;
    int prune;
    long stsl;
    long ltc;
    long smallest_large;
    int nox;
    int score;
    int premove;   /* the initial boards are taken as the first layer's positions */
} cge_rules;

const char *cge_last_error(void);

/* Builds the left-move row table for (free_formation, ignore_f) into tab
 * (65536 entries); right moves are derived by nibble reversal. */
void cge_build_row_table(int ignore_f, uint16_t *tab);

/* pinned (page-locked) host memory for layer transfers; NULL on failure */
void *cge_host_alloc(size_t bytes);
void cge_host_free(void *p);

/* ---- streaming generation (drives generate()) ---- */
int cge_gen_begin(const cge_rules *r, const uint64_t *initial, size_t n);
int cge_gen_step(size_t *n_layer);                 /* M_s for the next s */
int cge_gen_fetch(uint64_t *host, size_t n);       /* copy M_s to host */
void cge_gen_end(void);

/* ---- streaming solve (drives solve()) ---- */
int cge_solve_begin(const cge_rules *r, const uint64_t *win_images, size_t n_images, double p4);
/* Solves the layer `keys` (sorted unique, host memory) against the two
 * previously solved layers and rotates them. */
int cge_solve_step(const uint64_t *keys, size_t n, double *values);
void cge_solve_end(void);

#ifdef __cplusplus
}
#endif
#endif
