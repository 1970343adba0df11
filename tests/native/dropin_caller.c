/*
 * A C caller of the drop-in boundary (include/cablegen.h), linked to
 * libcablegen.so, written the way the reference's CLI drives the library:
 *
 *   gensolve CONF      change_config(CONF), then generate() from
 *                      Generate.initial to Generate.end into Generate.dir
 *                      and solve() from there down to Solve.end into
 *                      Solve.dir (frontends/cli/main.c:79-139: equal tile
 *                      sums required, start = the initial boards' sum)
 *   lookup CONF HEX S  change_config(CONF), then read_table() of the board's
 *                      sum, sum + 2 and sum + 4 from Solve.dir,
 *                      then, for every empty cell and spawn, the best move of
 *                      best() (main.c:174-189: the first strictly larger
 *                      value wins, starting from left at 0); S = 1 treats
 *                      HEX as a board with its spawn on it (lookup_spawn,
 *                      main.c:191-265)
 *   explore FILE       every board of a %d.tables file with its value
 *                      (main.c:275-287)
 *
 * Output lines follow the reference's formats ("Best move: %s (%.10lf)",
 * "Board(%0.20lf):").  Exits non-zero on a usage error or a mixed-sum
 * initial file.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "cablegen.h"

static const char *dir_name(dir d)
{
    static const char *n[4] = { "left", "right", "up", "down" };
    return n[(int)d & 3];
}

/* the first strict maximum over left, right, up, down (main.c:174-189) */
static dir best_move(uint64_t board, table *t, double *p)
{
    dir bd = left;
    double bp = 0.0;
    for (int d = left; d <= down; d++) {
        uint64_t b = board;
        if (!movedir(&b, (dir)d)) continue;
        double v = lookup(b, t, true);
        if (v > bp) {
            bd = (dir)d;
            bp = v;
        }
    }
    *p = bp;
    return bd;
}

static char *join(const char *a, const char *b)
{
    char *s = malloc(strlen(a) + strlen(b) + 1);
    strcpy(s, a);
    strcat(s, b);
    return s;
}

static int cmd_gensolve(char *conf)
{
    change_config(conf);
    min_settings_t st = *get_settings_min();
    static_arr_info init = read_boards(st.initial);
    if (!init.valid || init.size < 1) {
        fprintf(stderr, "no boards in %s\n", st.initial);
        return 1;
    }
    for (size_t i = 1; i < init.size; i++)
        if (get_sum(init.bp[i]) != get_sum(init.bp[0])) {
            fprintf(stderr, "initial boards have mixed sums\n");
            return 1;
        }
    char *bfmt = join(st.bdir, "%d.boards"), *tfmt = join(st.tdir, "%d.tables");
    generate(get_sum(init.bp[0]), (int)st.end_gen, bfmt, &init);
    static_arr_info ws = read_boards(st.winstates);
    if (!ws.valid || ws.size < 1) {
        fprintf(stderr, "no winstates in %s\n", st.winstates);
        return 1;
    }
    solve((unsigned)st.end_gen, (unsigned)st.end_solve, bfmt, tfmt, &ws);
    destroy_sarr(&init);
    destroy_sarr(&ws);
    free(bfmt);
    free(tfmt);
    return 0;
}

static int cmd_lookup(char *conf, const char *hex, int spawn)
{
    change_config(conf);
    const char *tdir = get_settings_min()->tdir;
    uint64_t board = strtoull(hex, NULL, 16), original = board;
    if (spawn)   /* the first cell holding a 2 or a 4 (or empty) is the spawn */
        for (int i = 0; i < 16; i++)
            if (GET_TILE(board, i) <= 2) {
                SET_TILE(board, i, 0);
                break;
            }
    generate_lut();
    int sum = get_sum(board);
    table t[3];
    char name[4096];
    for (int k = 0; k < 3; k++) {
        snprintf(name, sizeof name, "%s%d.tables", tdir, sum + 2 * k);
        read_table(&t[k], name);
    }
    if (!spawn) printf("Board(%0.10lf):\n", lookup(board, &t[0], true));
    for (int i = 0; i < 16; i++) {
        if (GET_TILE(board, i)) continue;
        for (int v = 1; v <= 2; v++) {
            SET_TILE(board, i, (uint64_t)v);
            double p;
            dir d = best_move(board, &t[v], &p);
            if (!spawn || original == board) printf("Best move: %s (%.10lf)\n", dir_name(d), p);
        }
        SET_TILE(board, i, 0);
    }
    for (int k = 0; k < 3; k++) {
        free(t[k].key.bp);
        free(t[k].value.bp);
    }
    return 0;
}

static int cmd_explore(const char *path)
{
    table t;
    read_table(&t, path);
    for (size_t i = 0; i < t.key.size; i++) {
        printf("Board(%0.20lf):\n", pun_uint64(t.value.bp[i]));
        output_board(t.key.bp[i]);
    }
    free(t.key.bp);
    free(t.value.bp);
    return 0;
}

int main(int argc, char **argv)
{
    if (argc >= 3 && !strcmp(argv[1], "gensolve")) return cmd_gensolve(argv[2]);
    if (argc >= 5 && !strcmp(argv[1], "lookup")) return cmd_lookup(argv[2], argv[3], atoi(argv[4]));
    if (argc >= 3 && !strcmp(argv[1], "explore")) return cmd_explore(argv[2]);
    fprintf(stderr, "usage: %s gensolve CONF | lookup CONF HEX SPAWN | explore FILE\n", argv[0]);
    return 2;
}
