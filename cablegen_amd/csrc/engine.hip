/*
 * engine.hip -- host side of the gfx950 engine: device state, workspace,
 * the per-layer generation step, the layer-by-layer solve and adjoint, the
 * device-resident tablebase, the streaming interface used by the drop-in
 * generate()/solve(), device primitives and kernel timers.
 *
 * One HIP stream (default priority) carries all work.  Layer arrays live in
 * a bump arena of large hipMalloc chunks (a C2-sized run keeps ~1e9 boards,
 * ~24 GB with values and argmax words, resident in HBM); per-step
 * temporaries live in workspace buffers that only grow.
 */
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <dlfcn.h>
#include <mutex>
#include <string>
#include <vector>
#include "board.h"
#include "engine.h"
#include "kernels.h"
#include "devutil.h"
#include "../../include/cablegen_gpu.h"

using namespace cg;

/* ------------------------------------------------------------------------ */
/* errors                                                                    */
/* ------------------------------------------------------------------------ */

static std::string g_err;

static int fail(const char *what, hipError_t e)
{
    char buf[512];
    snprintf(buf, sizeof buf, "%s: %s", what, hipGetErrorString(e));
    g_err = buf;
    return -1;
}

static int fail_msg(const char *what)
{
    g_err = what;
    return -1;
}

#define HC(x)                                      \
    do {                                           \
        hipError_t e_ = (x);                       \
        if (e_ != hipSuccess) return fail(#x, e_); \
    } while (0)

/* collects launch and runtime errors after a synchronisation point */
static int check_sync(hipStream_t st, const char *where)
{
    hipError_t e = hipStreamSynchronize(st);
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess) e = kernels_take_error();
    if (e != hipSuccess) return fail(where, e);
    return 0;
}

extern "C" const char *cge_last_error(void) { return g_err.c_str(); }
extern "C" const char *cg_last_error(void) { return g_err.c_str(); }

/* ------------------------------------------------------------------------ */
/* device state                                                              */
/* ------------------------------------------------------------------------ */

struct Buf {
    uint64_t *p = nullptr;
    size_t cap = 0;
};

struct Engine {
    bool init = false;
    hipStream_t st = nullptr;
    uint16_t *lut[2] = { nullptr, nullptr };     /* left-move table per ignore_f */
    /* device counters [CNT_N]: [0..2] a producer's count / OR / AND, [5..7]
     * sort counts (6: the carry C', 7: the layer M_s), [8..12] the spawn's
     * count / OR / AND pairs of a layer step */
    unsigned long long *d_cnt = nullptr;
    unsigned long long *h_cnt = nullptr;          /* host copy of the counters */
    /* error words: [0] the main stream's sorts and the batch's formation
     * flags, [1] the second stream's sorts */
    unsigned *d_err = nullptr;
    unsigned long long *h_pub = nullptr;          /* pinned, mapped: counters [0..CNT_N), flags [CNT_N..+2) */
    unsigned long long *hd_pub = nullptr;         /* its device address */
    Buf A, B, Cb;                                 /* candidate / sort buffers */
    /* spawn-4 buffers of the layer steps, used in turn: the pending carry
     * sort's input stays intact while the next step's spawns write the
     * other one, so a carry sort found oversize after them can be repeated */
    Buf S4[2];
    void *scratch = nullptr;
    size_t scratch_cap = 0;
    /* a second stream for the carry sort, which runs beside the next
     * step's moves: its own ping-pong buffer and scratch; ev2 orders the
     * main stream after it */
    hipStream_t st2 = nullptr;
    hipEvent_t ev2 = nullptr;
    hipEvent_t evf = nullptr;   /* orders the second stream after the main one (a fork) */
    Buf B2;
    void *scratch2 = nullptr;
    size_t scratch2_cap = 0;
    uint64_t *d_wmask = nullptr, *d_wval = nullptr;
    size_t wcap = 0;
    /* the sample sort's bucket buffers of the two lanes */
    Buf Pt, Pt2;
    /* a sample sort enqueued before the last counter read had an oversize
     * bucket: its output is incomplete and the sort must be repeated
     * (oversize: a sort reporting into either error word; oversize1: the
     * first word's, oversize2: the carry sorts' word).  Every read sets all
     * three afresh. */
    bool oversize = false, oversize1 = false, oversize2 = false;
};
constexpr int CNT_N = 24;
static Engine E;

static void tuning_load_default();
static std::string g_tune_path;   /* the table installed at initialisation, if any */

static int eng_init()
{
    if (E.init) return 0;
    HC(hipStreamCreate(&E.st));
    HC(hipStreamCreate(&E.st2));
    HC(hipEventCreateWithFlags(&E.ev2, hipEventDisableTiming));
    HC(hipEventCreateWithFlags(&E.evf, hipEventDisableTiming));
    HC(hipMalloc((void **)&E.d_cnt, CNT_N * sizeof(unsigned long long)));
    HC(hipMemset(E.d_cnt, 0, CNT_N * sizeof(unsigned long long)));
    E.h_cnt = (unsigned long long *)calloc(CNT_N, sizeof(unsigned long long));
    if (!E.h_cnt) return fail_msg("host allocation failed");
    HC(hipMalloc((void **)&E.d_err, 64));
    HC(hipMemset(E.d_err, 0, 64));
    HC(hipHostMalloc((void **)&E.h_pub, (CNT_N + 8) * sizeof(unsigned long long),
                     hipHostMallocMapped | hipHostMallocCoherent));
    memset(E.h_pub, 0, (CNT_N + 8) * sizeof(unsigned long long));
    HC(hipHostGetDevicePointer((void **)&E.hd_pub, E.h_pub, 0));
    tuning_load_default();
    E.init = true;
    return 0;
}

/* waits for both streams (work enqueued on them may still use a buffer
 * about to be freed) */
static int sync_streams()
{
    if (E.st) HC(hipStreamSynchronize(E.st));
    if (E.st2) HC(hipStreamSynchronize(E.st2));
    return 0;
}

static int ensure(Buf &b, size_t n)
{
    if (n <= b.cap && b.p) return 0;
    if (b.p && sync_streams()) return -1;
    if (b.p) HC(hipFree(b.p));
    b.p = nullptr;
    size_t c = std::max<size_t>(n + n / 4, 1024);
    HC(hipMalloc((void **)&b.p, c * sizeof(uint64_t)));
    b.cap = c;
    return 0;
}

static int ensure_scratch_bytes(size_t need)
{
    if (need <= E.scratch_cap && E.scratch) return 0;
    /* sorts or index builds enqueued on the streams may still use it */
    if (E.scratch && sync_streams()) return -1;
    if (E.scratch) HC(hipFree(E.scratch));
    E.scratch = nullptr;
    size_t c = need + need / 4 + 4096;
    HC(hipMalloc(&E.scratch, c));
    E.scratch_cap = c;
    return 0;
}

static int ensure_scratch(size_t n)
{
    size_t need = sort_scratch_bytes(n);
    if (need <= E.scratch_cap && E.scratch) return 0;
    if (E.scratch && sync_streams()) return -1;
    if (E.scratch) HC(hipFree(E.scratch));
    E.scratch = nullptr;
    size_t c = need + need / 4 + 4096;
    HC(hipMalloc(&E.scratch, c));
    E.scratch_cap = c;
    return 0;
}

static int ensure_scratch2(size_t n)
{
    size_t need = sort_scratch_bytes(n);
    if (need <= E.scratch2_cap && E.scratch2) return 0;
    if (E.scratch2 && sync_streams()) return -1;
    if (E.scratch2) HC(hipFree(E.scratch2));
    E.scratch2 = nullptr;
    size_t c = need + need / 4 + 4096;
    HC(hipMalloc(&E.scratch2, c));
    E.scratch2_cap = c;
    return 0;
}

static const uint16_t *get_lut(int ignore_f)
{
    int i = ignore_f ? 1 : 0;
    if (!E.lut[i]) {
        std::vector<uint16_t> h(65536);
        cge_build_row_table(i, h.data());
        if (hipMalloc((void **)&E.lut[i], 65536 * sizeof(uint16_t)) != hipSuccess) return nullptr;
        if (hipMemcpy(E.lut[i], h.data(), 65536 * sizeof(uint16_t), hipMemcpyHostToDevice) !=
            hipSuccess)
            return nullptr;
    }
    return E.lut[i];
}

/* ------------------------------------------------------------------------ */
/* kernel timers                                                             */
/* ------------------------------------------------------------------------ */

/* the order matches the kinds the sort's launchers pass (kernels.h) */
enum { PK_MOVE, PK_SPAWN, PK_SORT_HIST, PK_SORT_PASS, PK_UNIQUE, PK_SOLVE, PK_ADJOINT, PK_LOOKUP,
       PK_SORT_SPLIT, PK_SORT_PART, PK_SORT_LOCAL, PK_N };
static const char *PK_NAMES[PK_N] = { "move", "spawn", "sort_hist", "sort_pass",
                                      "unique", "solve", "adjoint", "lookup",
                                      "sort_split", "sort_part", "sort_local" };

struct ProfRec {
    int kind;
    hipStream_t st;
    hipEvent_t a, b;
    double bytes;
    int stats_slot;      /* solve launches: lookup counters read at collection */
    double base_bytes;
};
static bool g_prof = false;
static int g_prof_only = -1;   /* >= 0: only this kernel kind is timed */
static bool g_prof_main = false;   /* only launches on the main stream are timed */
/* cg_prof_select("sort"): every kernel of the sort stage is timed */
static constexpr int PK_GROUP_SORT = 1000;
static bool is_sort_kind(int kind)
{
    return kind == PK_SORT_HIST || kind == PK_SORT_PASS || kind == PK_UNIQUE || kind == PK_SORT_SPLIT ||
           kind == PK_SORT_PART || kind == PK_SORT_LOCAL;
}
static std::vector<ProfRec> g_recs;
static std::vector<hipEvent_t> g_evpool;

static hipEvent_t ev_get()
{
    if (!g_evpool.empty()) {
        hipEvent_t e = g_evpool.back();
        g_evpool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}
static double g_ms[PK_N], g_bytes[PK_N];
static long g_launches[PK_N];
/* the same per stream: 0 the main stream, 1 the second (carry sort) stream */
static double g_ms_s[2][PK_N], g_bytes_s[2][PK_N];
static long g_launches_s[2][PK_N];

/* whether launches of this kernel kind are being timed */
/* a stream capture is open: kernel timers (event records) stay out of it */
static bool g_capturing = false;

static bool prof_wants(int kind)
{
    return g_prof && !g_capturing && (g_prof_only < 0 || kind == g_prof_only ||
                      (g_prof_only == PK_GROUP_SORT && is_sort_kind(kind)));
}

static int prof_begin(int kind, hipStream_t st = nullptr)
{
    if (!prof_wants(kind)) return -1;
    if (g_prof_main && st && st != E.st) return -1;
    ProfRec r;
    r.kind = kind;
    r.st = st ? st : E.st;
    r.bytes = 0;
    r.stats_slot = -1;
    r.base_bytes = 0;
    r.a = ev_get();
    r.b = ev_get();
    if (!r.a || !r.b) return -1;
    hipEventRecord(r.a, r.st);
    g_recs.push_back(r);
    return (int)g_recs.size() - 1;
}

static void prof_end(int h, double bytes)
{
    if (h < 0) return;
    hipEventRecord(g_recs[h].b, g_recs[h].st);
    g_recs[h].bytes = bytes;
}

static void prof_set_bytes(int h, double bytes)
{
    if (h >= 0) g_recs[h].bytes = bytes;
}

/* per-launch lookup counters of the solve kernel (pairs of u64) */
static unsigned long long *g_stats = nullptr;
static int g_prof_count = 1;   /* timed solves count their lookups */
static int g_stats_cap = 0, g_stats_used = 0;

static void prof_collect()
{
    if (g_recs.empty()) return;
    hipStreamSynchronize(E.st);
    if (E.st2) hipStreamSynchronize(E.st2);
    std::vector<unsigned long long> st;
    if (g_stats_used) {
        st.resize(2 * g_stats_used);
        hipMemcpy(st.data(), g_stats, st.size() * 8, hipMemcpyDeviceToHost);
    }
    for (auto &r : g_recs) {
        if (r.stats_slot >= 0) {
            double L = (double)st[2 * r.stats_slot], H = (double)st[2 * r.stats_slot + 1];
            r.bytes = r.base_bytes + 24.0 * L + 8.0 * H;
        }
        float ms = 0;
        hipEventElapsedTime(&ms, r.a, r.b);
        g_ms[r.kind] += ms;
        g_bytes[r.kind] += r.bytes;
        g_launches[r.kind]++;
        const int lane = (E.st2 && r.st == E.st2) ? 1 : 0;
        g_ms_s[lane][r.kind] += ms;
        g_bytes_s[lane][r.kind] += r.bytes;
        g_launches_s[lane][r.kind]++;
        g_evpool.push_back(r.a);
        g_evpool.push_back(r.b);
    }
    g_recs.clear();
    g_stats_used = 0;
}

/* sort hooks: the radix passes run inside kernels.hip; these brackets let
 * the timers see each pass */
namespace cg {
int prof_hook_begin(int kind, hipStream_t st) { return prof_begin(kind, st); }
void prof_hook_end(int h, double bytes, hipStream_t) { prof_end(h, bytes); }
}

extern "C" void cg_prof_enable(int on)
{
    eng_init();
    g_prof = on != 0;
}

extern "C" int cg_prof_select(const char *kernel)
{
    g_prof_main = false;
    if (!kernel || !*kernel) {
        g_prof_only = -1;
        return 0;
    }
    /* "<kernel>@main": that kernel's launches on the main stream only */
    std::string k(kernel);
    const size_t at = k.find('@');
    if (at != std::string::npos) {
        if (k.substr(at) != "@main") return fail_msg("cg_prof_select: unknown stream");
        k = k.substr(0, at);
    }
    int sel = -2;
    for (int i = 0; i < PK_N; i++)
        if (k == PK_NAMES[i]) sel = i;
    if (k == "sort") sel = PK_GROUP_SORT;
    if (sel == -2) return fail_msg("cg_prof_select: unknown kernel");
    g_prof_only = sel;
    g_prof_main = at != std::string::npos;
    return 0;
}

extern "C" void cg_prof_reset(void)
{
    prof_collect();
    for (int i = 0; i < PK_N; i++) {
        g_ms[i] = g_bytes[i] = 0, g_launches[i] = 0;
        for (int l = 0; l < 2; l++) g_ms_s[l][i] = g_bytes_s[l][i] = 0, g_launches_s[l][i] = 0;
    }
}

extern "C" long cg_prof_query(const char *kernel, double *total_ms, double *total_bytes)
{
    prof_collect();
    for (int i = 0; i < PK_N; i++)
        if (!strcmp(kernel, PK_NAMES[i])) {
            if (total_ms) *total_ms = g_ms[i];
            if (total_bytes) *total_bytes = g_bytes[i];
            return g_launches[i];
        }
    return -1;
}

extern "C" long cg_prof_query_stream(const char *kernel, int stream, double *total_ms,
                                     double *total_bytes)
{
    prof_collect();
    if (stream < 0 || stream > 1) return -1;
    for (int i = 0; i < PK_N; i++)
        if (!strcmp(kernel, PK_NAMES[i])) {
            if (total_ms) *total_ms = g_ms_s[stream][i];
            if (total_bytes) *total_bytes = g_bytes_s[stream][i];
            return g_launches_s[stream][i];
        }
    return -1;
}

/* ------------------------------------------------------------------------ */
/* layer arena                                                               */
/* ------------------------------------------------------------------------ */

struct Arena {
    struct Chunk {
        char *p;
        size_t size;
    };
    std::vector<Chunk> chunks;
    size_t idx = 0;          /* chunk being filled */
    size_t used = 0;         /* bytes used in chunks[idx] */
    static constexpr size_t CHUNK = (size_t)2 << 30;

    void *alloc(size_t bytes)
    {
        bytes = (bytes + 255) & ~(size_t)255;
        if (bytes == 0) bytes = 256;
        while (idx < chunks.size() && used + bytes > chunks[idx].size) {
            idx++;
            used = 0;
        }
        if (idx == chunks.size()) {
            size_t c = std::max(CHUNK, bytes);
            void *p = nullptr;
            if (hipMalloc(&p, c) != hipSuccess) return nullptr;
            chunks.push_back({ (char *)p, c });
            used = 0;
        }
        void *r = chunks[idx].p + used;
        used += bytes;
        return r;
    }
    /* gives back the tail of the most recent allocation */
    void shrink_last(void *p, size_t old_bytes, size_t new_bytes)
    {
        old_bytes = (old_bytes + 255) & ~(size_t)255;
        new_bytes = (new_bytes + 255) & ~(size_t)255;
        if (old_bytes == 0) old_bytes = 256;
        if (new_bytes == 0) new_bytes = 256;
        if (idx < chunks.size() && (char *)p + old_bytes == chunks[idx].p + used && new_bytes <= old_bytes)
            used -= old_bytes - new_bytes;
    }
    /* rewinds without freeing: the chunks are reused by the next run */
    void reset()
    {
        idx = 0;
        used = 0;
    }
    void release()
    {
        for (auto &c : chunks) hipFree(c.p);
        chunks.clear();
        reset();
    }
};

/* ------------------------------------------------------------------------ */
/* the generation step                                                       */
/* ------------------------------------------------------------------------ */

/* canonical forms through the formation's stabiliser (knob canon_stab) */
static int g_canon_stab = 1;

/* the selection for a table whose boards have the formation of the
 * canonical board e */
static cg_canon_sel_t canon_sel_of(uint64_t e, const cg_rules &r)
{
    return g_canon_stab ? cg_make_canon_sel(e, r.free_formation) : cg_canon_all();
}

static MoveRules move_rules(const cg_rules &r)
{
    MoveRules m;
    m.ff = r.free_formation;
    m.prune = r.prune;
    m.prune_inputs = 1;
    m.prune_moves = 1;
    m.stsl = r.stsl;
    m.ltc = r.ltc;
    m.sl = r.smallest_large;
    m.cs = cg_canon_all();
    return m;
}

/* sorts repeated on the LSD path after an oversize bucket */
static unsigned long long g_oversize_repeats = 0;

/* the last sort enqueued on each error word, named in an error message */
struct SortErrNote {
    int slot = -1;
    size_t n = 0;
    uint64_t vary = 0, cell_max = 0;
    int sample = 0;
};
static SortErrNote g_err_note[2];

static int fail_sort_word(const char *what, unsigned e1, unsigned e2)
{
    char buf[512];
    const int w = e1 ? 0 : 1;
    const SortErrNote &en = g_err_note[w];
    snprintf(buf, sizeof buf,
             "%s (error words %#x %#x; last sort on word %d: count slot %d, %zu keys, key bits %#llx, "
             "cell bound %#llx, mode %d)",
             what, e1, e2, w, en.slot, en.n, (unsigned long long)en.vary, (unsigned long long)en.cell_max,
             en.sample);
    return fail_msg(buf);
}

/* copies device counters [lo, lo+n) to E.h_cnt, synchronises and checks the
 * sort error flag */
static void publish_counters(int lo, int n)
{
    /* one small kernel stores the counters and the flags into mapped host
     * memory (instead of two copy dispatches) */
    launch_publish(E.st, E.d_cnt, lo, n, E.d_err, E.hd_pub, CNT_N);
}

/* waits for the main stream and takes the counters [lo, lo+n) and the error
 * words that publish_counters (or a graph holding it) stored */
static int collect_counters(int lo, int n)
{
    /* the oversize flags describe this read only */
    E.oversize = E.oversize1 = E.oversize2 = false;
    if (check_sync(E.st, "counters")) return -1;
    for (int i = lo; i < lo + n; i++) E.h_cnt[i] = ((volatile unsigned long long *)E.h_pub)[i];
    const unsigned e1 = (unsigned)((volatile unsigned long long *)E.h_pub)[CNT_N];
    const unsigned e2 = (unsigned)((volatile unsigned long long *)E.h_pub)[CNT_N + 1];
    const unsigned e = e1 | e2;
    if (e) {
        launch_fill_u64(E.st, (uint64_t *)E.d_err, 1, 0);
        hipStreamSynchronize(E.st);
        if (e == SORT_ERR_OVERSIZE) {   /* the caller repeats that sort */
            E.oversize = true;
            E.oversize1 = e1 != 0;
            E.oversize2 = e2 != 0;
            g_oversize_repeats += (e1 != 0) + (e2 != 0);
            return 0;
        }
        if (e & 2u) return fail_msg("batch: a board's formation cells left the table's canonical place");
        if (e & SORT_ERR_CODE) return fail_sort_word("sort: a key's cell value exceeds the bound of its layer", e1, e2);
        return fail_sort_word("sort: a chained-scan wait exceeded its budget", e1, e2);
    }
    return 0;
}

static int read_counters(int lo, int n)
{
    publish_counters(lo, n);
    return collect_counters(lo, n);
}

static uint64_t vary_of(size_t n, uint64_t o, uint64_t a) { return n > 1 ? (o & ~a) : 0; }

/* enqueues sort + unique of buf[0..n) into dst; count lands in d_cnt[slot] */
/* lane 0: the main stream; lane 1: the second stream (its own ping-pong
 * buffer, scratch and sort context) -- the caller orders the main stream
 * after it (join_lane1) before using its output */
/* sample = SORT_LSD: the LSD path (the repeat of a sample sort whose
 * bucket overflowed); the sample sort leaves buf unchanged, so a repeat can
 * read it again */
static int sort_enqueue(uint64_t *buf, size_t n, uint64_t vary, uint64_t *dst, int slot, int *hu,
                        int lane = 0, uint64_t cell_max = CELLS_UNBOUNDED, int sample = SORT_SAMPLE,
                        KeyPrune kp = KeyPrune())
{
    const size_t pk = sort_part_keys(n, sample);
    /* the carry sorts (count slot 6) report into the second error word, the
     * others into the first, whichever stream runs them */
    const int ew = slot == 6 ? 1 : 0;
    unsigned *err = E.d_err + ew;
    {
        SortErrNote &en = g_err_note[ew];
        en.slot = slot;
        en.n = n;
        en.vary = vary;
        en.cell_max = cell_max;
        en.sample = sample;
    }
    if (lane == 1) {
        if (ensure(E.B2, n + 1) || ensure_scratch2(n) || (pk && ensure(E.Pt2, pk))) return -1;
        sort_unique(E.st2, buf, E.B2.p, n, vary, E.scratch2, dst, E.d_cnt + slot, err, hu, 1,
                    cell_max, E.Pt2.p, E.Pt2.cap, sample, kp);
    } else {
        if (ensure(E.B, n + 1) || ensure_scratch(n) || (pk && ensure(E.Pt, pk))) return -1;
        sort_unique(E.st, buf, E.B.p, n, vary, E.scratch, dst, E.d_cnt + slot, err, hu, 0,
                    cell_max, E.Pt.p, E.Pt.cap, sample, kp);
    }
    hipError_t e = kernels_take_error();
    if (e != hipSuccess) return fail("sort_unique", e);
    return 0;
}

/* later work on the main stream waits for everything enqueued on the second */
static int join_lane1()
{
    HC(hipEventRecord(E.ev2, E.st2));
    HC(hipStreamWaitEvent(E.st, E.ev2, 0));
    return 0;
}

/* later work on the second stream waits for everything enqueued on the main
 * one (inside a capture this brings the second stream into the graph) */
static int fork_lane1()
{
    HC(hipEventRecord(E.evf, E.st));
    HC(hipStreamWaitEvent(E.st2, E.evf, 0));
    return 0;
}

/* ---- layer steps captured in a graph ---- */

/* batched steps as one graph launch each (knob batch_graph; 0: the same
 * launches enqueued one by one) */
static int g_batch_graph = 1;
static unsigned long long g_graph_steps = 0, g_graph_instantiations = 0;

static int capture_begin()
{
    HC(hipStreamBeginCapture(E.st, hipStreamCaptureModeRelaxed));
    g_capturing = true;
    return 0;
}

/* ends the capture on the main stream and launches the graph through *exec:
 * updated in place when the new graph has the same shape, else instantiated
 * anew.  With ok = false the capture is only ended and dropped. */
static int capture_end_launch(hipGraphExec_t *exec, bool ok)
{
    g_capturing = false;
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(E.st, &g);
    if (!ok) {
        if (g) hipGraphDestroy(g);
        return -1;
    }
    if (e != hipSuccess) return fail("hipStreamEndCapture", e);
    if (*exec) {
        hipGraphNode_t bad = nullptr;
        hipGraphExecUpdateResult res = hipGraphExecUpdateSuccess;
        if (hipGraphExecUpdate(*exec, g, &bad, &res) != hipSuccess || res != hipGraphExecUpdateSuccess) {
            (void)hipGetLastError();
            hipGraphExecDestroy(*exec);
            *exec = nullptr;
        }
    }
    if (!*exec) {
        e = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            hipGraphDestroy(g);
            *exec = nullptr;
            return fail("hipGraphInstantiate", e);
        }
        g_graph_instantiations++;
    }
    hipGraphDestroy(g);
    HC(hipGraphLaunch(*exec, E.st));
    g_graph_steps++;
    return 0;
}

struct LayerAlloc {
    /* returns room for n keys; shrink() gives back the unused tail */
    virtual uint64_t *alloc(size_t n) = 0;
    virtual void shrink(uint64_t *p, size_t cap, size_t n) { (void)p, (void)cap, (void)n; }
    virtual ~LayerAlloc() {}
};

static int nonzero_nibbles(uint64_t x)
{
    int c = 0;
    for (int i = 0; i < 16; i++) c += ((x >> (4 * i)) & 0xf) != 0;
    return c;
}

/* The count of the carry C' is still on the device (d_cnt[6]) when a step
 * returns; the next step's first synchronisation brings it to the host. */
struct StepCounts {
    bool pending = false;
    size_t sumE = 0;         /* input size of the carry sort (upper bound of the count) */
    int hu4 = -1;            /* timer handle of its compaction, bytes set on arrival */
    size_t last_nM = 0;      /* previous layer size: sizes the spawn buffers */
    size_t emax = 16;        /* empty cells a board of the previous layer may have */
    /* the pending carry sort's input, key bits and cell bound: what its
     * repeat needs; s4 = the spawn-4 buffer the next step writes */
    const uint64_t *cin = nullptr;
    uint64_t vary4 = 0, cell4 = CELLS_UNBOUNDED;
    int s4 = 0;
    /* layer steps so far, and whether all of their spawn launches tested
     * the spawned boards (then the next pre-move set needs no test) */
    int steps = 0;
    bool spawns_pruned = true;
    KeyPrune kp4 = KeyPrune();   /* the prune test the carry sort applies */
    /* the carry C' = S4 kept as written (a multiset, like P), or sorted
     * and deduplicated on the second stream; fixed at the first step */
    bool carry_sorted = true;
};

/* spawn buffers are sized before |M_s| is known: for about twice the previous
 * layer (never more than the candidate count); a step whose spawns do not fit
 * is repeated with exact buffers */
static int g_spawn_est_mode = 0;   /* 1: undersized on purpose (exercises the repeat) */
static int g_carry_lane = 1;       /* stream of the carry sort: 1 the second, 0 the main */

static size_t spawn_rows_estimate(const StepCounts &pc, size_t ncand)
{
    if (g_spawn_est_mode == 1) return std::min(ncand, pc.last_nM / 2 + 1);
    return std::min(ncand, 2 * pc.last_nM + 65536);
}

static void counts_arrived(StepCounts &pc, size_t nC)
{
    prof_set_bytes(pc.hu4, 8.0 * pc.sumE + 8.0 * nC);
    pc.pending = false;
}

/* The pending carry sort (input still at pc.cin) had an oversize bucket: it
 * is repeated on the LSD path into C, on the main stream. */
static int carry_repeat(StepCounts &pc, uint64_t *C)
{
    E.oversize = false;
    int hu4;
    if (sort_enqueue(const_cast<uint64_t *>(pc.cin), pc.sumE, pc.vary4, C, 6, &hu4, 0, pc.cell4, SORT_LSD,
                     pc.kp4) ||
        read_counters(6, 1))
        return -1;
    if (E.oversize) return fail_msg("sort: an LSD sort reported an oversize bucket");
    pc.hu4 = hu4;
    return 0;
}

/* brings the pending count to the host (end of a generation) */
static int counts_finish(StepCounts &pc, uint64_t *C)
{
    if (!pc.pending) return 0;
    if (join_lane1() || read_counters(6, 1)) return -1;
    if (E.oversize && carry_repeat(pc, C)) return -1;
    counts_arrived(pc, E.h_cnt[6]);
    return 0;
}

/* Bound on the cell values of a generation's keys, for the sort's cell-pair
 * code.  Moves keep a board's sum of non-0xf tiles (0xf and 0xe cells never
 * merge) and a spawn adds a 2 or a 4, so when every initial board has the
 * same such sum, every board of the step that starts from sum s has sum s
 * and no tile above 2^floor(log2 s) outside its 0xf cells.  0xf cells are
 * never created or removed (spawns put 1 or 2), so every board keeps the
 * initial boards' count nf of them: when the AND of a key set has nf 0xf
 * cells, those are the 0xf cells of every key and the other cells are at
 * most floor(log2 s). */
struct ValueBound {
    bool ok = false;
    int nf = 0;
    long s = 0;   /* tile sum of the boards the next step starts from */
};

static ValueBound value_bound_of(const uint64_t *b, size_t n)
{
    ValueBound v;
    for (size_t i = 0; i < n; i++) {
        long sum = 0;
        int nf = 0;
        for (int c = 0; c < 16; c++) {
            unsigned t = (unsigned)(b[i] >> (4 * c)) & 15u;
            if (t == 15) nf++;
            else if (t) sum += 1L << t;
        }
        if (i == 0) {
            v.s = sum;
            v.nf = nf;
        } else if (sum != v.s || nf != v.nf) {
            return ValueBound();
        }
    }
    v.ok = n > 0;
    return v;
}

/* floor(log2 s): the largest tile of a board of tile sum s (s >= 1) */
static int tile_log2(long s) { return 63 - __builtin_clzl((unsigned long)s); }

/* largest values of the varying cells in a key set of tile sum s whose AND
 * is key_and (as a cell_max word), or CELLS_UNBOUNDED */
static uint64_t cell_bound(const ValueBound *v, long s, uint64_t key_and)
{
    if (!v || !v->ok || s < 1) return CELLS_UNBOUNDED;
    int nf = 0;
    for (int c = 0; c < 16; c++) nf += ((key_and >> (4 * c)) & 15u) == 15u;
    if (nf != v->nf) return CELLS_UNBOUNDED;
    int lg = tile_log2(s);
    return lg <= 14 ? cells_bounded_by(lg) : CELLS_UNBOUNDED;
}

/* the sample sort of the *dn keys at buf (at most nmax of them) into dst,
 * sized on the device, on the main stream; its count lands in d_cnt[slot] */
static int sort_enqueue_dev(uint64_t *buf, const unsigned long long *dn, size_t nmax, uint64_t *dst, int slot,
                            int *hu, KeyPrune kp = KeyPrune())
{
    const size_t pk = sort_part_keys(nmax, SORT_SAMPLE);
    if (ensure_scratch(nmax) || (pk && ensure(E.Pt, pk))) return -1;
    g_err_note[0] = SortErrNote();
    g_err_note[0].slot = slot;
    g_err_note[0].n = nmax;
    g_err_note[0].sample = 3;   /* sized on the device */
    if (sort_unique_dev(E.st, buf, dn, nmax, E.scratch, dst, E.d_cnt + slot, E.d_err, hu, 0, E.Pt.p, E.Pt.cap,
                        kp)) {
        hipError_t e = kernels_take_error();
        return e != hipSuccess ? fail("sort_unique_dev", e) : fail_msg("sort: device-sized sort refused");
    }
    hipError_t e = kernels_take_error();
    if (e != hipSuccess) return fail("sort_unique_dev", e);
    return 0;
}

/* one host synchronisation per layer step when the M sort can be sized on
 * the device (knob one_sync; 0: the count of the candidates comes to the
 * host before their sort, two synchronisations per step) */
static int g_one_sync = 1;

/* with prune: the moved boards are tested by the candidate sort and the
 * spawn-4 boards by the carry sort, once per distinct board, and the
 * spawn-2 boards by the spawn kernel, once per board, instead of all of
 * them in the move kernel (knob prune_in_sort) */
static int g_prune_in_sort = 1;

/* the spawn-4 carry of a layer step moved as written, duplicates included
 * (knob carry_sort 0), or sorted and deduplicated on the second stream (1):
 * P_{s+4} is a multiset either way, and M_{s+4} = uniq(moves(P_{s+4})) is
 * the same set.  A C2 spawn-4 set holds each board 1.3 to 1.45 times; the
 * moves and the candidate sort of those duplicates cost less than the carry
 * sort did: C2 generated in 181-182 ms against 212-214 ms. */
static int g_carry_sort = 0;

/* One layer step (generate.c:409-550): the pre-move boards P and the carried
 * spawn-4 set C give M_s = uniq(moves(P)) (from `la`), the next pre-move
 * boards P_{s+2} = S2 u C (written to nxtP) and the new carry C' = uniq(S4)
 * (written to nxtC).  P is a multiset: the moves of a multiset have the same
 * set of results, and M is sorted and deduplicated anyway, so P_{s+2} needs
 * no sort of its own (its duplicates, about 15% on DPDGap, cost less in the
 * move kernel and the M sort than sorting all of it would).
 *
 * The M sort takes the candidate count from the device (the sample sort
 * sized for 4 |P| candidates), M is allocated for that bound and shrunk
 * later, and the spawns read |M| on the device, so the step synchronises
 * once, after the spawns: that read brings the candidate count, |M|, the
 * spawn count and the previous step's carry count together, and a sort with
 * an oversize bucket or spawn buffers that were too small are repeated then.
 * When the bound is beyond the sample sort, the candidate count comes to the
 * host first (two synchronisations).  When pc.pending, nC is an upper bound
 * and the exact count is on the device; on return *nC2 is an upper bound,
 * pc.pending is set and *nP2 is exact.  vb (optional) bounds the cell values
 * for the LSD sorts and advances to the next step's sum. */
static int gen_step(const cg_rules &r, const uint64_t *P, size_t nP, uint64_t *C, size_t nC,
                    StepCounts &pc, LayerAlloc &la, uint64_t **M, size_t *nM,
                    Buf &nxtP, size_t *nP2, Buf &nxtC, size_t *nC2, ValueBound *vb,
                    uint64_t *m_or = nullptr, uint64_t *m_and = nullptr, cg_canon_sel_t cs = cg_canon_all(),
                    bool premoved = false)
{
    const long s = vb ? vb->s : 0;
    const uint16_t *lut = get_lut(r.ignore_f);
    if (!lut) return fail_msg("row table upload failed");
    /* moves: at most 4 candidates per board (premove: the boards themselves) */
    const size_t cmax = premoved ? nP : 4 * nP;
    if (ensure(E.A, cmax + 1)) return -1;
    /* the moved boards' prune test, in the sort of the candidates: a layer
     * step's candidates hold each distinct board about 3.5 times, and the
     * test depends only on the board's cells (premove takes no prune) */
    KeyPrune kp = KeyPrune(), kps = KeyPrune();
    if (r.prune && g_prune_in_sort) {
        kps.on = 1;
        kps.stsl = r.stsl;
        kps.ltc = r.ltc;
        kps.sl = r.smallest_large;
        if (!premoved) kp = kps;
    }
    /* the move's and the spawns' statistics, both cleared now (the spawns'
     * again only when they are repeated) */
    launch_init_stats(E.st, E.d_cnt, E.d_cnt + 8);
    int h = prof_begin(PK_MOVE);
    if (premoved) {
        /* premove: the boards of P are the layer's positions -- their
         * canonical forms are the candidates, counted as the move kernel
         * counts its own (no move, no prune) */
        if (nP) HC(hipMemcpyAsync(E.A.p, P, nP * 8, hipMemcpyDeviceToDevice, E.st));
        launch_canon(E.st, E.A.p, nP);
        launch_fill_u64(E.st, (uint64_t *)E.d_cnt, 1, nP);
        launch_or_and(E.st, E.A.p, nP, E.d_cnt);
    } else {
        MoveRules mr = move_rules(r);
        mr.cs = cs;
        mr.prune_moves = kp.on ? 0 : 1;
        /* after the first step the pre-move boards are this generation's
         * spawns, tested by the spawn kernel and the carry sort when every
         * step had them do so */
        mr.prune_inputs = (kps.on && pc.steps > 0 && pc.spawns_pruned) ? 0 : 1;
        launch_move(E.st, P, nP, nullptr, E.A.p, E.d_cnt, lut, mr, 1);
    }
    prof_end(h, 0);
    /* M_s = uniq(candidates); its count goes to d_cnt[7] */
    const bool dev = g_one_sync && cmax <= sort_sample_dev_limit();
    size_t ncand = 0, mcap = 0;
    uint64_t mor = 0, mand = ~0ULL;
    int hu = -1;
    if (dev) {
        mcap = cmax;
        *M = la.alloc(mcap);
        if (!*M) return fail_msg("layer allocation failed");
        if (sort_enqueue_dev(E.A.p, E.d_cnt, cmax, *M, 7, &hu, kp)) return -1;
    } else {
        /* the candidate count first; the previous step's carry sort
         * (second stream) overlapped the moves and is settled here too */
        if (pc.pending && join_lane1()) return -1;
        if (read_counters(0, pc.pending ? 7 : 3)) return -1;
        ncand = E.h_cnt[0];
        mor = E.h_cnt[1];
        mand = E.h_cnt[2];
        if (pc.pending) {
            if (E.oversize && E.oversize1) return fail_msg("sort: unexpected oversize");
            if (E.oversize && E.oversize2 && carry_repeat(pc, C)) return -1;
            nC = E.h_cnt[6];
            counts_arrived(pc, nC);
        }
        mcap = ncand;
        *M = la.alloc(mcap);
        if (!*M) return fail_msg("layer allocation failed");
        if (sort_enqueue(E.A.p, ncand, vary_of(ncand, mor, mand), *M, 7, &hu, 0, cell_bound(vb, s, mand),
                         SORT_SAMPLE, kp))
            return -1;
    }
    /* spawns of M_s (the kernel reads |M_s| on the device; the grid covers
     * the bound): S2 to the front of nxtP, S4 to the spawn-4 buffer the
     * pending carry sort does not read */
    if (pc.steps == 0) pc.carry_sorted = g_carry_sort != 0;
    /* unsorted, the 4-spawns go straight to the carry buffer */
    Buf &S4 = pc.carry_sorted ? E.S4[pc.s4] : nxtC;
    const size_t mrows = mcap;
    size_t emax = dev ? pc.emax : (size_t)(16 - nonzero_nibbles(ncand ? mand : ~0ULL));
    size_t cap = emax * spawn_rows_estimate(pc, mrows);
    const size_t cbound = pc.pending ? pc.sumE : nC;   /* bound on the carry count */
    int hs = -1;
    size_t sumE = 0;
    bool settled = false;
    for (int attempt = 0; attempt < 4 && !settled; attempt++) {
        if (ensure(nxtP, cap + cbound + 1) || ensure(S4, cap + 1)) return -1;
        if (attempt > 0) launch_init_stats(E.st, E.d_cnt + 8);
        hs = prof_begin(PK_SPAWN);
        launch_spawn(E.st, *M, mrows, E.d_cnt + 7, nxtP.p, S4.p, cap, E.d_cnt + 8, cs, kps,
                     pc.carry_sorted ? 0 : 1);
        prof_end(hs, 0);
        /* the previous step's carry sort (second stream) overlapped this
         * step so far; its count is read with the others */
        if (pc.pending && join_lane1()) return -1;
        if (read_counters(0, 13)) return -1;
        const bool ov_m = E.oversize && E.oversize1, ov_c = E.oversize && E.oversize2;
        E.oversize = false;
        if (dev && attempt == 0) {
            ncand = E.h_cnt[0];
            mor = E.h_cnt[1];
            mand = E.h_cnt[2];
        }
        if (pc.pending) {
            /* the carry sort of the previous step: repeated on the LSD path
             * when a bucket overflowed */
            if (ov_c && carry_repeat(pc, C)) return -1;
            nC = E.h_cnt[6];
            counts_arrived(pc, nC);
        } else if (ov_c) {
            return fail_msg("sort: unexpected oversize");
        }
        if (ov_m) {
            /* the M sort had an oversize bucket: repeat it on the LSD path
             * (the candidates are unchanged), then the spawns */
            if (sort_enqueue(E.A.p, ncand, vary_of(ncand, mor, mand), *M, 7, &hu, 0, cell_bound(vb, s, mand),
                             SORT_LSD, kp))
                return -1;
            continue;
        }
        *nM = E.h_cnt[7];
        sumE = E.h_cnt[8];
        if (sumE <= cap) settled = true;
        else cap = (size_t)(16 - nonzero_nibbles(*nM ? mand : ~0ULL)) * *nM;   /* exact bound, always enough */
    }
    if (!settled) return fail_msg("generation step: spawns or sort not settled");
    if (m_or) *m_or = mor;
    if (m_and) *m_and = mand;
    pc.last_nM = *nM;
    pc.emax = (size_t)(16 - nonzero_nibbles(*nM ? mand : ~0ULL));
    prof_set_bytes(h, 8.0 * nP + 8.0 * ncand);
    prof_set_bytes(hu, 8.0 * ncand + 8.0 * *nM);
    prof_set_bytes(hs, 8.0 * *nM + 16.0 * sumE);
    la.shrink(*M, mcap, *nM);
    /* P_{s+2} = S2 u C: the carry after the 2-spawns */
    if (nC) HC(hipMemcpyAsync(nxtP.p + sumE, C, nC * sizeof(uint64_t), hipMemcpyDeviceToDevice, E.st));
    if (!pc.carry_sorted) {
        /* C' = S4 as the spawn kernel wrote it into nxtC */
        pc.steps++;
        pc.spawns_pruned = pc.spawns_pruned && kps.on;
        if (vb) vb->s = s + 2;
        *nP2 = sumE + nC;
        *nC2 = sumE;
        return 0;
    }
    const uint64_t or4 = E.h_cnt[11], and4 = E.h_cnt[12];
    /* C' = uniq(S4) on the second stream (the spawns finished at the
     * synchronisation above); it overlaps the next step's moves, and its
     * count stays on the device until then */
    if (ensure(nxtC, sumE + 1)) return -1;
    int hu4;
    pc.vary4 = vary_of(sumE, or4, and4);
    pc.cell4 = cell_bound(vb, s + 4, and4);
    pc.kp4 = kps;
    if (sort_enqueue(S4.p, sumE, pc.vary4, nxtC.p, 6, &hu4, g_carry_lane, pc.cell4, SORT_SAMPLE_SPARSE, kps))
        return -1;
    pc.cin = S4.p;
    pc.s4 ^= 1;
    pc.steps++;
    pc.spawns_pruned = pc.spawns_pruned && kps.on;
    if (vb) vb->s = s + 2;
    pc.pending = true;
    pc.sumE = sumE;
    pc.hu4 = hu4;
    *nP2 = sumE + nC;
    *nC2 = sumE;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* winstates                                                                 */
/* ------------------------------------------------------------------------ */

static int upload_win(const uint64_t *images, size_t n, const cg_rules &r, WinRules *w)
{
    if (n > E.wcap || !E.d_wmask) {
        if (E.d_wmask) hipFree(E.d_wmask);
        if (E.d_wval) hipFree(E.d_wval);
        size_t c = std::max<size_t>(n, 64);
        HC(hipMalloc((void **)&E.d_wmask, c * 8));
        HC(hipMalloc((void **)&E.d_wval, c * 8));
        E.wcap = c;
    }
    std::vector<uint64_t> m(n);
    for (size_t i = 0; i < n; i++) m[i] = cg_win_mask(images[i]);
    if (n) {
        HC(hipMemcpy(E.d_wmask, m.data(), n * 8, hipMemcpyHostToDevice));
        HC(hipMemcpy(E.d_wval, images, n * 8, hipMemcpyHostToDevice));
    }
    w->mask = E.d_wmask;
    w->val = E.d_wval;
    w->n = (int)n;
    w->nox = r.nox;
    w->score = r.score;
    return 0;
}

/* the 8 images of a winstate, get_all_rots (board.c:349-374) */
static void win_images(const uint64_t *ws, size_t n, std::vector<uint64_t> &out)
{
    out.clear();
    for (size_t i = 0; i < n; i++) {
        uint64_t b = ws[i], x = b;
        out.push_back(b);
        for (int k = 1; k < 4; k++) out.push_back(x = cg_rot_cw(x));
        x = cg_flip_rows(b);
        out.push_back(x);
        for (int k = 1; k < 4; k++) out.push_back(x = cg_rot_cw(x));
    }
}


/* ------------------------------------------------------------------------ */
/* device-resident tablebase                                                 */
/* ------------------------------------------------------------------------ */

struct TbLayer {
    uint64_t *keys = nullptr;
    size_t n = 0;
    double *vals = nullptr;
    uint64_t *arg = nullptr;
    uint32_t *dir = nullptr;
    LayerIndexPlan plan;
    /* the arena blocks the values and argmax words use, kept for the next
     * solve of the same generation (a re-solve allocates nothing) */
    double *vals_buf = nullptr;
    uint64_t *arg_buf = nullptr;
};

static LayerView make_view(const uint64_t *keys, const double *vals, size_t n, const uint32_t *dir,
                           const LayerIndexPlan &p)
{
    LayerView v;
    v.keys = keys;
    v.vals = vals;
    v.n = n;
    v.dir = dir;
    v.bits = p.bits;
    v.shift = p.shift;
    v.cmask = p.cmask;
    v.cval = p.cval;
    v.x = p.x;
    return v;
}

static LayerView empty_view()
{
    LayerView v;
    memset(&v, 0, sizeof v);
    return v;
}

/* a layer the 32-bit directory can index (LAYER_INDEX_MAX) */
static int check_layer_size(size_t n)
{
    if (n >= LAYER_INDEX_MAX)
        return fail_msg("layer index: a layer of 2^32 or more boards does not fit the 32-bit directory");
    return 0;
}

/* builds the prefix-bucket directory of a sorted unique layer into dir
 * (plan.dir_entries u32) */
static int build_index(const uint64_t *keys, size_t n, const LayerIndexPlan &plan, uint32_t *dir)
{
    if (check_layer_size(n)) return -1;
    if (ensure_scratch_bytes(dir_scratch_bytes(plan.dir_entries))) return -1;
    launch_build_dir(E.st, keys, n, plan, dir, E.scratch);
    return 0;
}

/* OR / AND of a device key array (one synchronisation) */
static int key_or_and(const uint64_t *keys, size_t n, uint64_t *o, uint64_t *a)
{
    launch_init_stats(E.st, E.d_cnt);
    launch_or_and(E.st, keys, n, E.d_cnt);
    if (read_counters(1, 2)) return -1;
    *o = E.h_cnt[1];
    *a = E.h_cnt[2];
    return 0;
}

/* the layer directories of a generation built after its last step
 * (knob dir_deferred): built between the steps, their kernels share the GPU
 * with the carry sort of the second stream and take about 2.5 times longer
 * (C2: 21.8 ms per generation against 8.7 ms alone) */
static int g_dir_deferred = 1;

struct cg_tablebase {
    cg_rules r;
    int start = 0;
    std::vector<TbLayer> L;
    Arena arena;
    double p4 = 0.1;
    bool have_arg = false;
    bool solved = false;
    int solve_lo = 0;          /* lowest solved layer index */
    std::vector<uint64_t> W;   /* expanded winstates */
    Buf P, C, P2, C2;
    /* adjoint workspace, kept between calls: occupation of every board,
     * D per layer + P, the initial boards */
    Buf occ, dD, init;
    /* the generation's canonical forms through the formation's stabiliser
     * (from the canonical form of the first initial board) */
    cg_canon_sel_t cs = cg_canon_all();
};

static LayerView view_of(const cg_tablebase *tb, int idx)
{
    if (idx < 0 || idx >= (int)tb->L.size()) return empty_view();
    const TbLayer &l = tb->L[idx];
    if (!l.vals) return empty_view();
    return make_view(l.keys, l.vals, l.n, l.dir, l.plan);
}

struct ArenaLayerAlloc : LayerAlloc {
    Arena *a;
    explicit ArenaLayerAlloc(Arena *x) : a(x) {}
    uint64_t *alloc(size_t n) override
    {
        return (uint64_t *)a->alloc(std::max<size_t>(n, 1) * sizeof(uint64_t));
    }
    void shrink(uint64_t *p, size_t cap, size_t n) override
    {
        a->shrink_last(p, std::max<size_t>(cap, 1) * 8, std::max<size_t>(n, 1) * 8);
    }
};

extern "C" cg_tablebase *cg_tb_create(const cg_rules *rules)
{
    if (eng_init()) return nullptr;
    cg_tablebase *tb = new cg_tablebase();
    tb->r = *rules;
    return tb;
}

static void buf_free(Buf &b)
{
    if (b.p) hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
}

extern "C" void cg_tb_free(cg_tablebase *tb)
{
    if (!tb) return;
    hipStreamSynchronize(E.st);
    tb->arena.release();
    buf_free(tb->P);
    buf_free(tb->C);
    buf_free(tb->P2);
    buf_free(tb->C2);
    buf_free(tb->occ);
    buf_free(tb->dD);
    buf_free(tb->init);
    delete tb;
}

extern "C" int cg_tb_generate(cg_tablebase *tb, const uint64_t *initial, size_t n, int start, int end)
{
    return cg_tb_generate_bounded(tb, initial, n, start, end, 0, nullptr);
}

extern "C" int cg_tb_generate_bounded(cg_tablebase *tb, const uint64_t *initial, size_t n, int start,
                                      int end, size_t max_boards, int *last_sum)
{
    if (last_sum) *last_sum = start - 2;
    if (eng_init()) return -1;
    HC(hipStreamSynchronize(E.st));
    tb->arena.reset();
    tb->L.clear();
    tb->start = start;
    tb->solved = false;
    tb->have_arg = false;
    if (ensure(tb->P, n + 1)) return -1;
    if (n) HC(hipMemcpy(tb->P.p, initial, n * 8, hipMemcpyHostToDevice));
    tb->cs = n ? canon_sel_of(cg_canon(initial[0]), tb->r) : cg_canon_all();
    size_t nP = n, nC = 0;
    StepCounts pc;
    ArenaLayerAlloc la(&tb->arena);
    ValueBound vb = value_bound_of(initial, n);
    size_t total = 0;
    for (int s = start; s <= end; s += 2) {
        TbLayer l;
        size_t nP2 = 0, nC2 = 0;
        uint64_t mor = 0, mand = ~0ULL;
        if (gen_step(tb->r, tb->P.p, nP, tb->C.p, nC, pc, la, &l.keys, &l.n, tb->P2, &nP2,
                     tb->C2, &nC2, &vb, &mor, &mand, tb->cs, tb->r.premove && s == start))
            return -1;
        l.plan = plan_index(l.n, mor, mand);
        l.dir = (uint32_t *)tb->arena.alloc(l.plan.dir_entries * 4);
        if (!l.dir) return fail_msg("index allocation failed");
        if (!g_dir_deferred && build_index(l.keys, l.n, l.plan, l.dir)) return -1;
        tb->L.push_back(l);
        total += l.n;
        if (last_sum) *last_sum = s;
        std::swap(tb->P, tb->P2);
        std::swap(tb->C, tb->C2);
        nP = nP2;
        nC = nC2;
        if (max_boards && total > max_boards) break;
    }
    if (counts_finish(pc, tb->C.p)) return -1;
    if (g_dir_deferred)
        for (const TbLayer &l : tb->L)
            if (build_index(l.keys, l.n, l.plan, l.dir)) return -1;
    return check_sync(E.st, "cg_tb_generate");
}

static int solve_layer_dev(const cg_rules &r, const uint64_t *M, size_t n, double *V, uint64_t *arg,
                           LayerView t2, LayerView t4, const WinRules &w, double p4)
{
    const uint16_t *lut = get_lut(r.ignore_f);
    if (!lut) return fail_msg("row table upload failed");
    unsigned long long *stats = nullptr;
    int slot = -1;
    /* lookup counters (the algorithmic bytes) only when the solve is timed
     * and counting is on: they cost atomics at the end of every block and the
     * registers of the kernel instance that keeps them */
    if (prof_wants(PK_SOLVE) && g_prof_count) {
        if (g_stats_used == g_stats_cap) {
            /* grow: collect pending records first so the old array can go */
            prof_collect();
            if (g_stats) hipFree(g_stats);
            g_stats_cap = g_stats_cap ? 2 * g_stats_cap : 4096;
            HC(hipMalloc((void **)&g_stats, (size_t)g_stats_cap * 16));
        }
        slot = g_stats_used++;
        stats = g_stats + 2 * slot;
        launch_fill_u64(E.st, (uint64_t *)stats, 2, 0);
    }
    int h = prof_begin(PK_SOLVE);
    launch_solve(E.st, M, n, V, arg, t2, t4, w, p4, lut, r.free_formation, stats);
    prof_end(h, 0);
    if (h >= 0) {
        g_recs[h].stats_slot = slot;
        g_recs[h].base_bytes = 16.0 * n + (arg ? 8.0 * n : 0.0);
    }
    return 0;
}

extern "C" int cg_tb_solve(cg_tablebase *tb, const uint64_t *winstates, size_t n_win, int end_solve,
                           double p4, int keep_argmax)
{
    if (eng_init()) return -1;
    win_images(winstates, n_win, tb->W);
    WinRules w;
    if (upload_win(tb->W.data(), tb->W.size(), tb->r, &w)) return -1;
    tb->p4 = p4;
    int nl = (int)tb->L.size();
    for (int i = 0; i < nl; i++) tb->L[i].vals = nullptr, tb->L[i].arg = nullptr;
    tb->solve_lo = nl;
    /* solve.c:203: s runs down from the top while s >= end_solve and s != 0 */
    for (int i = nl - 1; i >= 0; i--) {
        int s = tb->start + 2 * i;
        if (s < end_solve || s <= 0) break;
        TbLayer &l = tb->L[i];
        if (!l.vals_buf) l.vals_buf = (double *)tb->arena.alloc(std::max<size_t>(l.n, 1) * 8);
        if (keep_argmax && !l.arg_buf) l.arg_buf = (uint64_t *)tb->arena.alloc(std::max<size_t>(l.n, 1) * 8);
        if (!l.vals_buf || (keep_argmax && !l.arg_buf)) return fail_msg("value allocation failed");
        l.vals = l.vals_buf;
        if (keep_argmax) l.arg = l.arg_buf;
        if (solve_layer_dev(tb->r, l.keys, l.n, l.vals, l.arg, view_of(tb, i + 1), view_of(tb, i + 2),
                            w, p4))
            return -1;
        tb->solve_lo = i;
    }
    tb->solved = true;
    tb->have_arg = keep_argmax != 0;
    return check_sync(E.st, "cg_tb_solve");
}

extern "C" int cg_tb_adjoint(cg_tablebase *tb, const uint64_t *initial, size_t n, double *P, double *dP,
                             double *D)
{
    if (eng_init()) return -1;
    if (!tb->solved || !tb->have_arg) return fail_msg("cg_tb_adjoint needs cg_tb_solve(keep_argmax=1)");
    const uint16_t *lut = get_lut(tb->r.ignore_f);
    if (!lut) return fail_msg("row table upload failed");
    int nl = (int)tb->L.size();
    WinRules w;
    if (upload_win(tb->W.data(), tb->W.size(), tb->r, &w)) return -1;
    /* occupation per layer + D per layer + P: workspace of the tablebase,
     * allocated on the first call (and when a larger table needs more) */
    size_t tot = 0;
    std::vector<size_t> off(nl + 1);
    for (int i = 0; i < nl; i++) off[i] = tot, tot += tb->L[i].n;
    off[nl] = tot;
    if (ensure(tb->occ, tot + 1) || ensure(tb->dD, nl + 2) || ensure(tb->init, n + 1)) return -1;
    double *occ = (double *)tb->occ.p, *dD = (double *)tb->dD.p;
    uint64_t *dinit = tb->init.p;
    launch_fill_u64(E.st, tb->occ.p, tot + 1, 0);
    launch_fill_u64(E.st, tb->dD.p, nl + 2, 0);
    if (n) HC(hipMemcpyAsync(dinit, initial, n * 8, hipMemcpyHostToDevice, E.st));
    double *dPp = dD + nl;
    if (nl > 0 && tb->solve_lo == 0)
        launch_adjoint_init(E.st, dinit, n, view_of(tb, 0), occ, lut, tb->r.free_formation, dPp,
                            tb->r.premove);
    for (int i = 0; i < nl; i++) {
        if (!tb->L[i].vals) continue;
        LayerView t2 = view_of(tb, i + 1), t4 = view_of(tb, i + 2);
        double *o2 = (i + 1 < nl) ? occ + off[i + 1] : nullptr;
        double *o4 = (i + 2 < nl) ? occ + off[i + 2] : nullptr;
        int h = prof_begin(PK_ADJOINT);
        launch_adjoint(E.st, tb->L[i].keys, tb->L[i].n, occ + off[i], tb->L[i].arg, t2, o2, t4, o4, w,
                       tb->p4, lut, tb->r.free_formation, dD + i);
        prof_end(h, 24.0 * tb->L[i].n);
    }
    std::vector<double> h(nl + 1);
    HC(hipMemcpyAsync(h.data(), dD, (nl + 1) * 8, hipMemcpyDeviceToHost, E.st));
    if (check_sync(E.st, "cg_tb_adjoint")) return -1;
    double s = 0;
    for (int i = 0; i < nl; i++) {
        s += h[i];
        if (D) D[i] = h[i];
    }
    if (P) *P = h[nl];
    if (dP) *dP = s;
    return 0;
}

extern "C" int cg_tb_num_layers(const cg_tablebase *tb) { return (int)tb->L.size(); }
extern "C" int cg_tb_start(const cg_tablebase *tb) { return tb->start; }

extern "C" size_t cg_tb_layer_size(const cg_tablebase *tb, int idx)
{
    return (idx >= 0 && idx < (int)tb->L.size()) ? tb->L[idx].n : 0;
}

extern "C" size_t cg_tb_total_boards(const cg_tablebase *tb)
{
    size_t s = 0;
    for (auto &l : tb->L) s += l.n;
    return s;
}

extern "C" int cg_tb_copy_layer(const cg_tablebase *tb, int idx, uint64_t *keys, double *values)
{
    if (idx < 0 || idx >= (int)tb->L.size()) return fail_msg("layer index out of range");
    const TbLayer &l = tb->L[idx];
    if (keys && l.n) HC(hipMemcpy(keys, l.keys, l.n * 8, hipMemcpyDeviceToHost));
    if (values && l.n) {
        if (!l.vals) return fail_msg("layer not solved");
        HC(hipMemcpy(values, l.vals, l.n * 8, hipMemcpyDeviceToHost));
    }
    return 0;
}

extern "C" int cg_tb_layer_device(const cg_tablebase *tb, int idx, const uint64_t **keys,
                                  const double **values)
{
    if (idx < 0 || idx >= (int)tb->L.size()) return fail_msg("layer index out of range");
    if (keys) *keys = tb->L[idx].keys;
    if (values) *values = tb->L[idx].vals;
    return 0;
}

extern "C" int cg_tb_write_files(const cg_tablebase *tb, const char *boards_fmt, const char *tables_fmt)
{
    std::vector<uint64_t> k;
    std::vector<double> v;
    for (int i = 0; i < (int)tb->L.size(); i++) {
        const TbLayer &l = tb->L[i];
        int s = tb->start + 2 * i;
        k.resize(l.n);
        if (cg_tb_copy_layer(tb, i, k.data(), nullptr)) return -1;
        char name[4096];
        if (boards_fmt) {
            snprintf(name, sizeof name, boards_fmt, s);
            FILE *f = fopen(name, "wb");
            if (!f) return fail_msg("cannot open boards file");
            fwrite(k.data(), 8, l.n, f);
            fclose(f);
        }
        if (tables_fmt && l.vals) {
            v.resize(l.n);
            if (cg_tb_copy_layer(tb, i, nullptr, v.data())) return -1;
            snprintf(name, sizeof name, tables_fmt, s);
            FILE *f = fopen(name, "wb");
            if (!f) return fail_msg("cannot open tables file");
            fwrite(k.data(), 8, l.n, f);
            fwrite(v.data(), 8, l.n, f);
            fclose(f);
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* device primitives                                                         */
/* ------------------------------------------------------------------------ */

extern "C" int cg_dev_move(const cg_rules *rules, const uint64_t *d_in, size_t n, uint64_t *d_out,
                           size_t *n_out)
{
    if (eng_init()) return -1;
    const uint16_t *lut = get_lut(rules->ignore_f);
    if (!lut) return fail_msg("row table upload failed");
    launch_init_stats(E.st, E.d_cnt);
    int h = prof_begin(PK_MOVE);
    launch_move(E.st, d_in, n, nullptr, d_out, E.d_cnt, lut, move_rules(*rules));
    prof_end(h, 0);
    if (read_counters(0, 1)) return -1;
    *n_out = E.h_cnt[0];
    prof_set_bytes(h, 8.0 * n + 8.0 * *n_out);
    return 0;
}

extern "C" int cg_dev_spawn(const uint64_t *d_in, size_t n, uint64_t *d_out2, uint64_t *d_out4,
                            size_t *n_out)
{
    if (eng_init()) return -1;
    launch_init_stats(E.st, E.d_cnt);
    int h = prof_begin(PK_SPAWN);
    launch_spawn(E.st, d_in, n, nullptr, d_out2, d_out4, 16 * n, E.d_cnt);
    prof_end(h, 0);
    if (read_counters(0, 1)) return -1;
    *n_out = E.h_cnt[0];
    prof_set_bytes(h, 8.0 * n + 16.0 * *n_out);
    return 0;
}

extern "C" int cg_dev_sort_unique_bounded(uint64_t *d_keys, size_t n, uint64_t *d_out,
                                          size_t *n_unique, uint64_t cell_max)
{
    if (eng_init()) return -1;
    launch_init_stats(E.st, E.d_cnt);
    launch_or_and(E.st, d_keys, n, E.d_cnt);
    if (read_counters(0, 3)) return -1;
    int hu;
    const uint64_t vary = vary_of(n, E.h_cnt[1], E.h_cnt[2]);
    if (sort_enqueue(d_keys, n, vary, d_out, 5, &hu, 0, cell_max)) return -1;
    if (read_counters(5, 1)) return -1;
    if (E.oversize) {   /* a bucket of the sample sort overflowed: the LSD path */
        E.oversize = false;
        if (sort_enqueue(d_keys, n, vary, d_out, 5, &hu, 0, cell_max, SORT_LSD) || read_counters(5, 1))
            return -1;
    }
    *n_unique = E.h_cnt[5];
    prof_set_bytes(hu, 8.0 * n + 8.0 * *n_unique);
    return 0;
}

extern "C" int cg_dev_sort_unique(uint64_t *d_keys, size_t n, uint64_t *d_out, size_t *n_unique)
{
    return cg_dev_sort_unique_bounded(d_keys, n, d_out, n_unique, CELLS_UNBOUNDED);
}

extern "C" int cg_sort_code_probe(const uint64_t *keys, size_t n, uint64_t vary, uint64_t cell_max,
                                  uint64_t *codes, uint64_t *back)
{
    SortCode c = make_pair_code(vary, cell_max);
    if (!c.cells) return -1;
    for (size_t i = 0; i < n; i++) {
        uint32_t dm = 0;
        codes[i] = sc_encode(keys[i], c, dm);
        back[i] = sc_decode(codes[i], c) | (keys[i] & ~c.vmask);
    }
    return c.nbits;
}

extern "C" int cg_canon_sel_probe(uint64_t e, int ff, const uint64_t *b, size_t n, int o, uint64_t *out,
                                  uint32_t *sel)
{
    const cg_canon_sel_t cs = cg_make_canon_sel(e, ff);
    for (size_t i = 0; i < n; i++) out[i] = cg_canon_fast(b[i], cs, o & 1);
    if (sel) {
        sel[0] = cs.sel[0];
        sel[1] = cs.sel[1];
    }
    return 0;
}

extern "C" int cg_layer_index_probe(size_t n, uint64_t key_or, uint64_t key_and, int *bits)
{
    if (check_layer_size(n)) return -1;
    LayerIndexPlan p = plan_index(n, key_or, key_and);
    if (bits) *bits = p.bits;
    return 0;
}

/* a temporary index for caller-owned layers */
struct TmpIndex {
    uint32_t *dir = nullptr;
    LayerIndexPlan plan;
    ~TmpIndex()
    {
        if (dir) hipFree(dir);
    }
    int build(const uint64_t *keys, size_t n)
    {
        uint64_t o = 0, a = ~0ULL;
        if (n && key_or_and(keys, n, &o, &a)) return -1;
        plan = plan_index(n, o, a);
        HC(hipMalloc((void **)&dir, plan.dir_entries * 4));
        return build_index(keys, n, plan, dir);
    }
};

extern "C" int cg_dev_lookup(const uint64_t *d_keys, const double *d_vals, size_t n, const uint64_t *d_q,
                             size_t nq, double *d_out, int canonicalize)
{
    if (eng_init()) return -1;
    TmpIndex ix;
    if (ix.build(d_keys, n)) return -1;
    LayerView t = make_view(d_keys, d_vals, n, ix.dir, ix.plan);
    int h = prof_begin(PK_LOOKUP);
    launch_lookup(E.st, t, d_q, nq, d_out, canonicalize);
    prof_end(h, 16.0 * nq);
    return check_sync(E.st, "cg_dev_lookup");
}

/* ------------------------------------------------------------------------ */
/* lookup without allocation                                                 */
/* ------------------------------------------------------------------------ */
/* cg_dev_lookup builds a temporary directory (hipMalloc) and its count and
 * scan buffers in the engine's workspace.  The variant below keeps all of it
 * in one file-scope device array, so it calls no allocator: the layer's OR /
 * AND counters (64 bytes), a directory of at most 2^NA_DIR_BITS buckets,
 * and the directory build's counts and scan sums.  The array lives on every
 * GPU the library runs on until the process ends; its address on the
 * current device is taken once per device with hipGetSymbolAddress (the
 * array's name is a host-side stand-in, never a device address).  Work goes
 * to the null stream; calls that overlap on one GPU share the array and
 * race. */
constexpr int NA_DIR_BITS = 20;
constexpr size_t NA_DIR_ENTRIES = ((size_t)1 << NA_DIR_BITS) + 1;
constexpr size_t NA_STATS_BYTES = 256;
constexpr size_t NA_DIR_BYTES = (NA_DIR_ENTRIES * 4 + 255) & ~(size_t)255;
/* the build's counts (one word per bucket, rounded up) and per-tile sums:
 * at most two words per bucket plus slack, whatever the scan's tile */
constexpr size_t NA_BUILD_BYTES = (NA_DIR_ENTRIES * 8 + 4096 + 255) & ~(size_t)255;
constexpr size_t NA_SCRATCH_BYTES = NA_STATS_BYTES + NA_DIR_BYTES + NA_BUILD_BYTES;

static __device__ __align__(256) unsigned char scratch[NA_SCRATCH_BYTES];

static std::mutex g_na_mu;
static void *g_na_ptr[64];   /* scratch's address per device */

/* the array's device address on the current device */
static int na_scratch(unsigned char **p)
{
    int dev = 0;
    HC(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return fail_msg("cg_dev_lookup_noalloc: device ordinal out of range");
    std::lock_guard<std::mutex> g(g_na_mu);
    if (!g_na_ptr[dev]) {
        void *q = nullptr;
        HC(hipGetSymbolAddress(&q, HIP_SYMBOL(scratch)));
        g_na_ptr[dev] = q;
    }
    *p = (unsigned char *)g_na_ptr[dev];
    return 0;
}

extern "C" int cg_dev_lookup_noalloc(const uint64_t *d_keys, const double *d_vals, size_t n,
                                     const uint64_t *d_q, size_t nq, double *d_out, int canonicalize)
{
    if (check_layer_size(n)) return -1;
    unsigned char *base = nullptr;
    if (na_scratch(&base)) return -1;
    unsigned long long *stats = (unsigned long long *)base;
    uint32_t *dir = (uint32_t *)(base + NA_STATS_BYTES);
    void *build = base + NA_STATS_BYTES + NA_DIR_BYTES;
    const hipStream_t st = 0;
    (void)kernels_take_error();
    uint64_t o = 0, a = ~0ULL;
    if (n) {
        launch_init_stats(st, stats);
        launch_or_and(st, d_keys, n, stats);
        unsigned long long h[3];
        HC(hipMemcpy(h, stats, sizeof h, hipMemcpyDeviceToHost));
        o = h[1];
        a = h[2];
    }
    LayerIndexPlan plan = plan_index(n, o, a);
    if (plan.bits > NA_DIR_BITS) {   /* a coarser directory: longer searches, the same results */
        plan.bits = NA_DIR_BITS;
        plan.shift = plan.x.nbits - NA_DIR_BITS;
        plan.dir_entries = NA_DIR_ENTRIES;
    }
    if (plan.dir_entries > NA_DIR_ENTRIES || dir_scratch_bytes(plan.dir_entries) > NA_BUILD_BYTES)
        return fail_msg("cg_dev_lookup_noalloc: the directory exceeds the fixed scratch array");
    launch_build_dir(st, d_keys, n, plan, dir, build);
    launch_lookup(st, make_view(d_keys, d_vals, n, dir, plan), d_q, nq, d_out, canonicalize);
    return check_sync(st, "cg_dev_lookup_noalloc");
}

extern "C" int cg_dev_solve_layer(const cg_rules *rules, const uint64_t *d_M, size_t n, double *d_V,
                                  const uint64_t *d_K2, const double *d_V2, size_t n2,
                                  const uint64_t *d_K4, const double *d_V4, size_t n4,
                                  const uint64_t *win, size_t n_images, double p4)
{
    if (eng_init()) return -1;
    WinRules w;
    if (upload_win(win, n_images, *rules, &w)) return -1;
    TmpIndex i2, i4;
    if (i2.build(d_K2, n2) || i4.build(d_K4, n4)) return -1;
    LayerView t2 = make_view(d_K2, d_V2, n2, i2.dir, i2.plan);
    LayerView t4 = make_view(d_K4, d_V4, n4, i4.dir, i4.plan);
    int rc = solve_layer_dev(*rules, d_M, n, d_V, nullptr, t2, t4, w, p4);
    if (!rc) rc = check_sync(E.st, "cg_dev_solve_layer");
    return rc;
}

extern "C" int cg_dev_random_boards(uint64_t *d_out, size_t n, uint64_t seed, uint64_t fmask,
                                    double p_empty, int kmax, int uniform)
{
    if (eng_init()) return -1;
    launch_random_boards(E.st, d_out, n, seed, fmask, p_empty, kmax, uniform);
    return check_sync(E.st, "cg_dev_random_boards");
}

extern "C" int cg_dev_canon(uint64_t *d_keys, size_t n)
{
    if (eng_init()) return -1;
    launch_canon(E.st, d_keys, n);
    return check_sync(E.st, "cg_dev_canon");
}

/* ------------------------------------------------------------------------ */
/* tuning knobs and the autotuner                                            */
/* ------------------------------------------------------------------------ */

extern "C" int cg_tune_set(const char *knob, int value)
{
    SortCfg f = sort_forced();
    if (!strcmp(knob, "sort_bits")) {
        if (value && (value < 8 || value > 10)) return fail_msg("sort_bits: 0 or 8..10");
        return sort_force(value, f.kpt, f.lbw);
    }
    if (!strcmp(knob, "sort_kpt")) {
        if (value && value != 7 && value != 8 && value != 16 && value != 32)
            return fail_msg("sort_kpt: 0, 7, 8, 16 or 32");
        return sort_force(f.bits, value, f.lbw);
    }
    if (!strcmp(knob, "sort_lbw")) {
        if (value && value != 16 && value != 32) return fail_msg("sort_lbw: 0, 16 or 32");
        return sort_force(f.bits, f.kpt, value);
    }
    if (!strcmp(knob, "dir_delta")) {
        if (value < -4 || value > 3) return fail_msg("dir_delta: -4..3");
        set_dir_delta(value);
        return 0;
    }
    if (!strcmp(knob, "spawn_estimate")) {
        if (value != 0 && value != 1) return fail_msg("spawn_estimate: 0 or 1");
        g_spawn_est_mode = value;
        return 0;
    }
    if (!strcmp(knob, "cell_code")) {
        if (value != 0 && value != 1) return fail_msg("cell_code: 0 or 1");
        sort_set_cell_code(value);
        return 0;
    }
    if (!strcmp(knob, "carry_stream")) {
        if (value != 0 && value != 1) return fail_msg("carry_stream: 0 or 1");
        g_carry_lane = value;
        return 0;
    }
    if (!strcmp(knob, "canon_stab")) {
        if (value != 0 && value != 1) return fail_msg("canon_stab: 0 or 1");
        g_canon_stab = value;
        return 0;
    }
    if (!strcmp(knob, "one_sync")) {
        if (value != 0 && value != 1) return fail_msg("one_sync: 0 or 1");
        g_one_sync = value;
        return 0;
    }
    if (!strcmp(knob, "sample_bmax")) {
        if (value != 2048 && value != 8192 && value != 16384) return fail_msg("sample_bmax: 2048, 8192 or 16384");
        sort_set_sample_limits(value, sort_sample_probes());
        return 0;
    }
    if (!strcmp(knob, "prof_count")) {
        if (value != 0 && value != 1) return fail_msg("prof_count: 0 or 1");
        g_prof_count = value;
        return 0;
    }
    if (!strcmp(knob, "part_prefetch")) {
        if (value != 0 && value != 1) return fail_msg("part_prefetch: 0 or 1");
        sort_set_part_prefetch(value);
        return 0;
    }
    if (!strcmp(knob, "part_blocks")) {
        if (value < 0 || value > (1 << 20)) return fail_msg("part_blocks: 0..2^20");
        sort_set_part_blocks(value);
        return 0;
    }
    if (!strcmp(knob, "bucket_bitonic")) {
        if (value != 0 && value != 1) return fail_msg("bucket_bitonic: 0 or 1");
        sort_set_bucket_bitonic(value);
        return 0;
    }
    if (!strcmp(knob, "solve_waves")) {
        if (value != 0 && value != 5 && value != 6) return fail_msg("solve_waves: 0, 5 or 6");
        solve_set_waves(value);
        return 0;
    }
    if (!strcmp(knob, "spawn_blocks")) {
        if (value < 64 || value > 65536) return fail_msg("spawn_blocks: 64..65536");
        spawn_set_blocks(value);
        return 0;
    }
    if (!strcmp(knob, "carry_sort")) {
        if (value != 0 && value != 1) return fail_msg("carry_sort: 0 or 1");
        g_carry_sort = value;
        return 0;
    }
    if (!strcmp(knob, "dir_deferred")) {
        if (value != 0 && value != 1) return fail_msg("dir_deferred: 0 or 1");
        g_dir_deferred = value;
        return 0;
    }
    if (!strcmp(knob, "prune_in_sort")) {
        if (value != 0 && value != 1) return fail_msg("prune_in_sort: 0 or 1");
        g_prune_in_sort = value;
        return 0;
    }
    if (!strcmp(knob, "distinct_bitonic")) {
        if (value < 0 || value > 16384) return fail_msg("distinct_bitonic: 0 (off) or at most 16384 values");
        sort_set_distinct_bitonic(value, sort_distinct_bitonic_bits());
        return 0;
    }
    if (!strcmp(knob, "distinct_bitonic_bits")) {
        if (value < 0 || value > 64) return fail_msg("distinct_bitonic_bits: 0..64");
        sort_set_distinct_bitonic(sort_distinct_bitonic_max(), value);
        return 0;
    }
    if (!strcmp(knob, "sample_probes")) {
        if (value < 0 || value > 16384) return fail_msg("sample_probes: 0 (default) or 1..16384");
        sort_set_sample_limits(sort_sample_bmax(), value);
        return 0;
    }
    if (!strcmp(knob, "batch_graph")) {
        if (value != 0 && value != 1) return fail_msg("batch_graph: 0 or 1");
        g_batch_graph = value;
        return 0;
    }
    if (!strcmp(knob, "sample_sort")) {
        if (value != 0 && value != 1) return fail_msg("sample_sort: 0 or 1");
        sort_set_sample(value, sort_sample_bucket(), sort_sample_cap());
        return 0;
    }
    if (!strcmp(knob, "sample_bucket")) {
        if (value < 1000 || value > 16384) return fail_msg("sample_bucket: 1000..16384");
        sort_set_sample(sort_sample_on(), value, sort_sample_cap());
        return 0;
    }
    if (!strcmp(knob, "sample_cap")) {
        if (value < 0 || value > 131072) return fail_msg("sample_cap: 0 (131072) or 1..131072");
        sort_set_sample(sort_sample_on(), sort_sample_bucket(), value);
        return 0;
    }
    if (!strcmp(knob, "reset") || !strcmp(knob, "builtin")) {
        sort_table_reset();
        sort_force(0, 0, 0);
        set_dir_delta(0);
        g_spawn_est_mode = 0;
        sort_set_cell_code(1);
        g_carry_lane = 1;
        g_canon_stab = 1;
        g_one_sync = 1;
        g_prune_in_sort = 1;
        g_dir_deferred = 1;
        g_carry_sort = 0;
        spawn_set_blocks(1024);
        solve_set_waves(6);
        g_batch_graph = 1;
        sort_set_sample(1, 5000, 0);
        sort_set_sample_limits(16384, 0);
        sort_set_bucket_bitonic(0);
        sort_set_part_blocks(512);
        sort_set_part_prefetch(1);
        g_prof_count = 1;
        sort_set_distinct_bitonic(2048, 9);
        if (!strcmp(knob, "reset") && !g_tune_path.empty()) return cg_tune_load(g_tune_path.c_str());
        return 0;
    }
    return fail_msg("unknown tuning knob");
}

extern "C" int cg_tune_get(const char *knob)
{
    SortCfg f = sort_forced();
    if (!strcmp(knob, "sort_bits")) return f.bits;
    if (!strcmp(knob, "sort_kpt")) return f.kpt;
    if (!strcmp(knob, "sort_lbw")) return f.lbw;
    if (!strcmp(knob, "dir_delta")) return get_dir_delta();
    if (!strcmp(knob, "spawn_estimate")) return g_spawn_est_mode;
    if (!strcmp(knob, "cell_code")) return sort_cell_code();
    if (!strcmp(knob, "carry_stream")) return g_carry_lane;
    if (!strcmp(knob, "canon_stab")) return g_canon_stab;
    if (!strcmp(knob, "one_sync")) return g_one_sync;
    if (!strcmp(knob, "prune_in_sort")) return g_prune_in_sort;
    if (!strcmp(knob, "dir_deferred")) return g_dir_deferred;
    if (!strcmp(knob, "carry_sort")) return g_carry_sort;
    if (!strcmp(knob, "spawn_blocks")) return spawn_blocks();
    if (!strcmp(knob, "solve_waves")) return solve_waves();
    if (!strcmp(knob, "batch_graph")) return g_batch_graph;
    if (!strcmp(knob, "sample_bmax")) return sort_sample_bmax();
    if (!strcmp(knob, "bucket_bitonic")) return sort_bucket_bitonic();
    if (!strcmp(knob, "part_blocks")) return sort_part_blocks();
    if (!strcmp(knob, "part_prefetch")) return sort_part_prefetch();
    if (!strcmp(knob, "prof_count")) return g_prof_count;
    if (!strcmp(knob, "distinct_bitonic")) return sort_distinct_bitonic_max();
    if (!strcmp(knob, "distinct_bitonic_bits")) return sort_distinct_bitonic_bits();
    if (!strcmp(knob, "sample_probes")) return sort_sample_probes();
    if (!strcmp(knob, "graph_steps")) return (int)std::min<unsigned long long>(g_graph_steps, 0x7fffffff);
    if (!strcmp(knob, "graph_instantiations"))
        return (int)std::min<unsigned long long>(g_graph_instantiations, 0x7fffffff);
    if (!strcmp(knob, "cell_code_sorts")) return (int)std::min<unsigned long long>(sort_code_count(), 0x7fffffff);
    if (!strcmp(knob, "sample_sort")) return sort_sample_on();
    if (!strcmp(knob, "sample_bucket")) return sort_sample_bucket();
    if (!strcmp(knob, "sample_cap")) return sort_sample_cap();
    if (!strcmp(knob, "sample_sorts")) return (int)std::min<unsigned long long>(sort_sample_count(), 0x7fffffff);
    if (!strcmp(knob, "sample_repeats")) return (int)std::min<unsigned long long>(g_oversize_repeats, 0x7fffffff);
    return fail_msg("unknown tuning knob");
}

/* Reads the picks of an autotuner report (the JSON cg_autotune writes):
 * the "pick" objects of the three "sort" classes ("<nbits>": "<bits>x<kpt>/
 * <window>") and the "dir" section's "pick" (directory depth).  Anything
 * malformed leaves the defaults in place and returns -1. */
static int tuning_parse(const std::string &js)
{
    size_t dpos = js.find("\"dir\"");
    if (dpos == std::string::npos) return fail_msg("tuning: no dir section");
    /* position just after `"pick"` + spaces + ':' + spaces, or npos */
    auto after_pick = [&](size_t from) -> size_t {
        size_t q = js.find("\"pick\"", from);
        if (q == std::string::npos) return q;
        q += 6;
        while (q < js.size() && (js[q] == ' ' || js[q] == ':' || js[q] == '\n')) q++;
        return q;
    };
    size_t pos = 0;
    std::vector<std::pair<int, SortCfg>> entries[SORT_SIZE_CLASSES];
    for (int cls = 0; cls < SORT_SIZE_CLASSES; cls++) {
        pos = after_pick(pos);
        if (pos == std::string::npos || pos > dpos || js[pos] != '{')
            return fail_msg("tuning: missing sort picks");
        pos++;
        size_t end = js.find('}', pos);
        if (end == std::string::npos) return fail_msg("tuning: unterminated picks");
        while (pos < end) {
            int nb, b, k, w, used = 0;
            if (sscanf(js.c_str() + pos, " \"%d\" : \"%dx%d/%d\"%n", &nb, &b, &k, &w, &used) != 4 || !used)
                return fail_msg("tuning: malformed pick");
            entries[cls].push_back({nb, SortCfg{b, k, w}});
            pos += used;
            while (pos < end && (js[pos] == ',' || js[pos] == ' ' || js[pos] == '\n')) pos++;
        }
    }
    size_t ppos = after_pick(dpos);
    int delta = 0;
    if (ppos == std::string::npos || sscanf(js.c_str() + ppos, "%d", &delta) != 1 || delta < -4 ||
        delta > 3)
        return fail_msg("tuning: malformed dir pick");
    sort_table_reset();
    for (int cls = 0; cls < SORT_SIZE_CLASSES; cls++)
        for (auto &e : entries[cls])
            if (sort_table_set(cls, e.first, e.second)) {
                sort_table_reset();
                return fail_msg("tuning: a pick names no kernel");
            }
    set_dir_delta(delta);
    return 0;
}

extern "C" int cg_tune_load(const char *path)
{
    FILE *f = fopen(path, "rb");
    if (!f) return fail_msg("cg_tune_load: cannot open the file");
    std::string js;
    char buf[4096];
    size_t r;
    while ((r = fread(buf, 1, sizeof buf, f)) > 0) js.append(buf, r);
    fclose(f);
    return tuning_parse(js);
}

/* At initialisation the library installs the table the autotuner wrote on
 * an MI355X, tuned/mi355x.json next to the library, unless CABLEGEN_TUNING
 * names another file (an empty value keeps the built-in defaults). */
static void tuning_load_default()
{
    const char *env = getenv("CABLEGEN_TUNING");
    std::string path;
    if (env) {
        path = env;
    } else {
        Dl_info info;
        if (!dladdr((void *)&tuning_load_default, &info) || !info.dli_fname) return;
        path = info.dli_fname;
        size_t slash = path.rfind('/');
        path = (slash == std::string::npos ? std::string(".") : path.substr(0, slash)) + "/tuned/mi355x.json";
    }
    if (path.empty()) return;
    FILE *f = fopen(path.c_str(), "rb");
    if (!f) return;
    fclose(f);
    if (cg_tune_load(path.c_str()))
        fprintf(stderr, "cablegen: ignoring tuning file %s (%s)\n", path.c_str(), g_err.c_str());
    else
        g_tune_path = path;
}

extern "C" int cg_tune_sort_entry(int size_class, int nbits, int bits, int kpt, int lbw)
{
    SortCfg c = {bits, kpt, lbw};
    if (sort_table_set(size_class, nbits, c)) return fail_msg("cg_tune_sort_entry: no such entry or configuration");
    return 0;
}

extern "C" int cg_tune_sort_query(size_t n, int nbits, int *bits, int *kpt, int *lbw)
{
    SortCfg c = sort_choose(n, nbits);
    if (bits) *bits = c.bits;
    if (kpt) *kpt = c.kpt;
    if (lbw) *lbw = c.lbw;
    return 0;
}

namespace {

struct TuneBufs {
    uint64_t *src = nullptr, *work = nullptr, *tmp = nullptr, *dst = nullptr;
    ~TuneBufs()
    {
        for (uint64_t *p : {src, work, tmp, dst})
            if (p) hipFree(p);
    }
    int alloc(size_t n)
    {
        for (uint64_t **p : {&src, &work, &tmp, &dst}) HC(hipMalloc((void **)p, (n + 1) * 8));
        return 0;
    }
};

/* histogram + pass time (ms) of sorting src & mask, best of reps */
int tune_sort_once(TuneBufs &b, size_t n, uint64_t mask, int reps, double *best)
{
    *best = 1e30;
    for (int r = 0; r < reps; r++) {
        launch_copy_and_u64(E.st, b.work, b.src, n, mask);
        cg_prof_reset();
        int hu;
        sort_unique(E.st, b.work, b.tmp, n, mask, E.scratch, b.dst, E.d_cnt + 7, E.d_err, &hu, 0,
                    CELLS_UNBOUNDED, nullptr, 0, false);   /* the LSD configurations */
        hipError_t e = kernels_take_error();
        if (e != hipSuccess) return fail("autotune sort", e);
        double mh = 0, mp = 0;
        cg_prof_query("sort_hist", &mh, nullptr);
        cg_prof_query("sort_pass", &mp, nullptr);
        *best = std::min(*best, mh + mp);
    }
    return read_counters(7, 1);
}

void jcat(std::string &s, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    s += buf;
}

} // namespace

/* Measures every sort configuration at every pass count on board-like keys
 * (random boards over 16 free cells, empty with p = 0.4, else exponent k with
 * P(k) ~ 2^-k, masked to the pass count's bits) for three sizes, and fills
 * the (size class, varying bits) table with the fastest; then times the
 * directory depths on a lookup workload (a 2^22-board DPDGap layer, 2^25
 * queries) and sets the fastest.  A configuration replaces the default only
 * when it is at least 3% faster.  effort scales the sizes (0: a quick pass
 * with sizes / 8, 1: full).  report (may be NULL) receives a JSON summary. */
extern "C" int cg_autotune(int effort, char *report, size_t cap)
{
    if (eng_init()) return -1;
    const bool was_prof = g_prof;
    const SortCfg forced = sort_forced();
    sort_force(0, 0, 0);
    sort_table_reset();
    g_prof = true;
    cg_prof_reset();
    const size_t sizes[SORT_SIZE_CLASSES] = {(size_t)1 << 18, (size_t)1 << 21,
                                             (size_t)1 << (effort > 0 ? 25 : 24)};
    const int reps = effort > 0 ? 3 : 2;
    const int ncfg = sort_num_configs();
    std::string js = "{\"sort\":[";
    int rc = 0;
    for (int cls = 0; cls < SORT_SIZE_CLASSES && !rc; cls++) {
        size_t n = sizes[cls];
        TuneBufs b;
        if ((rc = b.alloc(n)) || (rc = ensure_scratch(n))) break;
        launch_random_boards(E.st, b.src, n, 0x5eed0000ULL + cls, 0, 0.4, 11, 0);
        /* T[i][k]: configuration i with k passes */
        std::vector<std::vector<double>> T(ncfg, std::vector<double>(9, 1e30));
        jcat(js, "%s{\"n\":%zu,\"ms\":{", cls ? "," : "", n);
        for (int i = 0; i < ncfg && !rc; i++) {
            SortCfg c = sort_config(i);
            sort_force(c.bits, c.kpt, c.lbw);
            int kmax = (64 + c.bits - 1) / c.bits;
            jcat(js, "%s\"%dx%d/%d\":[", i ? "," : "", c.bits, c.kpt, c.lbw);
            for (int k = 1; k <= kmax && !rc; k++) {
                int nb = std::min(64, k * c.bits);
                uint64_t mask = nb >= 64 ? ~0ULL : ((1ULL << nb) - 1);
                rc = tune_sort_once(b, n, mask, reps, &T[i][k]);
                jcat(js, "%s%.4f", k > 1 ? "," : "", T[i][k]);
            }
            jcat(js, "]");
        }
        sort_force(0, 0, 0);
        if (rc) break;
        jcat(js, "},\"pick\":{");
        for (int nbits = 1; nbits <= 64; nbits++) {
            SortCfg d = sort_config(0);
            int best = 0;
            double tb = T[0][(nbits + d.bits - 1) / d.bits];
            for (int i = 1; i < ncfg; i++) {
                SortCfg c = sort_config(i);
                double t = T[i][(nbits + c.bits - 1) / c.bits];
                if (t < tb * (best == 0 ? 0.97 : 1.0)) tb = t, best = i;
            }
            SortCfg c = sort_config(best);
            sort_table_set(cls, nbits, c);
            jcat(js, "%s\"%d\":\"%dx%d/%d\"", nbits > 1 ? "," : "", nbits, c.bits, c.kpt, c.lbw);
        }
        jcat(js, "}}");
    }
    /* directory depth */
    if (!rc) {
        const size_t nl = (size_t)1 << (effort > 0 ? 22 : 20), nq = nl << 3;
        TuneBufs b;
        uint64_t *q = nullptr;
        double *out = nullptr;
        uint32_t *dir = nullptr;
        const size_t dir_cap = ((size_t)4 << (63 - __builtin_clzll(nl))) + 1;   /* delta <= +2 */
        rc = b.alloc(nl);
        if (!rc && (hipMalloc((void **)&q, nq * 8) != hipSuccess ||
                    hipMalloc((void **)&out, nq * 8) != hipSuccess ||
                    hipMalloc((void **)&dir, dir_cap * 4) != hipSuccess))
            rc = fail_msg("autotune: allocation failed");
        const uint64_t dpd = 0x00000000FFF0FFF0ULL;
        if (!rc) {
            /* canonical forms move the formation cells: the varying bits come
             * from the keys themselves */
            launch_random_boards(E.st, b.src, nl, 0x5eed1000ULL, dpd, 0.4, 11, 0);
            launch_canon(E.st, b.src, nl);
            uint64_t ko = 0, ka = ~0ULL;
            rc = key_or_and(b.src, nl, &ko, &ka);
            int hu;
            if (!rc) {
                sort_unique(E.st, b.src, b.tmp, nl, vary_of(nl, ko, ka), E.scratch, b.dst, E.d_cnt + 7,
                            E.d_err, &hu);
                rc = read_counters(7, 1);
            }
        }
        size_t nu = rc ? 0 : E.h_cnt[7];
        uint64_t o = 0, a = ~0ULL;
        if (!rc) rc = key_or_and(b.dst, nu, &o, &a);
        if (!rc) {
            launch_random_boards(E.st, q, nq, 0x5eed2000ULL, dpd, 0.4, 11, 0);
            launch_fill_u64(E.st, (uint64_t *)b.work, nu, 0);   /* values */
        }
        double best_t = 1e30;
        int best_d = 0;
        double t0 = 0;
        jcat(js, "],\"dir\":{\"layer\":%zu,\"queries\":%zu,\"ms\":{", nu, nq);
        for (int d = -3; d <= 2 && !rc; d++) {
            set_dir_delta(d);
            LayerIndexPlan plan = plan_index(nu, o, a);
            if (plan.dir_entries > dir_cap) continue;
            if (ensure_scratch_bytes(dir_scratch_bytes(plan.dir_entries))) {
                rc = -1;
                break;
            }
            double tbest = 1e30;
            for (int r = 0; r < reps && !rc; r++) {
                hipEvent_t e0 = ev_get(), e1 = ev_get();
                hipEventRecord(e0, E.st);
                launch_build_dir(E.st, b.dst, nu, plan, dir, E.scratch);
                LayerView t = make_view(b.dst, (const double *)b.work, nu, dir, plan);
                launch_lookup(E.st, t, q, nq, out, 1);
                hipEventRecord(e1, E.st);
                hipError_t e = hipEventSynchronize(e1);
                if (e != hipSuccess) { rc = fail("autotune lookup", e); break; }
                float ms = 0;
                hipEventElapsedTime(&ms, e0, e1);
                g_evpool.push_back(e0);
                g_evpool.push_back(e1);
                tbest = std::min(tbest, (double)ms);
            }
            if (d == 0) t0 = tbest;
            if (tbest < best_t) best_t = tbest, best_d = d;
            jcat(js, "%s\"%d\":%.4f", d > -3 ? "," : "", d, tbest);
        }
        if (!rc && t0 > 0 && best_t > 0.97 * t0) best_d = 0;
        set_dir_delta(best_d);
        jcat(js, "},\"pick\":%d}}", best_d);
        if (q) hipFree(q);
        if (out) hipFree(out);
        if (dir) hipFree(dir);
    }
    cg_prof_reset();
    g_prof = was_prof;
    if (forced.bits || forced.kpt || forced.lbw) sort_force(forced.bits, forced.kpt, forced.lbw);
    if (rc) {
        sort_table_reset();
        set_dir_delta(0);
        return -1;
    }
    if (report && cap) {
        strncpy(report, js.c_str(), cap - 1);
        report[cap - 1] = 0;
        if (js.size() >= cap) return fail_msg("cg_autotune: report buffer too small");
    }
    return 0;
}

extern "C" int cg_dev_sync(void)
{
    if (eng_init()) return -1;
    return check_sync(E.st, "cg_dev_sync");
}

/* ------------------------------------------------------------------------ */
/* streaming interface for the drop-in generate() / solve()                  */
/* ------------------------------------------------------------------------ */

static cg_rules to_rules(const cge_rules *r)
{
    cg_rules x;
    x.free_formation = r->free_formation;
    x.ignore_f = r->ignore_f;
    x.prune = r->prune;
    x.stsl = r->stsl;
    x.ltc = r->ltc;
    x.smallest_large = r->smallest_large;
    x.nox = r->nox;
    x.score = r->score;
    x.premove = r->premove;
    return x;
}

static struct {
    cg_rules r;
    Buf P, C, P2, C2, M;
    size_t nP = 0, nC = 0, nM = 0;
    StepCounts pc;
    ValueBound vb;
    cg_canon_sel_t cs;
    bool first;   /* the next step is the first (premove applies to it) */
} G;

struct BufLayerAlloc : LayerAlloc {
    uint64_t *alloc(size_t n) override
    {
        if (ensure(G.M, n + 1)) return nullptr;
        return G.M.p;
    }
};

extern "C" int cge_gen_begin(const cge_rules *r, const uint64_t *initial, size_t n)
{
    if (eng_init()) return -1;
    G.r = to_rules(r);
    if (ensure(G.P, n + 1)) return -1;
    if (n) HC(hipMemcpy(G.P.p, initial, n * 8, hipMemcpyHostToDevice));
    G.nP = n;
    G.nC = 0;
    G.nM = 0;
    G.pc = StepCounts();
    G.vb = value_bound_of(initial, n);
    G.cs = n ? canon_sel_of(cg_canon(initial[0]), G.r) : cg_canon_all();
    G.first = true;
    return 0;
}

extern "C" int cge_gen_step(size_t *n_layer)
{
    uint64_t *M = nullptr;
    size_t nP2 = 0, nC2 = 0;
    BufLayerAlloc la;
    const bool premoved = G.r.premove && G.first;
    G.first = false;
    if (gen_step(G.r, G.P.p, G.nP, G.C.p, G.nC, G.pc, la, &M, &G.nM, G.P2, &nP2, G.C2,
                 &nC2, &G.vb, nullptr, nullptr, G.cs, premoved))
        return -1;
    std::swap(G.P, G.P2);
    std::swap(G.C, G.C2);
    G.nP = nP2;
    G.nC = nC2;
    *n_layer = G.nM;
    return 0;
}

extern "C" void *cge_host_alloc(size_t bytes)
{
    if (eng_init()) return nullptr;
    void *p = nullptr;
    if (hipHostMalloc(&p, bytes ? bytes : 8, hipHostMallocDefault) != hipSuccess) {
        fail_msg("pinned host allocation failed");
        return nullptr;
    }
    return p;
}

extern "C" void cge_host_free(void *p)
{
    if (p) (void)hipHostFree(p);
}

extern "C" int cge_gen_fetch(uint64_t *host, size_t n)
{
    if (n > G.nM) return fail_msg("cge_gen_fetch: too many boards");
    if (n) HC(hipMemcpy(host, G.M.p, n * 8, hipMemcpyDeviceToHost));
    return 0;
}

extern "C" void cge_gen_end(void)
{
    counts_finish(G.pc, G.C.p);
    buf_free(G.P);
    buf_free(G.C);
    buf_free(G.P2);
    buf_free(G.C2);
    buf_free(G.M);
}

struct SolveLayerBuf {
    Buf keys, vals, dir;
    size_t n = 0;
    LayerIndexPlan plan;
};

static struct {
    cg_rules r;
    WinRules w;
    double p4 = 0.1;
    SolveLayerBuf L[3];   /* rotating: cur, s+2, s+4 */
    int i2 = -1, i4 = -1;
} S;

static LayerView sview(int i)
{
    if (i < 0) return empty_view();
    const SolveLayerBuf &b = S.L[i];
    return make_view(b.keys.p, (const double *)b.vals.p, b.n, (const uint32_t *)b.dir.p, b.plan);
}

extern "C" int cge_solve_begin(const cge_rules *r, const uint64_t *win, size_t n_images, double p4)
{
    if (eng_init()) return -1;
    S.r = to_rules(r);
    S.p4 = p4;
    S.i2 = S.i4 = -1;
    return upload_win(win, n_images, S.r, &S.w);
}

extern "C" int cge_solve_step(const uint64_t *keys, size_t n, double *values)
{
    int cur = 0;
    while (cur == S.i2 || cur == S.i4) cur++;
    SolveLayerBuf &b = S.L[cur];
    if (ensure(b.keys, n + 1) || ensure(b.vals, n + 1)) return -1;
    b.n = n;
    if (n) HC(hipMemcpyAsync(b.keys.p, keys, n * 8, hipMemcpyHostToDevice, E.st));
    uint64_t o = 0, a = ~0ULL;
    if (n && key_or_and(b.keys.p, n, &o, &a)) return -1;
    b.plan = plan_index(n, o, a);
    if (ensure(b.dir, b.plan.dir_entries / 2 + 1)) return -1;
    if (build_index(b.keys.p, n, b.plan, (uint32_t *)b.dir.p)) return -1;
    if (solve_layer_dev(S.r, b.keys.p, n, (double *)b.vals.p, nullptr, sview(S.i2), sview(S.i4), S.w,
                        S.p4))
        return -1;
    if (n) HC(hipMemcpyAsync(values, b.vals.p, n * 8, hipMemcpyDeviceToHost, E.st));
    if (check_sync(E.st, "cge_solve_step")) return -1;
    S.i4 = S.i2;
    S.i2 = cur;
    return 0;
}

extern "C" void cge_solve_end(void)
{
    for (auto &l : S.L) {
        buf_free(l.keys);
        buf_free(l.vals);
        buf_free(l.dir);
        l.n = 0;
    }
    S.i2 = S.i4 = -1;
}

/* ------------------------------------------------------------------------ */
/* batched tables                                                            */
/* ------------------------------------------------------------------------ */
/* Many small tables share every kernel launch: per layer step one move, one
 * sort/dedupe, one spawn, the carry sort and one directory build serve all the
 * tables (composite keys carry the table id), with two host
 * synchronisations per step whatever the number of tables.  The solve of a
 * step is one kernel over a per-table descriptor array; the whole downward
 * sweep is captured once in a hipGraph and replayed. */

struct cg_batch {
    cg_rules r;
    struct Tab {
        std::vector<uint64_t> init;
        int start = 0, end = 0, nsteps = 0;
        std::vector<uint64_t> wimg;
        BatchTable dev;
    };
    std::vector<Tab> tabs;
    Arena arena;
    BatchTable *d_tabs = nullptr;
    struct Step {
        uint64_t *keys = nullptr;
        double *vals = nullptr;
        uint32_t *dir = nullptr;
        size_t n = 0;
        std::vector<size_t> off;            /* ntab + 1 */
        std::vector<LayerIndexPlan> plan;   /* per table */
        std::vector<size_t> dir_off;        /* per table */
        std::vector<BatchIndexPlan> bplan;  /* per table (source of an asynchronous copy) */
    };
    std::vector<Step> steps;
    int K = 0;
    Buf P, C, P2, C2, M;
    unsigned long long *d_tstat = nullptr;
    std::vector<unsigned long long> h_tstat;
    /* solve */
    BSolveDesc *d_desc = nullptr;
    std::vector<size_t> tiles;
    uint64_t *d_wmask = nullptr, *d_wval = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    bool solved = false;
    double graph_p4 = -1;
    /* generation: the layer step's graph, and the per-table statistics in
     * mapped host memory (3 words per table) */
    hipGraphExec_t gstep = nullptr;
    unsigned long long *h_tpub = nullptr, *hd_tpub = nullptr;
    size_t tpub_cap = 0;
};

extern "C" cg_batch *cg_batch_create(const cg_rules *rules)
{
    if (eng_init()) return nullptr;
    if (rules->free_formation) {
        fail_msg("batched tables need free_formation = false (formation cells must stay put)");
        return nullptr;
    }
    if (rules->premove) {
        fail_msg("batched tables do not take premove (run such a table alone)");
        return nullptr;
    }
    cg_batch *b = new cg_batch();
    b->r = *rules;
    return b;
}

static void batch_drop_graph(cg_batch *b)
{
    if (b->gexec) hipGraphExecDestroy(b->gexec);
    if (b->graph) hipGraphDestroy(b->graph);
    b->gexec = nullptr;
    b->graph = nullptr;
}

extern "C" void cg_batch_free(cg_batch *b)
{
    if (!b) return;
    hipStreamSynchronize(E.st);
    batch_drop_graph(b);
    if (b->gstep) hipGraphExecDestroy(b->gstep);
    if (b->h_tpub) hipHostFree(b->h_tpub);
    b->arena.release();
    buf_free(b->P);
    buf_free(b->C);
    buf_free(b->P2);
    buf_free(b->C2);
    buf_free(b->M);
    if (b->d_tabs) hipFree(b->d_tabs);
    if (b->d_tstat) hipFree(b->d_tstat);
    delete b;
}

extern "C" int cg_batch_add(cg_batch *b, const uint64_t *initial, size_t n, int start, int end,
                            const uint64_t *winstates, size_t n_win)
{
    if (!n) return fail_msg("cg_batch_add: no initial boards");
    if (b->tabs.size() >= (1u << (64 - BATCH_TID_SHIFT))) return fail_msg("cg_batch_add: too many tables");
    if (end < start) return fail_msg("cg_batch_add: end < start");
    cg_batch::Tab t;
    uint64_t c0 = cg_canon(initial[0]);
    uint64_t fnib = cg_fmask(c0) * 0xf;
    int nfree = 0;
    for (int x = 0; x < 16; x++)
        if (!((fnib >> (60 - 4 * x)) & 0xf)) t.dev.pos[nfree++] = (uint8_t)x;
    if (nfree * 4 > BATCH_TID_SHIFT)
        return fail_msg("cg_batch_add: more than 12 free cells do not fit a composite key");
    t.dev.fnib = fnib;
    t.dev.nfree = nfree;
    {
        /* runs of free-cell bits: at most 4 lets xform pack them directly */
        uint64_t free_bits = ~fnib;
        int runs = 0;
        for (int bit = 63; bit >= 0; bit--)
            if (((free_bits >> bit) & 1) && (bit == 63 || !((free_bits >> (bit + 1)) & 1))) runs++;
        t.dev.runs = free_bits && runs <= 4;
        t.dev.fx = make_xform(t.dev.runs ? free_bits : 0);
    }
    t.start = start;
    t.end = end;
    t.nsteps = (end - start) / 2 + 1;
    t.dev.nsteps = t.nsteps;
    for (size_t i = 0; i < n; i++) {
        uint64_t c = cg_canon(initial[i]);
        if ((c & fnib) != fnib) return fail_msg("cg_batch_add: initial boards differ in formation");
        if (t.dev.runs && (t.dev.fnib | xform_inv(xform(c, t.dev.fx), t.dev.fx)) != c)
            return fail_msg("cg_batch_add: free-cell packing does not round-trip");
        t.init.push_back(c);
    }
    win_images(winstates, n_win, t.wimg);
    b->tabs.push_back(t);
    b->solved = false;
    return (int)b->tabs.size() - 1;
}

struct BatchMAlloc : LayerAlloc {
    Buf *buf;
    explicit BatchMAlloc(Buf *x) : buf(x) {}
    uint64_t *alloc(size_t n) override
    {
        if (ensure(*buf, n + 1)) return nullptr;
        return buf->p;
    }
};

static uint64_t host_benc(uint64_t c, const BatchTable &T, uint64_t tid)
{
    uint64_t k = 0;
    for (int j = 0; j < T.nfree; j++) k = (k << 4) | ((c >> (60 - 4 * T.pos[j])) & 0xf);
    return (tid << BATCH_TID_SHIFT) | k;
}

extern "C" int cg_batch_generate(cg_batch *b)
{
    if (eng_init()) return -1;
    const int ntab = (int)b->tabs.size();
    if (!ntab) return fail_msg("cg_batch_generate: no tables");
    HC(hipStreamSynchronize(E.st));
    batch_drop_graph(b);
    b->arena.reset();
    b->steps.clear();
    b->solved = false;
    b->d_desc = nullptr;
    const uint16_t *lut = get_lut(b->r.ignore_f);
    if (!lut) return fail_msg("row table upload failed");
    /* device table descriptors and per-table statistics */
    if (b->d_tabs) hipFree(b->d_tabs);
    if (b->d_tstat) hipFree(b->d_tstat);
    std::vector<BatchTable> ht(ntab);
    b->K = 0;
    for (int t = 0; t < ntab; t++) {
        ht[t] = b->tabs[t].dev;
        b->K = std::max(b->K, b->tabs[t].nsteps);
    }
    HC(hipMalloc((void **)&b->d_tabs, ntab * sizeof(BatchTable)));
    HC(hipMemcpy(b->d_tabs, ht.data(), ntab * sizeof(BatchTable), hipMemcpyHostToDevice));
    HC(hipMalloc((void **)&b->d_tstat, 3 * ntab * sizeof(unsigned long long)));
    b->h_tstat.assign(3 * ntab, 0);
    /* pre-move layer 0: canonical initial boards as composite keys */
    std::vector<uint64_t> p0;
    for (int t = 0; t < ntab; t++)
        for (uint64_t c : b->tabs[t].init) p0.push_back(host_benc(c, b->tabs[t].dev, (uint64_t)t));
    if (ensure(b->P, p0.size() + 1)) return -1;
    HC(hipMemcpy(b->P.p, p0.data(), p0.size() * 8, hipMemcpyHostToDevice));
    size_t nP = p0.size(), nC = 0;
    StepCounts pc;
    std::vector<uint64_t> init_tstat(3 * ntab);
    for (int t = 0; t < ntab; t++) init_tstat[3 * t + 2] = ~0ULL;
    int maxfree = 0;
    for (int t = 0; t < ntab; t++) maxfree = std::max(maxfree, b->tabs[t].dev.nfree);
    /* cell bounds of the composite keys: the table id's nibbles take the
     * values of 0..ntab-1; a free cell holds at most the largest tile of its
     * table's tile sum (ValueBound) when every table's initial boards share
     * one tile sum and hold no 0xf outside the formation (0xf never enters a
     * free cell: a formation cell would be vacated, which the kernels
     * report) */
    uint64_t tid_max = 0;
    for (int j = 0; j < 4; j++) {
        const int lo = BATCH_TID_SHIFT + 4 * j;
        const uint64_t rest = (uint64_t)(ntab - 1) >> (4 * j);
        tid_max |= (rest > 15 ? 15ULL : rest) << lo;
    }
    long sum_max = 0;
    bool bounded = true;
    for (int t = 0; t < ntab && bounded; t++) {
        const cg_batch::Tab &tab = b->tabs[t];
        ValueBound v = value_bound_of(tab.init.data(), tab.init.size());
        for (uint64_t c : tab.init)
            for (int x = 0; x < 16; x++)
                if (!((tab.dev.fnib >> (60 - 4 * x)) & 15) && ((c >> (60 - 4 * x)) & 15) == 15) bounded = false;
        if (!v.ok) bounded = false;
        /* the sum reached by the carry of the table's last step */
        sum_max = std::max(sum_max, v.s + 2L * (tab.nsteps - 1) + 4);
    }
    uint64_t batch_cells = CELLS_UNBOUNDED;
    if (bounded && sum_max >= 1 && tile_log2(sum_max) <= 14)
        batch_cells = tid_max | (cells_bounded_by(tile_log2(sum_max)) >> (64 - BATCH_TID_SHIFT));
    if (b->tpub_cap < (size_t)3 * ntab) {
        if (b->h_tpub) hipHostFree(b->h_tpub);
        b->h_tpub = nullptr;
        b->tpub_cap = 0;
        HC(hipHostMalloc((void **)&b->h_tpub, 3 * ntab * sizeof(unsigned long long),
                         hipHostMallocMapped | hipHostMallocCoherent));
        HC(hipHostGetDevicePointer((void **)&b->hd_tpub, b->h_tpub, 0));
        b->tpub_cap = 3 * ntab;
    }
    /* One host synchronisation per step.  Everything a step enqueues is one
     * graph launch (knob batch_graph), which also carries the work the
     * previous step left: the copy of the carry C into the pre-move boards,
     * the carry sort of the previous spawn-4 boards on the second stream (a
     * branch of the graph, joined before the spawns) and the directories of
     * the previous layer.  Sizes the host cannot know before the launch (the
     * candidate count, |M|, the spawn count) are read on the device: the M
     * sort is the sample sort sized for 4 |P| candidates, the layer and the
     * spawn buffers are allocated for bounds. */
    struct Deferred {
        /* the carry sort of the previous step's spawn-4 boards */
        bool carry = false;
        const uint64_t *c_in = nullptr;
        size_t c_n = 0;
        uint64_t c_vary = 0;
        /* the copy of the carry before it into the pre-move boards */
        const uint64_t *cp_src = nullptr;
        uint64_t *cp_dst = nullptr;
        size_t cp_n = 0;
        /* this step's launch carried a carry sort (settled at its read) */
        bool carry_done = false;
        /* the directories of the previous layer */
        int dir_step = -1;
        size_t dir_entries = 0;
        const size_t *d_seg = nullptr;
        const BatchIndexPlan *d_plan = nullptr;
    } df;
    /* enqueues the deferred work (main stream, and the carry sort forked to
     * the second one) */
    /* carry buffers: the carry sort writes cbuf[cw]; the carry copied into
     * the pre-move boards is in the other */
    Buf *cbuf[2] = { &b->C, &b->C2 };
    int cw = 0;
    /* carry_sort 0: a step's spawn-4 boards are moved two steps later as
     * written; prev_s4 is the count of the previous step's */
    const bool carry_sorted = g_carry_sort != 0;
    size_t prev_s4 = 0;
    auto enqueue_deferred = [&](void) -> int {
        if (df.cp_n) launch_copy_and_u64(E.st, df.cp_dst, df.cp_src, df.cp_n, ~0ULL);
        if (df.dir_step >= 0) {
            const cg_batch::Step &ps = b->steps[df.dir_step];
            launch_bdir(E.st, ps.keys, ps.n, df.d_seg, ntab, df.d_plan, df.dir_entries, ps.dir, E.scratch);
        }
        if (df.carry) {
            if (g_carry_lane == 1 && fork_lane1()) return -1;
            int hu4;
            if (sort_enqueue(const_cast<uint64_t *>(df.c_in), df.c_n, df.c_vary, cbuf[cw]->p, 6, &hu4,
                             g_carry_lane, batch_cells, SORT_SAMPLE_SPARSE))
                return -1;
            pc.hu4 = hu4;
        }
        hipError_t e = kernels_take_error();
        return e != hipSuccess ? fail("batch step", e) : 0;
    };
    for (int k = 0; k < b->K; k++) {
        cg_batch::Step st;
        const size_t cmax = 4 * nP;
        const bool dev = g_one_sync && cmax <= sort_sample_dev_limit();
        const bool graph = dev && g_batch_graph;
        /* buffers and allocations first: nothing may allocate inside a
         * capture */
        /* the carry this step's spawns sit before */
        const size_t cbound = carry_sorted ? (df.carry ? df.c_n : 0) : prev_s4;
        size_t cap = (size_t)maxfree * spawn_rows_estimate(pc, cmax);
        Buf &S4 = E.S4[pc.s4];
        st.keys = (uint64_t *)b->arena.alloc(std::max<size_t>(cmax, 1) * 8);
        if (!st.keys) return fail_msg("batch layer allocation failed");
        if (ensure(E.A, cmax + 1) || ensure(b->M, cmax + 1) || ensure(b->P2, cap + cbound + 1) ||
            ensure(S4, cap + 1) || (carry_sorted && ensure(*cbuf[cw], cbound + 1)))
            return -1;
        {
            size_t scr = sort_scratch_bytes(cmax);
            if (df.dir_step >= 0) scr = std::max(scr, dir_scratch_bytes(df.dir_entries + 1));
            if (ensure_scratch_bytes(scr)) return -1;
        }
        if (dev) {
            const size_t pk = sort_part_keys(cmax, SORT_SAMPLE);
            /* look-back words for the sample sort's most buckets, epochs
             * for its launch */
            if ((pk && ensure(E.Pt, pk)) || sort_reserve(E.st, 0, (size_t)sort_sample_bmax(), 4)) return -1;
        }
        if (df.carry) {
            /* the carry sort's buffers and look-back words (second stream) */
            const int lane = g_carry_lane;
            const size_t pk = sort_part_keys(df.c_n, SORT_SAMPLE_SPARSE);
            if (lane == 1 ? (ensure(E.B2, df.c_n + 1) || ensure_scratch2(df.c_n) || (pk && ensure(E.Pt2, pk)))
                          : (ensure(E.B, df.c_n + 1) || (pk && ensure(E.Pt, pk))))
                return -1;
            if (sort_prepare(lane == 1 ? E.st2 : E.st, lane, df.c_n, df.c_vary, batch_cells, SORT_SAMPLE_SPARSE))
                return fail_msg("sort: look-back words");
        }
        if (graph && capture_begin()) return -1;
        int rc = enqueue_deferred();
        int hm = -1;
        if (!rc) {
            launch_init_stats(E.st, E.d_cnt);
            hm = prof_begin(PK_MOVE);
            launch_bmove(E.st, b->P.p, nP, nullptr, E.A.p, E.d_cnt, lut, move_rules(b->r), b->d_tabs, k, E.d_err);
            prof_end(hm, 0);
        }
        size_t ncand = 0;
        uint64_t mor = 0, mand = ~0ULL;
        int hu = -1;
        bool carry_ov = false;
        if (!rc && dev) {
            rc = sort_enqueue_dev(E.A.p, E.d_cnt, cmax, b->M.p, 7, &hu);
        } else if (!rc) {
            /* the bound is beyond the sample sort: the candidate count first */
            if (df.carry && g_carry_lane == 1) rc = join_lane1();
            if (!rc) rc = read_counters(0, 3);
            if (!rc && E.oversize) {
                /* only the carry sort can have overflowed; it is repeated
                 * with the counts of the step below */
                if (E.oversize1) rc = fail_msg("sort: unexpected oversize");
                carry_ov = E.oversize && E.oversize2;
                E.oversize = false;
            }
            if (!rc) {
                ncand = E.h_cnt[0];
                mor = E.h_cnt[1];
                mand = E.h_cnt[2];
                rc = sort_enqueue(E.A.p, ncand, vary_of(ncand, mor, mand), b->M.p, 7, &hu, 0, batch_cells);
            }
        }
        const size_t mrows = dev ? cmax : ncand;
        int hs = -1;
        auto enqueue_decode_spawn = [&](size_t capk) {
            launch_tstat_init(E.st, b->d_tstat, ntab);
            launch_bdecode(E.st, b->M.p, mrows, E.d_cnt + 7, st.keys, b->d_tabs, b->d_tstat);
            launch_init_stats(E.st, E.d_cnt + 8);
            hs = prof_begin(PK_SPAWN);
            launch_bspawn(E.st, b->M.p, mrows, E.d_cnt + 7, b->P2.p, S4.p, capk, E.d_cnt + 8, b->d_tabs, k,
                          E.d_err);
            prof_end(hs, 0);
            launch_copy_and_u64(E.st, (uint64_t *)b->hd_tpub, (const uint64_t *)b->d_tstat, 3 * ntab, ~0ULL);
        };
        if (!rc) {
            enqueue_decode_spawn(cap);
            if (df.carry && g_carry_lane == 1) rc = join_lane1();
        }
        if (!rc) publish_counters(0, 13);
        if (graph && capture_end_launch(&b->gstep, rc == 0)) return -1;
        if (rc || collect_counters(0, 13)) return -1;
        /* the counts of the step (and of the carry sort it carried) */
        size_t nM = 0, sumE = 0;
        bool settled = false;
        for (int attempt = 0; attempt < 4; attempt++) {
            const bool ov_m = E.oversize && E.oversize1, ov_c = (E.oversize && E.oversize2) || carry_ov;
            E.oversize = false;
            carry_ov = false;
            if (attempt == 0 && dev) {
                ncand = E.h_cnt[0];
                mor = E.h_cnt[1];
                mand = E.h_cnt[2];
            }
            if (df.carry) {
                if (ov_c && carry_repeat(pc, cbuf[cw]->p)) return -1;
                nC = E.h_cnt[6];
                counts_arrived(pc, nC);
                df.carry = false;
                df.carry_done = true;
            } else if (ov_c) {
                return fail_msg("sort: unexpected oversize");
            }
            if (!ov_m) {
                nM = E.h_cnt[7];
                sumE = E.h_cnt[8];
                if (sumE <= cap) {
                    settled = true;
                    break;
                }
                cap = (size_t)maxfree * nM;   /* exact bound, always enough */
                if (ensure(b->P2, cap + cbound + 1) || ensure(S4, cap + 1)) return -1;
            } else if (sort_enqueue(E.A.p, ncand, vary_of(ncand, mor, mand), b->M.p, 7, &hu, 0, batch_cells,
                                    SORT_LSD)) {
                /* the M sort had an oversize bucket: repeated on the LSD path */
                return -1;
            }
            enqueue_decode_spawn(cap);
            if (read_counters(0, 13)) return -1;
        }
        if (!settled) return fail_msg("batch step: spawns or sort not settled");
        pc.last_nM = nM;
        prof_set_bytes(hm, 8.0 * nP + 8.0 * ncand);
        prof_set_bytes(hu, 8.0 * ncand + 8.0 * nM);
        prof_set_bytes(hs, 8.0 * nM + 16.0 * sumE);
        b->arena.shrink_last(st.keys, std::max<size_t>(cmax, 1) * 8, std::max<size_t>(nM, 1) * 8);
        for (int t = 0; t < ntab; t++)
            for (int j = 0; j < 3; j++) b->h_tstat[3 * t + j] = ((volatile unsigned long long *)b->h_tpub)[3 * t + j];
        const uint64_t or4 = E.h_cnt[11], and4 = E.h_cnt[12];
        /* directories of every table of this step: planned now, built at the
         * head of the next launch */
        st.n = nM;
        st.off.assign(ntab + 1, 0);
        st.plan.resize(ntab);
        st.dir_off.resize(ntab);
        st.bplan.resize(ntab);
        size_t entries = 0;
        for (int t = 0; t < ntab; t++) {
            size_t nt = b->h_tstat[3 * t];
            st.off[t + 1] = st.off[t] + nt;
            st.plan[t] = plan_index(nt, b->h_tstat[3 * t + 1], b->h_tstat[3 * t + 2]);
            st.dir_off[t] = entries;
            st.bplan[t].x = st.plan[t].x;
            st.bplan[t].shift = st.plan[t].shift;
            st.bplan[t].dir_base = entries;
            entries += st.plan[t].dir_entries;
        }
        if (st.off[ntab] != nM) return fail_msg("batch: per-table counts disagree with the layer");
        if (check_layer_size(nM)) return -1;
        st.dir = (uint32_t *)b->arena.alloc((entries + 1) * 4);
        size_t *d_seg = (size_t *)b->arena.alloc((ntab + 1) * sizeof(size_t));
        BatchIndexPlan *d_plan = (BatchIndexPlan *)b->arena.alloc(ntab * sizeof(BatchIndexPlan));
        if (!st.dir || !d_seg || !d_plan) return fail_msg("batch index allocation failed");
        /* the host arrays live in the step, so the asynchronous copies may
         * read them later; they precede the next launch on the stream */
        HC(hipMemcpyAsync(d_seg, st.off.data(), (ntab + 1) * sizeof(size_t), hipMemcpyHostToDevice, E.st));
        HC(hipMemcpyAsync(d_plan, st.bplan.data(), ntab * sizeof(BatchIndexPlan), hipMemcpyHostToDevice, E.st));
        b->steps.push_back(std::move(st));
        if (!carry_sorted) {
            /* P_{k+1} = S2 u the previous step's spawn-4 boards as written
             * (the spawn-4 buffers alternate between steps); the next
             * launch carries this layer's directories */
            if (prev_s4)
                HC(hipMemcpyAsync(b->P2.p + sumE, E.S4[pc.s4 ^ 1].p, prev_s4 * sizeof(uint64_t),
                                  hipMemcpyDeviceToDevice, E.st));
            df.carry = false;
            df.cp_n = 0;
            df.dir_step = k;
            df.dir_entries = entries;
            df.d_seg = d_seg;
            df.d_plan = d_plan;
            pc.s4 ^= 1;
            const size_t n2 = sumE + prev_s4;
            prev_s4 = sumE;
            std::swap(b->P, b->P2);
            nP = n2;
            nC = 0;
            continue;
        }
        /* what the next launch carries: P_{k+1} = S2 u C (the carry before
         * this step's, now settled), the carry sort of this step's spawn-4
         * boards, this layer's directories */
        const bool had_carry = df.carry_done;
        df.carry_done = false;
        df.cp_src = had_carry ? cbuf[cw]->p : nullptr;
        df.cp_dst = b->P2.p + sumE;
        df.cp_n = had_carry ? nC : 0;
        if (had_carry) cw ^= 1;
        df.carry = true;
        df.c_in = S4.p;
        df.c_n = sumE;
        df.c_vary = vary_of(sumE, or4, and4);
        df.dir_step = k;
        df.dir_entries = entries;
        df.d_seg = d_seg;
        df.d_plan = d_plan;
        pc.pending = true;
        pc.sumE = sumE;
        pc.cin = S4.p;
        pc.vary4 = df.c_vary;
        pc.cell4 = batch_cells;
        pc.s4 ^= 1;
        const size_t n2 = sumE + df.cp_n;
        std::swap(b->P, b->P2);
        nP = n2;
        nC = 0;
    }
    /* the last layer's directories (its carry sort has no use) */
    df.carry = false;
    df.cp_n = 0;
    pc.pending = false;
    if (enqueue_deferred()) return -1;
    return check_sync(E.st, "cg_batch_generate");
}

static LayerView batch_view(const cg_batch *b, int t, int layer)
{
    if (layer < 0 || layer >= b->tabs[t].nsteps) return empty_view();
    const cg_batch::Step &st = b->steps[layer];
    if (!st.vals) return empty_view();
    size_t nt = st.off[t + 1] - st.off[t];
    return make_view(st.keys, st.vals, nt, st.dir + st.dir_off[t], st.plan[t]);
}

extern "C" int cg_batch_solve(cg_batch *b, double p4, int use_graph)
{
    if (eng_init()) return -1;
    const int ntab = (int)b->tabs.size();
    if ((int)b->steps.size() != b->K || !ntab) return fail_msg("cg_batch_solve needs cg_batch_generate");
    const uint16_t *lut = get_lut(b->r.ignore_f);
    if (!lut) return fail_msg("row table upload failed");
    if (!b->d_desc) {
        /* values, winstate images and the per-step descriptors, once per generation */
        for (auto &st : b->steps) {
            st.vals = (double *)b->arena.alloc(std::max<size_t>(st.n, 1) * 8);
            if (!st.vals) return fail_msg("batch value allocation failed");
        }
        std::vector<uint64_t> wm, wv;
        std::vector<size_t> woff(ntab + 1, 0);
        for (int t = 0; t < ntab; t++) {
            for (uint64_t w : b->tabs[t].wimg) {
                wm.push_back(cg_win_mask(w));
                wv.push_back(w);
            }
            woff[t + 1] = wm.size();
        }
        b->d_wmask = (uint64_t *)b->arena.alloc(std::max<size_t>(wm.size(), 1) * 8);
        b->d_wval = (uint64_t *)b->arena.alloc(std::max<size_t>(wv.size(), 1) * 8);
        if (!wm.empty()) {
            HC(hipMemcpy(b->d_wmask, wm.data(), wm.size() * 8, hipMemcpyHostToDevice));
            HC(hipMemcpy(b->d_wval, wv.data(), wv.size() * 8, hipMemcpyHostToDevice));
        }
        std::vector<BSolveDesc> all((size_t)b->K * ntab);
        b->tiles.assign(b->K, 0);
        for (int k = 0; k < b->K; k++) {
            size_t tile0 = 0;
            for (int t = 0; t < ntab; t++) {
                BSolveDesc &d = all[(size_t)k * ntab + t];
                memset(&d, 0, sizeof d);
                int layer = b->tabs[t].nsteps - 1 - k;   /* each table from its own top */
                d.tile0 = tile0;
                d.win.mask = b->d_wmask + woff[t];
                d.win.val = b->d_wval + woff[t];
                d.win.n = (int)(woff[t + 1] - woff[t]);
                d.win.nox = b->r.nox;
                d.win.score = b->r.score;
                if (layer < 0) continue;
                const cg_batch::Step &st = b->steps[layer];
                d.M = st.keys + st.off[t];
                d.n = st.off[t + 1] - st.off[t];
                d.V = st.vals + st.off[t];
                d.t2 = batch_view(b, t, layer + 1);
                d.t4 = batch_view(b, t, layer + 2);
                tile0 += (d.n + 255) / 256;
            }
            b->tiles[k] = tile0;
        }
        b->d_desc = (BSolveDesc *)b->arena.alloc(all.size() * sizeof(BSolveDesc));
        if (!b->d_desc) return fail_msg("batch descriptor allocation failed");
        HC(hipMemcpy(b->d_desc, all.data(), all.size() * sizeof(BSolveDesc), hipMemcpyHostToDevice));
    }
    if (!use_graph) {
        for (int k = 0; k < b->K; k++) {
            int h = prof_begin(PK_SOLVE);
            launch_bsolve(E.st, b->d_desc + (size_t)k * ntab, ntab, b->tiles[k], p4, lut,
                          b->r.free_formation);
            prof_end(h, 0);
        }
    } else {
        if (!b->gexec || b->graph_p4 != p4) {
            batch_drop_graph(b);
            HC(hipStreamSynchronize(E.st));
            HC(hipStreamBeginCapture(E.st, hipStreamCaptureModeThreadLocal));
            for (int k = 0; k < b->K; k++)
                launch_bsolve(E.st, b->d_desc + (size_t)k * ntab, ntab, b->tiles[k], p4, lut,
                              b->r.free_formation);
            hipError_t e = hipStreamEndCapture(E.st, &b->graph);
            if (e != hipSuccess) return fail("hipStreamEndCapture", e);
            HC(hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
            b->graph_p4 = p4;
        }
        HC(hipGraphLaunch(b->gexec, E.st));
    }
    b->solved = true;
    return check_sync(E.st, "cg_batch_solve");
}

extern "C" int cg_batch_num_tables(const cg_batch *b) { return (int)b->tabs.size(); }
extern "C" int cg_batch_num_layers(const cg_batch *b, int t)
{
    return (t >= 0 && t < (int)b->tabs.size()) ? b->tabs[t].nsteps : -1;
}

extern "C" size_t cg_batch_layer_size(const cg_batch *b, int t, int layer)
{
    if (t < 0 || t >= (int)b->tabs.size() || layer < 0 || layer >= b->tabs[t].nsteps ||
        layer >= (int)b->steps.size())
        return 0;
    return b->steps[layer].off[t + 1] - b->steps[layer].off[t];
}

extern "C" size_t cg_batch_total_boards(const cg_batch *b)
{
    size_t s = 0;
    for (auto &st : b->steps) s += st.n;
    return s;
}

extern "C" int cg_batch_copy_layer(const cg_batch *b, int t, int layer, uint64_t *keys, double *values)
{
    if (t < 0 || t >= (int)b->tabs.size() || layer < 0 || layer >= (int)b->steps.size() ||
        layer >= b->tabs[t].nsteps)
        return fail_msg("cg_batch_copy_layer: index out of range");
    const cg_batch::Step &st = b->steps[layer];
    size_t o = st.off[t], n = st.off[t + 1] - o;
    if (keys && n) HC(hipMemcpy(keys, st.keys + o, n * 8, hipMemcpyDeviceToHost));
    if (values && n) {
        if (!b->solved) return fail_msg("cg_batch_copy_layer: not solved");
        HC(hipMemcpy(values, st.vals + o, n * 8, hipMemcpyDeviceToHost));
    }
    return 0;
}
