/*
 * sort.hip -- onesweep LSD radix sort + unique compaction of u64 keys (gfx950).
 *
 * The bits that differ between keys (OR & ~AND of the key set, produced by
 * the kernels that wrote the keys) are packed into a dense key on the fly
 * (an order-preserving bijection on the key set), so a DPDGap layer with 40
 * varying bits spread over bits 0..51 sorts in 5 passes of 8 bits, or 4 of
 * 10.  When the caller bounds the cell values (tiles of a layer of tile sum s
 * are at most 2^floor(log2 s)), pairs of varying cells are coded as
 * a * base + b: the DPDGap layer's 10 cells of values 0..10 then take 35 bits,
 * 4 passes of 9.  One histogram kernel computes every pass's digit counts in a single
 * read; each pass is one kernel whose tiles take their global offsets from a
 * chained (decoupled look-back) scan over per-tile digit counts; the unique
 * compaction is one more chained-scan kernel.
 *
 * Digit width is a template parameter (8, 9 or 10 bits): a pass block has one
 * thread per digit (256, 512 or 1024 threads), ranks its keys with one wave
 * ballot per digit bit (match-any), stages the tile in LDS in digit order and
 * writes each digit's run contiguously.  Keys per thread sets the tile
 * (4096 to 16384 keys).  Which (digit width, tile, look-back window) a sort
 * uses is looked up per (size class, number of varying bits) in a table the
 * autotuner fills; the default is 8 bits, 8192-key tiles.
 *
 * Look-back flags live in a sorter-owned status array: each word holds a
 * state (aggregate / inclusive prefix), the epoch of the launch that wrote it
 * and the value, so a launch ignores words left by earlier launches and the
 * array is cleared only when it grows or the epoch counter wraps.  A
 * look-back wait is bounded: past its budget it sets an error flag and the
 * kernel finishes; the host reports a failure.
 */
#include <hip/hip_runtime.h>
#include <algorithm>
#include <string.h>
#include "kernels.h"
#include "devutil.h"

namespace cg {

constexpr int OS_MAXPASS = 8;
constexpr int OS_MAXR = 1024;
constexpr int OS_VALBITS = 42, OS_EPOCHBITS = 20;
constexpr uint64_t OS_VAL = (1ULL << OS_VALBITS) - 1;
constexpr uint64_t OS_AGG = 1ULL << 62, OS_INC = 2ULL << 62;
constexpr uint32_t OS_EPOCH_MAX = (1u << OS_EPOCHBITS) - 1;
constexpr int OS_SPIN_LIMIT = 1 << 22;
constexpr int UQ_THREADS = 256, UQ_KPT = 32, UQ_TILE = UQ_THREADS * UQ_KPT;

__device__ __forceinline__ uint64_t ld_flag(const uint64_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void st_flag(uint64_t *p, uint64_t state, uint32_t epoch, uint64_t v)
{
    __hip_atomic_store(p, state | ((uint64_t)epoch << OS_VALBITS) | v, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}

/* 0 (not published by this launch), 1 (aggregate) or 2 (inclusive prefix) */
__device__ __forceinline__ uint32_t fstate(uint64_t f, uint32_t epoch)
{
    return (uint32_t)((f >> OS_VALBITS) & OS_EPOCH_MAX) == epoch ? (uint32_t)(f >> 62) : 0u;
}

/* waits until the flag at p is published by this launch; bounded */
__device__ __forceinline__ uint64_t wait_flag(const uint64_t *p, uint32_t epoch, unsigned *err)
{
    uint64_t v;
    int spins = 0;
    while (fstate(v = ld_flag(p), epoch) == 0) {
        if (++spins > OS_SPIN_LIMIT) {
            atomicOr(err, 1u);
            return OS_INC | ((uint64_t)epoch << OS_VALBITS);
        }
        __builtin_amdgcn_s_sleep(2);
    }
    return v;
}

/* Sum for one chained-scan lane over the entries t_from, t_from - 1, ...,
 * lo (status of entry i at status[i * stride]), nearest first, up to and
 * including the first inclusive prefix (*done then set).  The flags are
 * coherent across XCDs, so every load is a long round trip and the loads
 * themselves cost bandwidth: the window grows from win0 by doubling up to
 * LB_W entries per step, and the walk waits only on the first entry that
 * has not published yet. */
template <int LB_W, int PF>
__device__ __forceinline__ uint64_t lookback_range(const uint64_t *status, size_t stride, long t_from,
                                                   long lo, uint32_t epoch, unsigned *err, int win0,
                                                   bool *done_out, const uint64_t (&pre)[PF],
                                                   bool use_pre)
{
    uint64_t excl = 0;
    long t = t_from;
    int win = win0;
    bool first = use_pre;
    *done_out = false;
    while (t >= lo) {
        uint64_t v[LB_W];
        /* the first PF entries of the first window may have been loaded
         * earlier (pre, same entries) */
#pragma unroll
        for (int j = 0; j < LB_W; j++)
            v[j] = (j < win && t - j >= lo)
                       ? ((j < PF && first) ? pre[j < PF ? j : 0] : ld_flag(status + (size_t)(t - j) * stride))
                       : 0;
        first = false;
        /* fully unrolled and predicated, so v stays in registers */
        int used = 0;
        bool blocked = false, done = false;
#pragma unroll
        for (int j = 0; j < LB_W; j++) {
            if (!blocked && !done && j < win && t - j >= lo) {
                uint32_t f = fstate(v[j], epoch);
                if (f == 0) {
                    blocked = true;
                } else {
                    excl += v[j] & OS_VAL;
                    used = j + 1;
                    done = f == 2;
                }
            }
        }
        if (done) {
            *done_out = true;
            return excl;
        }
        t -= used;
        if (!blocked) {
            win = win * 2 < LB_W ? win * 2 : LB_W;
            continue;
        }
        /* t is the nearest entry that has not published yet */
        uint64_t w = wait_flag(status + (size_t)t * stride, epoch, err);
        excl += w & OS_VAL;
        if (fstate(w, epoch) == 2) {
            *done_out = true;
            return excl;
        }
        t -= 1;
    }
    return excl;
}

/* Two-level look-back of a pass's digit lane.  Tiles form groups of
 * LB_GROUP; the last tile of a group publishes the group's status (gst):
 * its aggregate as soon as it has summed its group's tile aggregates, and the
 * inclusive prefix once known.  A tile walks its in-group predecessors, then
 * previous groups, so when all tiles start at once (a sort of a few hundred
 * tiles) the walk takes a few round trips instead of one per LB_W tiles.
 * Group statuses depend only on tile aggregates, which never wait, so every
 * wait ends.  Returns the tile's exclusive prefix. */
constexpr int LB_GROUP = 16;

/* loads the first in-group window of lookback_grouped (issued early, used
 * later: any snapshot of aggregates and inclusive prefixes is valid input) */
template <int PF>
__device__ __forceinline__ void lookback_prefetch(const uint64_t *tst, size_t stride, size_t tile,
                                                  uint64_t (&pre)[PF])
{
    const long g0 = (long)(tile / LB_GROUP) * LB_GROUP, t = (long)tile - 1;
#pragma unroll
    for (int j = 0; j < PF; j++) pre[j] = t - j >= g0 ? ld_flag(tst + (size_t)(t - j) * stride) : 0;
}

template <int LB_W, int PF>
__device__ __forceinline__ uint64_t lookback_grouped(const uint64_t *tst, uint64_t *gst, size_t stride,
                                                     size_t tile, uint64_t tot, uint32_t epoch,
                                                     unsigned *err, const uint64_t (&pre)[PF])
{
    const long g = (long)(tile / LB_GROUP), g0 = g * LB_GROUP;
    const bool gend = (long)tile - g0 == LB_GROUP - 1;
    bool done;
    uint64_t excl =
        lookback_range<LB_W, PF>(tst, stride, (long)tile - 1, g0, epoch, err, LB_W, &done, pre, true);
    if (gend) st_flag(gst + (size_t)g * stride, done ? OS_INC : OS_AGG, epoch, excl + tot);
    if (done) return excl;
    excl += lookback_range<LB_W, PF>(gst, stride, g - 1, 0, epoch, err, LB_W, &done, pre, false);
    if (gend) st_flag(gst + (size_t)g * stride, OS_INC, epoch, excl + tot);
    return excl;
}

/* entries of a pass's status array: tiles, then groups, per digit */
__host__ __device__ constexpr size_t pass_status_entries(size_t tiles, int bits)
{
    return (tiles + (tiles + LB_GROUP - 1) / LB_GROUP) << bits;
}

/* digit counts of every pass in one read of the keys, 16 waves per block.
 * Board keys are skewed (the top digits take few values), so lanes of one
 * wave often hit the same counter: NC copies of the counters indexed by
 * lane % NC, each starting in a different LDS bank, cap a wave's conflict on
 * one counter at 64 / NC lanes. */
constexpr int HIST_THREADS = 1024;

template <int BITS>
__global__ __launch_bounds__(HIST_THREADS) void k_os_hist(const uint64_t *__restrict__ keys, size_t n,
                                                 SortCode c, int npass, unsigned long long *ghist,
                                                 uint64_t *cval, unsigned *err)
{
    constexpr int R = 1 << BITS, NC = BITS == 8 ? 8 : (BITS == 9 ? 4 : 2);
    constexpr int STRIDE = OS_MAXPASS * R + 1;   /* 64 KiB of counters */
    __shared__ uint32_t h[NC * STRIDE];
    /* the bits outside the packed ones are the same in every key */
    if (blockIdx.x == 0 && threadIdx.x == 0) *cval = keys[0] & ~c.vmask;
    for (int i = threadIdx.x; i < NC * STRIDE; i += HIST_THREADS) h[i] = 0;
    __syncthreads();
    uint32_t *hc = h + (lane_id() % NC) * STRIDE;
    uint32_t bad = 0;   /* a cell outside its base (cell-pair code) */
    constexpr int HB = 8;   /* keys loaded per thread before counting */
    for (size_t b0 = (size_t)blockIdx.x * HIST_THREADS * HB; b0 < n;
         b0 += (size_t)gridDim.x * HIST_THREADS * HB) {
        uint64_t kk[HB];
#pragma unroll
        for (int j = 0; j < HB; j++) {
            size_t i = b0 + (size_t)j * HIST_THREADS + threadIdx.x;
            kk[j] = i < n ? keys[i] : 0;
        }
#pragma unroll
        for (int j = 0; j < HB; j++) {
            size_t i = b0 + (size_t)j * HIST_THREADS + threadIdx.x;
            if (i >= n) continue;
            uint64_t v = sc_encode(kk[j], c, bad);
            for (int p = 0; p < npass; p++) atomicAdd(&hc[p * R + ((v >> (BITS * p)) & (R - 1))], 1u);
        }
    }
    /* a cell value outside the code would be sorted wrongly */
    if (bad) atomicOr(err, SORT_ERR_CODE);
    __syncthreads();
    for (int i = threadIdx.x; i < npass * R; i += HIST_THREADS) {
        uint32_t t = 0;
#pragma unroll
        for (int k = 0; k < NC; k++) t += h[k * STRIDE + i];
        if (t) atomicAdd(&ghist[i], (unsigned long long)t);
    }
}

/* exclusive scan of every pass's digit counts (64-bit: a sort may hold
 * 2^32 keys or more) */
template <int BITS>
__global__ __launch_bounds__(1 << BITS) void k_os_scan(const unsigned long long *ghist, int npass,
                                                       uint64_t *gbase)
{
    constexpr int R = 1 << BITS;
    __shared__ uint64_t wtmp[17];
    for (int p = 0; p < npass; p++) {
        uint64_t tot;
        uint64_t ex = block_excl_scan<uint64_t>(ghist[p * R + threadIdx.x], &tot, wtmp);
        gbase[p * R + threadIdx.x] = ex;
    }
}

/* One LSD pass: a block of R = 2^BITS threads (thread d owns digit d) sorts
 * a tile of R * KPT keys by digit (k >> shift) & (R - 1), stably, where k is
 * the key's code.  The first pass (RAW) reads the original keys and encodes
 * them (sc_encode) once; every pass writes codes, so later passes extract a
 * digit with a shift and a mask. */
/* threads of a pass block: one per digit, except the 8-keys-per-thread
 * configurations, which use 1024 threads (the first 2^bits own the digits) so
 * that an 8192-key tile is spread over 16 waves */
constexpr int pass_threads(int bits, int kpt) { return kpt <= 8 ? 1024 : (1 << bits); }

/* at least 2 waves per SIMD: the registers of a KPT = 32 block then match the
 * two blocks per CU its 70 KiB of LDS allows (asking for 4 with KPT = 16
 * spills) */
/* 8-bit digits x 8 keys and 9-bit digits x 7 keys (7168-key tiles, 78 KiB of
 * LDS) with 1024 threads and window 16: two blocks (32 waves) per CU, which
 * needs <= 64 VGPRs (a small spill in the look-back); window 32: one block per
 * CU, no spill.  The autotuner picks between them. */
constexpr int pass_min_waves(int bits, int kpt, int lbw)
{
    return (((bits == 8 && kpt == 8) || (bits == 9 && kpt == 7)) && lbw == 16) ? 8 : 2;
}

template <int BITS, int KPT, int LB_W, bool RAW>
__global__ __launch_bounds__(pass_threads(BITS, KPT), pass_min_waves(BITS, KPT, LB_W)) void k_os_pass(const uint64_t *__restrict__ in,
                                                       uint64_t *__restrict__ out, size_t n,
                                                       SortCode c, int shift,
                                                       const uint64_t *__restrict__ gbase,
                                                       uint64_t *status, unsigned *tile_ctr,
                                                       uint32_t epoch, unsigned *err)
{
    constexpr int R = 1 << BITS, T = pass_threads(BITS, KPT), NW = T / 64, TILE = T * KPT,
                  WCHUNK = 64 * KPT;
    static_assert(TILE < 65536, "per-wave counters are 16-bit");
    static_assert(T >= R, "a thread per digit");
    __shared__ uint16_t wcnt[NW * R];   /* per-wave digit counts, then wave prefixes */
    __shared__ uint32_t tstart[R];
    __shared__ uint64_t obase[R];
    __shared__ uint64_t stage[TILE];
    __shared__ uint32_t wtmp[17];
    __shared__ unsigned tile_sh;

    const int w = threadIdx.x >> 6, lane = lane_id();
    if (threadIdx.x == 0) tile_sh = atomicAdd(tile_ctr, 1u);
    for (int i = threadIdx.x; i < NW * R; i += T) wcnt[i] = 0;
    __syncthreads();
    const size_t tile = tile_sh;
    const size_t base = tile * TILE + (size_t)w * WCHUNK;
    /* keys of this wave: rows j with j * 64 + lane < rem */
    const int rem = base < n ? (int)((n - base) < (size_t)WCHUNK ? (n - base) : (size_t)WCHUNK) : 0;
    uint64_t key[KPT];
    uint32_t rank2[(KPT + 1) / 2];   /* ranks (< 64 KPT) packed two per register */
    const uint64_t lt = lanemask_lt();
#pragma unroll
    for (int j = 0; j < KPT; j++) key[j] = j * 64 + lane < rem ? in[base + j * 64 + lane] : 0;
    if (RAW) {
        uint32_t bad = 0;   /* checked by the histogram kernel */
#pragma unroll
        for (int j = 0; j < KPT; j++) key[j] = sc_encode(key[j], c, bad);
    }
#pragma unroll
    for (int j = 0; j < KPT; j++) {
        bool valid = j * 64 + lane < rem;
        unsigned d = (unsigned)(key[j] >> shift) & (R - 1);
        /* lanes whose digit differs from this lane's in some bit: OR over the
         * bits of (ballot of the bit) XOR (the lane's own bit as 0 / all ones) */
        uint64_t mism = 0;
#pragma unroll
        for (int bit = 0; bit < BITS; bit++) {
            uint32_t t = 0u - ((d >> bit) & 1u);
            uint64_t bb = __ballot(t != 0);
            mism |= bb ^ (((uint64_t)t << 32) | t);
        }
        uint64_t peers = __ballot(valid) & ~mism;
        uint32_t before = (uint32_t)__popcll(peers & lt);
        uint32_t cnt = wcnt[w * R + d];
        __builtin_amdgcn_wave_barrier();
        if (valid && before == 0) wcnt[w * R + d] = (uint16_t)(cnt + (uint32_t)__popcll(peers));
        __builtin_amdgcn_wave_barrier();
        uint32_t r = cnt + before;
        if (j & 1) rank2[j >> 1] |= r << 16;
        else rank2[j >> 1] = r;
    }
    __syncthreads();
    /* thread d < R: wave prefixes (in place) and the tile's count of digit d */
    const int d = threadIdx.x;
    uint32_t tot = 0;
    /* the 8-waves-per-SIMD variant keeps its look-back window small */
    constexpr int LBW = pass_min_waves(BITS, KPT, LB_W) > 2 ? LB_W / 2 : LB_W;
    constexpr int PF = LBW < 16 ? LBW : 16;   /* words loaded ahead of the staging */
    uint64_t pre[PF];
    uint64_t *my = status + tile * R + d;
    if (d < R) {
#pragma unroll
        for (int ww = 0; ww < NW; ww++) {
            uint32_t c = wcnt[ww * R + d];
            wcnt[ww * R + d] = (uint16_t)tot;
            tot += c;
        }
        /* chained scan over tiles for digit d: publish this tile's count,
         * and start loading the predecessors' words; they are used after
         * the staging below, which hides one round trip */
        if (tile == 0) {
            st_flag(my, OS_INC, epoch, tot);
        } else {
            st_flag(my, OS_AGG, epoch, tot);
            lookback_prefetch<PF>(status + d, R, tile, pre);
        }
    }
    uint32_t all;
    uint32_t ts = block_excl_scan<uint32_t>(tot, &all, wtmp);   /* threads >= R add 0 */
    if (d < R) tstart[d] = ts;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < KPT; j++) {
        if (j * 64 + lane < rem) {
            /* recompute the digit rather than keep KPT digits live across the scan */
            uint64_t kk = key[j];
            asm volatile("" : "+v"(kk));
            unsigned dd = (unsigned)(kk >> shift) & (R - 1);
            uint32_t r = (rank2[j >> 1] >> ((j & 1) * 16)) & 0xffffu;
            stage[tstart[dd] + wcnt[w * R + dd] + r] = kk;
        }
    }
    if (d < R) {
        uint64_t excl = 0;
        if (tile != 0) {
            const size_t ntiles = (n + TILE - 1) / TILE;
            excl = lookback_grouped<LBW, PF>(status + d, status + ntiles * R + d, R, tile, tot, epoch,
                                             err, pre);
            st_flag(my, OS_INC, epoch, excl + tot);
        }
        obase[d] = gbase[d] + excl;
    }
    __syncthreads();
    size_t tile0 = tile * TILE;
    uint32_t tn = (uint32_t)((n - tile0) < (size_t)TILE ? (n - tile0) : (size_t)TILE);
    for (uint32_t p = threadIdx.x; p < tn; p += T) {
        uint64_t kk = stage[p];
        unsigned dd = (unsigned)(kk >> shift) & (R - 1);
        out[obase[dd] + (p - tstart[dd])] = kk;
    }
}

/* unique compaction of a sorted array with a chained scan over tiles;
 * the last tile writes the total to *total */
/* PRUNE: the sort drops the distinct keys (boards) that fail kp's prune
 * test (a separate instance keeps the plain compaction's registers) */
template <bool PRUNE>
__global__ __launch_bounds__(UQ_THREADS) void k_os_unique(const uint64_t *__restrict__ k, size_t n,
                                                          uint64_t *__restrict__ dst,
                                                          uint64_t *status, unsigned *tile_ctr,
                                                          uint32_t epoch,
                                                          unsigned long long *total, unsigned *err,
                                                          int packed, SortCode code,
                                                          const uint64_t *cval, KeyPrune kp)
{
    constexpr int WCHUNK = 64 * UQ_KPT;
    __shared__ unsigned tile_sh;
    __shared__ uint32_t wc[UQ_THREADS / 64];
    __shared__ uint64_t excl_sh;
    const int w = threadIdx.x >> 6, lane = lane_id();
    if (threadIdx.x == 0) tile_sh = atomicAdd(tile_ctr, 1u);
    __syncthreads();
    const size_t tile = tile_sh;
    const size_t base = tile * UQ_TILE + (size_t)w * WCHUNK;
    const uint64_t lt = lanemask_lt();
    uint32_t c = 0;
    uint64_t m[UQ_KPT], kv[UQ_KPT];
    /* one load per key; the predecessor comes from the neighbouring lane,
     * lane 0 of each row reads the key before the row */
#pragma unroll
    for (int j = 0; j < UQ_KPT; j++) {
        size_t i = base + (size_t)j * 64 + lane;
        kv[j] = i < n ? k[i] : 0;
    }
    uint64_t prev0[UQ_KPT];
#pragma unroll
    for (int j = 0; j < UQ_KPT; j++) {
        size_t i0 = base + (size_t)j * 64;
        prev0[j] = (lane == 0 && i0 > 0 && i0 <= n) ? k[i0 - 1] : 0;
    }
    /* codes go back to the original bits: sc_decode plus the constant bits */
    const uint64_t cv = packed ? *cval : 0;
#pragma unroll
    for (int j = 0; j < UQ_KPT; j++) {
        size_t i = base + (size_t)j * 64 + lane;
        uint64_t pv = __shfl_up(kv[j], 1, 64);
        if (lane == 0) pv = prev0[j];
        bool f = i < n && (i == 0 || kv[j] != pv);
        if (PRUNE && f) f = !cg_prune(packed ? (sc_decode(kv[j], code) | cv) : kv[j], kp.stsl, kp.ltc, kp.sl);
        m[j] = __ballot(f);
        c += (uint32_t)__popcll(m[j]);
    }
    if (lane == 0) wc[w] = c;
    __syncthreads();
    if (w == 0) {
        /* wave 0 publishes and looks back 64 tiles per step (lane i reads tile t - i) */
        uint32_t tot = 0;
#pragma unroll
        for (int ww = 0; ww < UQ_THREADS / 64; ww++) tot += wc[ww];
        uint64_t *my = status + tile;
        uint64_t excl = 0;
        if (tile == 0) {
            if (lane == 0) st_flag(my, OS_INC, epoch, tot);
        } else {
            if (lane == 0) st_flag(my, OS_AGG, epoch, tot);
            long t = (long)tile - 1;
            while (t >= 0) {
                long ti = t - lane;
                uint64_t v = ti >= 0 ? ld_flag(status + ti) : 0;
                uint32_t f = ti >= 0 ? fstate(v, epoch) : 2u;
                uint64_t notready = __ballot(f == 0), inc = __ballot(f == 2);
                /* nearest lane that is not ready, and nearest inclusive one */
                int nr = notready ? __ffsll((long long)notready) - 1 : 64;
                int ni = inc ? __ffsll((long long)inc) - 1 : 64;
                int stop = nr < ni ? nr : ni;          /* lanes before `stop` are aggregates */
                uint64_t part =
                    (ti >= 0 && (lane < stop || (lane == ni && ni < nr))) ? (v & OS_VAL) : 0;
#pragma unroll
                for (int dd = 32; dd >= 1; dd >>= 1) part += __shfl_xor(part, dd, 64);
                excl += part;
                if (ni < nr) break;                    /* reached an inclusive prefix */
                if (nr < 64) {
                    /* wait for the nearest unpublished tile, then continue past it */
                    uint64_t wv = 0;
                    if (lane == 0) wv = wait_flag(status + (t - nr), epoch, err);
                    wv = __shfl(wv, 0, 64);
                    excl += wv & OS_VAL;
                    if (fstate(wv, epoch) == 2) break;
                    t = t - nr - 1;
                } else {
                    t -= 64;
                }
            }
            if (lane == 0) st_flag(my, OS_INC, epoch, excl + tot);
        }
        if (lane == 0) {
            excl_sh = excl;
            if ((tile + 1) * UQ_TILE >= n) *total = excl + tot;
        }
    }
    __syncthreads();
    uint64_t o = excl_sh;
    for (int ww = 0; ww < w; ww++) o += wc[ww];
#pragma unroll
    for (int j = 0; j < UQ_KPT; j++) {
        if ((m[j] >> lane) & 1ULL)
            dst[o + __popcll(m[j] & lt)] = packed ? (sc_decode(kv[j], code) | cv) : kv[j];
        o += __popcll(m[j]);
    }
}

/* ------------------------------------------------------------------------ */
/* sample sort: one partition pass into LDS-sized buckets, then each bucket  */
/* sorted and deduplicated in LDS                                            */
/* ------------------------------------------------------------------------ */
/* For mid-size sorts (the move candidates and spawn-4 sets of a layer step,
 * 10^4 to 10^7 keys) the LSD passes are bound by tile latency rather than
 * bandwidth, and every duplicate crosses HBM once per pass.  Here:
 *   k_ss_gather, k_ss_runs, k_ss_select  a regular sample of SS_S keys,
 *                sorted in 16 runs of 1024 by 16 blocks; B - 1 splitters
 *                (every (S/B)-th sample) by each value's rank over the runs;
 *   k_ss_part    one read of the keys: each key goes to the bucket of the
 *                splitters around it (all copies of a key to one bucket),
 *                tiles rank their keys per bucket in LDS and reserve room
 *                with one atomic per (tile, bucket) in a bucket-major
 *                buffer (SS_NX regions per bucket, one per XCD);
 *   k_ss_local   one block per bucket (in ticket order): the bucket's keys
 *                are deduplicated through an LDS hash set, the distinct
 *                values sorted in LDS by stable 8-bit LSD passes over the
 *                bits in which they differ, and written at their place,
 *                which a chained scan of distinct counts over the buckets
 *                gives.
 * So a key is read twice and written once, and only unique keys reach the
 * output.  A bucket above SS_LCAP keys (a sample far from its data, or one
 * key repeated tens of thousands of times) sets SORT_ERR_OVERSIZE: the
 * caller then repeats the sort on the LSD path. */
constexpr int SS_T = 1024;                  /* threads of every sample-sort block */
constexpr int SS_NW = SS_T / 64;
constexpr int SS_LCAP = 16384;              /* keys of a bucket (32-bit relative values) */
constexpr int SS_LCAP64 = 8192;             /* keys of a bucket that spans 2^32 or more */
constexpr int SS_KS32 = 16;                 /* rows per thread of a bucket's sort: 16384 values */
constexpr int SS_KS64 = 8;                  /* 64-bit values: 8192 */
constexpr int SS_PK = 8;                    /* keys per thread of a partition tile */
constexpr int SS_TILE = SS_T * SS_PK;       /* 8192 */
constexpr int SS_BMAX = 16384;              /* buckets */
constexpr int SS_BMID = 8192;               /* buckets of the partition with 64-bit splitters in LDS */
constexpr int SS_BSMALL = 2048;             /* buckets of the partition with two blocks per CU */
/* A bucket's keys land in SS_NX regions, one per group of partition blocks
 * (blockIdx % SS_NX, which on MI355X is the XCD the block runs on): SS_NX
 * counters per bucket, so no counter takes every tile's reservation, and a
 * region's partly written lines stay in one XCD's L2.  Each region holds a
 * whole bucket, since the keys of one bucket may all come from a few
 * neighbouring tiles (moves of a sorted layer keep its order). */
constexpr int SS_NX = 8;
constexpr int SS_SUBCAP = SS_LCAP;
constexpr int SS_BSTRIDE = SS_NX * SS_SUBCAP;   /* keys of a bucket's regions */
static_assert((uint64_t)SS_BMAX * SS_BSTRIDE < 0xffffffffULL, "places in part are 32-bit");
/* splitter guide: the number of splitters below smin + (p << gshift) for
 * p = 0..SS_GUIDE, which narrows a key's splitter search to a few entries */
constexpr int SS_GUIDE = 4096;
constexpr int SS_S = SS_LCAP;               /* samples up to SS_BSMALL buckets (8 per bucket) */
constexpr int SS_SMAX = 8 * SS_S;            /* samples of the largest sorts */
/* keys a bucket's regions hold; the default cap on a bucket's keys (only its
 * distinct values must fit the LDS) */
constexpr int SS_KEYCAP = SS_NX * SS_SUBCAP;

/* Buckets of a sample sort of n keys: 1 (one block takes every key) up to
 * SS_LCAP64 keys, which a bucket of keys spanning 2^32 or more holds as
 * distinct 64-bit values, else ceil(n / sh.bavg) with at least 2, and for
 * the largest sorts buckets of sh.bhi keys up to sh.bmax of them.
 * The host and the kernels compute it alike, so a sort whose size is only
 * known on the device sizes its grids for the bound and the kernels agree on
 * B. */
struct SsShape {
    int bavg;   /* mean keys per bucket up to SS_BSMALL buckets */
    int bhi;    /* mean keys per bucket of the larger sorts (>= bavg) */
    int bmax;   /* most buckets */
};

__host__ __device__ __forceinline__ size_t ss_nbuckets(size_t n, SsShape sh)
{
    if (n <= (size_t)SS_LCAP64) return 1;
    size_t b = (n + (size_t)sh.bavg - 1) / (size_t)sh.bavg;
    b = b < 2 ? 2 : b;
    if (b <= (size_t)SS_BSMALL) return b;
    /* SS_BSMALL buckets while their mean stays within bhi, then buckets of
     * bhi keys: B never decreases as n grows */
    if (n <= (size_t)SS_BSMALL * (size_t)sh.bhi) return SS_BSMALL;
    b = (n + (size_t)sh.bhi - 1) / (size_t)sh.bhi;
    return b > (size_t)sh.bmax ? (size_t)sh.bmax : b;
}

/* the sort's key count: *dn (read on the device) bounded by nmax, or nmax */
__device__ __forceinline__ size_t ss_count(const unsigned long long *dn, size_t nmax)
{
    if (!dn) return nmax;
    const size_t v = (size_t)*dn;
    return v < nmax ? v : nmax;
}

/* 64 KiB of values plus counters: two blocks per CU */
struct SsLds {
    union {
        uint32_t v32[SS_LCAP];
        uint64_t v64[SS_LCAP64];
    } stage;
    /* digit-major counters: [digit][wave], 8-byte aligned for the scan's
     * 4-counter loads; counts, then each (digit, wave)'s first position */
    uint16_t cnt[256 * SS_NW] __attribute__((aligned(8)));
    uint32_t wtmp[17];
    uint64_t red[2 * SS_NW];
    uint64_t vlo, vhi;      /* extremes of a bucket's values */
    uint32_t segc[SS_NX];   /* keys of the bucket's regions */
    uint64_t excl;
    unsigned ticket;
};

template <typename V> __device__ __forceinline__ V *ss_stage(SsLds &L);
template <> __device__ __forceinline__ uint32_t *ss_stage<uint32_t>(SsLds &L) { return L.stage.v32; }
template <> __device__ __forceinline__ uint64_t *ss_stage<uint64_t>(SsLds &L) { return L.stage.v64; }

/* minimum and maximum of this thread's keys folded over the block */
__device__ __forceinline__ void ss_min_max(uint64_t a, uint64_t b, SsLds &L, uint64_t *mn, uint64_t *mx)
{
    const int w = threadIdx.x >> 6, lane = lane_id();
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        uint64_t x = __shfl_xor(a, d, 64), y = __shfl_xor(b, d, 64);
        a = x < a ? x : a;
        b = y > b ? y : b;
    }
    if (lane == 0) L.red[w] = a, L.red[SS_NW + w] = b;
    __syncthreads();
    a = ~0ULL;
    b = 0;
#pragma unroll
    for (int ww = 0; ww < SS_NW; ww++) {
        a = L.red[ww] < a ? L.red[ww] : a;
        b = L.red[SS_NW + ww] > b ? L.red[SS_NW + ww] : b;
    }
    *mn = a;
    *mx = b;
    __syncthreads();   /* red may be reused */
}

/* The block's m values are spread over the waves in chunks of `chunk` (a
 * multiple of 64): value i belongs to wave i / chunk, row (i % chunk) / 64,
 * lane i % 64; v[j] of a thread is its row j.  ss_chunk gives the chunk
 * size for m values and K rows per wave at most. */
template <int K>
__device__ __forceinline__ int ss_chunk(int m)
{
    const int rows = (m + SS_NW * 64 - 1) / (SS_NW * 64);
    return (rows < 1 ? 1 : (rows > K ? K : rows)) * 64;
}

__device__ __forceinline__ int ss_rem(int m, int chunk)
{
    const int w = threadIdx.x >> 6;
    return m - w * chunk < 0 ? 0 : (m - w * chunk > chunk ? chunk : m - w * chunk);
}

/* Sorts the m values (layout as above) ascending over their low nb bits,
 * which hold every valid value.  Stable 8-bit LSD passes: each wave ranks its
 * values with one ballot per digit bit against its own counter of each digit
 * (the rows of a wave in order), one block scan over the (digit, wave)
 * counters gives every (digit, wave) its first position, and the values are
 * scattered into the stage.  On return the stage's first m entries hold the
 * sorted values and v[] the values at the thread's positions. */
template <typename V, int K>
__device__ __forceinline__ void ss_block_sort(V (&v)[K], int m, int chunk, int nb, SsLds &L)
{
    V *stage = ss_stage<V>(L);
    const int w = threadIdx.x >> 6, lane = lane_id();
    const uint64_t lt = lanemask_lt();
    const int rem = ss_rem(m, chunk);
    if (nb <= 0) {   /* every valid value is equal: any order is sorted */
#pragma unroll
        for (int j = 0; j < K; j++)
            if (j * 64 + lane < rem) stage[w * chunk + j * 64 + lane] = v[j];
        __syncthreads();
        return;
    }
    uint64_t *cnt64 = reinterpret_cast<uint64_t *>(L.cnt);
    for (int sh = 0; sh < nb; sh += 8) {
        cnt64[threadIdx.x] = 0;   /* 4096 counters: four per thread */
        __syncthreads();
        uint32_t rank2[K / 2];
#pragma unroll
        for (int j = 0; j < K; j++) {
            if (j * 64 >= rem) break;   /* wave-uniform */
            const bool valid = j * 64 + lane < rem;
            const unsigned d = (unsigned)(v[j] >> sh) & 255u;
            uint64_t mism = 0;
#pragma unroll
            for (int bit = 0; bit < 8; bit++) {
                uint32_t t = 0u - ((d >> bit) & 1u);
                uint64_t bb = __ballot(t != 0);
                mism |= bb ^ (((uint64_t)t << 32) | t);
            }
            const uint64_t peers = __ballot(valid) & ~mism;
            const uint32_t before = (uint32_t)__popcll(peers & lt);
            /* the wave's LDS operations execute in order: the leader's
             * update is seen by the next row's read */
            const uint32_t c = L.cnt[d * SS_NW + w];
            if (valid && before == 0) L.cnt[d * SS_NW + w] = (uint16_t)(c + (uint32_t)__popcll(peers));
            const uint32_t r = c + before;
            if (j & 1) rank2[j >> 1] |= r << 16;
            else rank2[j >> 1] = r;
        }
        __syncthreads();
        {
            /* thread t: digit t / 4, waves 4 (t % 4) .. + 3 */
            const uint64_t q = cnt64[threadIdx.x];
            const uint32_t c0 = (uint32_t)(q & 0xffff), c1 = (uint32_t)((q >> 16) & 0xffff),
                           c2 = (uint32_t)((q >> 32) & 0xffff), c3 = (uint32_t)(q >> 48);
            uint32_t all;
            const uint32_t ex = block_excl_scan_dpp(c0 + c1 + c2 + c3, &all, L.wtmp);
            cnt64[threadIdx.x] = (uint64_t)ex | ((uint64_t)(ex + c0) << 16) | ((uint64_t)(ex + c0 + c1) << 32) |
                                 ((uint64_t)(ex + c0 + c1 + c2) << 48);
        }
        __syncthreads();
        /* positions (< 16384) replace the ranks in place; all counter reads
         * are issued before the first store */
#pragma unroll
        for (int j = 0; j < K; j++) {
            if (j * 64 >= rem) break;
            const unsigned d = (unsigned)(v[j] >> sh) & 255u;
            const uint32_t p = L.cnt[d * SS_NW + w] + ((rank2[j >> 1] >> ((j & 1) * 16)) & 0xffffu);
            rank2[j >> 1] = (j & 1) ? ((rank2[j >> 1] & 0xffffu) | (p << 16)) : ((rank2[j >> 1] & 0xffff0000u) | p);
        }
#pragma unroll
        for (int j = 0; j < K; j++) {
            if (j * 64 >= rem) break;
            if (j * 64 + lane < rem) stage[(rank2[j >> 1] >> ((j & 1) * 16)) & 0xffffu] = v[j];
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < K; j++)
            if (j * 64 + lane < rem) v[j] = stage[w * chunk + j * 64 + lane];
    }
    __syncthreads();
}

/* the regular sample of SS_S keys, gathered by many blocks (one block's
 * scattered loads would take ~20 us) into smp, with its smallest and
 * largest key folded into mnx[0] (as ~min) and mnx[1] by atomic max (both
 * cleared to 0 beforehand) */
__global__ __launch_bounds__(SS_T) void k_ss_gather(const uint64_t *__restrict__ keys,
                                                    const unsigned long long *dn, size_t nmax, SsShape sh,
                                                    int S, uint64_t *__restrict__ smp,
                                                    unsigned long long *mnx)
{
    __shared__ uint64_t ra[SS_NW], rb[SS_NW];
    const size_t n = ss_count(dn, nmax);
    if (ss_nbuckets(n, sh) <= 1) return;   /* one block sorts them: no sample */
    const size_t s = (size_t)blockIdx.x * SS_T + threadIdx.x;
    uint64_t a = 0, b = 0;   /* ~min and max */
    if (s < (size_t)S) {
        const uint64_t k = keys[((2 * s + 1) * n) / (2 * (size_t)S)];
        smp[s] = k;
        a = ~k;
        b = k;
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        const uint64_t x = __shfl_xor(a, d, 64), y = __shfl_xor(b, d, 64);
        a = x > a ? x : a;
        b = y > b ? y : b;
    }
    /* one atomic pair per block: same-address atomics from every wave of the
     * grid serialise */
    if (lane_id() == 0) ra[threadIdx.x >> 6] = a, rb[threadIdx.x >> 6] = b;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < SS_NW; w++) {
            a = ra[w] > a ? ra[w] : a;
            b = rb[w] > b ? rb[w] : b;
        }
        atomicMax(&mnx[0], (unsigned long long)a);
        atomicMax(&mnx[1], (unsigned long long)b);
    }
}

/* The sample's order: its values are the top 32 bits of its range, (key -
 * min) >> z, and a splitter is min + (value << z): monotone, which is all a
 * splitter needs.  Ties in those bits only move a bucket boundary to a
 * multiple of 2^z.  Fewer bits would not do: a layer's keys fill a tiny part
 * of their range, and one 2^24-th of it can hold more keys than a bucket. */
__device__ __forceinline__ void ss_sample_range(const unsigned long long *mnx, uint64_t *mn, int *nb, int *z)
{
    const uint64_t lo = ~(uint64_t)mnx[0], hi = mnx[1];
    *mn = lo;
    *nb = hi > lo ? 64 - __clzll((long long)(hi - lo)) : 0;
    *z = *nb > 32 ? *nb - 32 : 0;
}

/* SS_NRUN blocks each sort SS_RUN of the sample's values into a run */
constexpr int SS_RUN = SS_T;
constexpr int SS_NRUN = SS_S / SS_RUN;

__global__ __launch_bounds__(SS_T) void k_ss_runs(const uint64_t *__restrict__ smp, const unsigned long long *mnx,
                                                  uint32_t *__restrict__ runs, const unsigned long long *dn,
                                                  size_t nmax, SsShape sh)
{
    __shared__ SsLds L;
    if (ss_nbuckets(ss_count(dn, nmax), sh) <= 1) return;
    uint64_t mn;
    int nb, z;
    ss_sample_range(mnx, &mn, &nb, &z);
    /* one row per wave: value t of the run is thread t's */
    uint32_t v[2];
    v[0] = (uint32_t)((smp[(size_t)blockIdx.x * SS_RUN + threadIdx.x] - mn) >> z);
    v[1] = 0;
    int m = SS_RUN;
    asm volatile("" : "+s"(m));
    ss_block_sort<uint32_t, 2>(v, m, 64, nb - z, L);
    runs[(size_t)blockIdx.x * SS_RUN + threadIdx.x] = L.stage.v32[threadIdx.x];
}

/* Splitter i is the sample value of rank ((i + 1) SS_S) / B.  Each thread
 * takes one value of one run and finds its rank in the whole sample: its
 * place in its run plus, in every other run, the values before it (equal
 * values of earlier runs count as before, of later runs as after, so the
 * ranks are 0..SS_S - 1 once each); the thread holding a splitter's rank
 * writes it. */
__global__ __launch_bounds__(SS_T) void k_ss_select(const uint32_t *__restrict__ runs,
                                                    const unsigned long long *mnx,
                                                    const unsigned long long *dn, size_t nmax, SsShape sh,
                                                    uint64_t *__restrict__ spl)
{
    __shared__ uint32_t R[SS_S];
    const int B = (int)ss_nbuckets(ss_count(dn, nmax), sh);
    if (B <= 1) return;
    for (int i = threadIdx.x; i < SS_S; i += SS_T) R[i] = runs[i];
    __syncthreads();
    const int r = blockIdx.x;
    const uint32_t e = R[r * SS_RUN + threadIdx.x];
    uint32_t rank = threadIdx.x;
    for (int q = 0; q < SS_NRUN; q++) {
        if (q == r) continue;
        const uint32_t *X = R + q * SS_RUN;
        uint32_t lo = 0;
#pragma unroll
        for (uint32_t step = SS_RUN / 2; step; step >>= 1) {
            const uint32_t x = X[lo + step - 1];
            if (q < r ? x <= e : x < e) lo += step;
        }
        const uint32_t x = X[lo < SS_RUN ? lo : SS_RUN - 1];
        if (lo < SS_RUN && (q < r ? x <= e : x < e)) lo++;
        rank += lo;
    }
    const uint64_t c = ((uint64_t)rank * (uint64_t)B + SS_S - 1) / SS_S;   /* i + 1 */
    if (c >= 1 && c + 1 <= (uint64_t)B && (c * SS_S) / (uint64_t)B == rank) {
        uint64_t mn;
        int nb, z;
        ss_sample_range(mnx, &mn, &nb, &z);
        spl[c - 1] = mn + ((uint64_t)e << z);
    }
}

/* The splitters of a larger sample (S = 32768 or 65536 values in S / 1024
 * sorted runs), whose values do not fit one block's LDS: block (r, g) loads
 * the SS_RGRP runs of group g and adds to each value of run r its count of
 * values before it in those runs (the tie rule of k_ss_select); k_ss_pick
 * then writes the values whose summed rank is a splitter's. */
constexpr int SS_RGRP = SS_S / SS_RUN;   /* runs per group: 64 KiB of LDS */

__global__ __launch_bounds__(SS_T) void k_ss_rank(const uint32_t *__restrict__ runs, int nrun,
                                                  const unsigned long long *dn, size_t nmax, SsShape sh,
                                                  uint32_t *__restrict__ rank)
{
    __shared__ uint32_t R[SS_RGRP * SS_RUN];
    if (ss_nbuckets(ss_count(dn, nmax), sh) <= 1) return;
    const int r = (int)(blockIdx.x % (unsigned)nrun), g = (int)(blockIdx.x / (unsigned)nrun);
    const int q0 = g * SS_RGRP;
    for (int i = threadIdx.x; i < SS_RGRP * SS_RUN; i += SS_T) R[i] = runs[(size_t)q0 * SS_RUN + i];
    __syncthreads();
    const uint32_t e = runs[(size_t)r * SS_RUN + threadIdx.x];
    uint32_t c = 0;
    for (int qq = 0; qq < SS_RGRP; qq++) {
        const int q = q0 + qq;
        if (q == r) {
            c += threadIdx.x;
            continue;
        }
        const uint32_t *X = R + qq * SS_RUN;
        uint32_t lo = 0;
#pragma unroll
        for (uint32_t step = SS_RUN / 2; step; step >>= 1) {
            const uint32_t x = X[lo + step - 1];
            if (q < r ? x <= e : x < e) lo += step;
        }
        const uint32_t x = X[lo < SS_RUN ? lo : SS_RUN - 1];
        if (lo < SS_RUN && (q < r ? x <= e : x < e)) lo++;
        c += lo;
    }
    atomicAdd(&rank[(size_t)r * SS_RUN + threadIdx.x], c);
}

__global__ __launch_bounds__(SS_T) void k_ss_pick(const uint32_t *__restrict__ runs,
                                                  const uint32_t *__restrict__ rank,
                                                  const unsigned long long *mnx,
                                                  const unsigned long long *dn, size_t nmax, SsShape sh, int S,
                                                  uint64_t *__restrict__ spl)
{
    const int B = (int)ss_nbuckets(ss_count(dn, nmax), sh);
    if (B <= 1) return;
    const size_t i = (size_t)blockIdx.x * SS_T + threadIdx.x;
    const uint32_t e = runs[i], rk = rank[i];
    const uint64_t c = ((uint64_t)rk * (uint64_t)B + S - 1) / (uint64_t)S;   /* splitter c - 1 */
    if (c >= 1 && c + 1 <= (uint64_t)B && (c * (uint64_t)S) / (uint64_t)B == rk) {
        uint64_t mn;
        int nb, z;
        ss_sample_range(mnx, &mn, &nb, &z);
        spl[c - 1] = mn + ((uint64_t)e << z);
    }
}

/* Splitters as the partition keeps them in LDS: 64-bit keys up to SS_BMID
 * buckets; above that their sample values e = (splitter - min) >> z, 32 bits
 * and exact (every splitter is min + (e << z)), so 16384 splitters, their
 * counters and the guide fit 136 KiB. */
template <int BM> struct SsSplT { using type = uint64_t; };
template <> struct SsSplT<SS_BMAX> { using type = uint32_t; };

template <int BM>
struct SsPartLds {
    typename SsSplT<BM>::type spl[BM];
    uint32_t cnt[BM];        /* counts, then each bucket's place in part (~0u: region full) */
    uint16_t guide[SS_GUIDE + 1];
};

/* Bucket of every key (the number of splitters <= key, searched between the
 * guide's bounds) and its rank among the tile's keys of that bucket (an LDS
 * atomic); one reservation per (tile, bucket) on the bucket's counter of
 * this block's region group; every key written from registers to part[b *
 * SS_BSTRIDE + group * SS_SUBCAP + reservation + rank].  A block loads the
 * splitters and builds the guide once, then strides over the tiles (a grid
 * of fewer blocks than tiles); its group stays blockIdx % SS_NX, so with a
 * grid that is a multiple of SS_NX tile t writes group t % SS_NX; with PF
 * the next tile's keys are loaded while this tile reserves.  Without
 * a staging tile the LDS holds only the splitters, guide and counters, so
 * two blocks share a CU up to SS_BSMALL buckets (one block above it: 104
 * KiB at SS_BMID, 136 KiB at SS_BMAX with 32-bit splitters); a tile's run
 * of one bucket is a few keys either way.  A tile
 * whose reservation would pass the region's end writes none of that
 * bucket's keys (the bucket is then oversize). */
template <int BM, bool PF = false>
__global__ __launch_bounds__(SS_T, BM <= SS_BSMALL ? 2 : 1) void k_ss_part(const uint64_t *__restrict__ keys,
                                                     const unsigned long long *dn, size_t nmax, SsShape sh,
                                                     const uint64_t *__restrict__ spl_g,
                                                     const unsigned long long *mnx,
                                                     uint64_t *__restrict__ part,
                                                     unsigned long long *gcur, unsigned *err)
{
    __shared__ SsPartLds<BM> P;
    const size_t n = ss_count(dn, nmax);
    const int B = (int)ss_nbuckets(n, sh);
    const size_t ntiles = (n + SS_TILE - 1) / SS_TILE;
    /* the grid covers nmax keys or fewer blocks that stride over the tiles:
     * blocks past n's tiles (and every block of a one-block sort) have
     * nothing to do */
    if (B <= 1 || (size_t)blockIdx.x >= ntiles) return;
    const int nspl = B - 1;
    const int grp = blockIdx.x % SS_NX;
    uint64_t smin;
    int nbs, zs;
    ss_sample_range(mnx, &smin, &nbs, &zs);
    constexpr bool S32 = sizeof(P.spl[0]) == 4;
    for (int i = threadIdx.x; i < nspl; i += SS_T) {
        if constexpr (S32) P.spl[i] = (uint32_t)((spl_g[i] - smin) >> zs);
        else P.spl[i] = spl_g[i];
    }
    /* guide[p]: splitters below smin + (p << gs); the sample's range fits
     * SS_GUIDE steps of 2^gs */
    const int gs = nbs > 12 ? nbs - 12 : 0;
    __syncthreads();
    for (int p = threadIdx.x; p <= SS_GUIDE; p += SS_T) {
        int lo = 0, hi = nspl;
        if (p == SS_GUIDE) {
            lo = nspl;
        } else {
            const uint64_t bound = smin + ((uint64_t)p << gs);
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                uint64_t sv;
                if constexpr (S32) sv = smin + ((uint64_t)P.spl[mid] << zs);
                else sv = P.spl[mid];
                if (sv < bound) lo = mid + 1;
                else hi = mid;
            }
        }
        P.guide[p] = (uint16_t)lo;
    }
    uint64_t k[SS_PK];
    if constexpr (PF) {
#pragma unroll
        for (int j = 0; j < SS_PK; j++) {
            const size_t i = (size_t)blockIdx.x * SS_TILE + (size_t)j * SS_T + threadIdx.x;
            k[j] = i < n ? keys[i] : 0;
        }
    }
    for (size_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        for (int i = threadIdx.x; i < B; i += SS_T) P.cnt[i] = 0;
        __syncthreads();   /* the guide (first tile), the cleared counters */
        const size_t tile0 = tile * SS_TILE;
        uint32_t br[SS_PK];   /* bucket << 16 | rank in the tile's bucket */
        if constexpr (!PF) {
#pragma unroll
            for (int j = 0; j < SS_PK; j++) {
                const size_t i = tile0 + (size_t)j * SS_T + threadIdx.x;
                k[j] = i < n ? keys[i] : 0;
            }
        }
#pragma unroll
        for (int j = 0; j < SS_PK; j++) {
            const size_t i = tile0 + (size_t)j * SS_T + threadIdx.x;
            if (i < n) {
                uint64_t pp = k[j] < smin ? 0 : (k[j] - smin) >> gs;
                const int p = pp < (uint64_t)SS_GUIDE ? (int)pp : SS_GUIDE - 1;
                int lo = P.guide[p], hi = P.guide[p + 1];
                if constexpr (S32) {
                    /* for a key >= smin, min + (e << z) <= key exactly when e <=
                     * (key - min) >> z (capped at 2^32 - 1, which every e is
                     * within); a key below smin is below every splitter */
                    const uint64_t q = (k[j] - smin) >> zs;
                    const uint32_t kv = q > 0xffffffffULL ? 0xffffffffu : (uint32_t)q;
                    if (k[j] < smin) hi = lo;
                    while (lo < hi) {
                        const int mid = (lo + hi) >> 1;
                        if (P.spl[mid] <= kv) lo = mid + 1;
                        else hi = mid;
                    }
                } else {
                    while (lo < hi) {
                        const int mid = (lo + hi) >> 1;
                        if (P.spl[mid] <= k[j]) lo = mid + 1;
                        else hi = mid;
                    }
                }
                br[j] = ((uint32_t)lo << 16) | atomicAdd(&P.cnt[lo], 1u);
            }
        }
        /* the next tile's keys load while this one reserves and writes */
        uint64_t kn[PF ? SS_PK : 1];
        if constexpr (PF) {
            const size_t t1 = (tile + gridDim.x) * SS_TILE;
#pragma unroll
            for (int j = 0; j < SS_PK; j++) {
                const size_t i = t1 + (size_t)j * SS_T + threadIdx.x;
                kn[j] = i < n ? keys[i] : 0;
            }
        }
        __syncthreads();
        /* BM / SS_T buckets per thread: the reservations */
        constexpr int PER = BM / SS_T;
        static_assert(BM % SS_T == 0 && BM <= 65536, "buckets per thread; bucket << 16 in br");
        bool over = false;
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const int b = threadIdx.x + q * SS_T;
            const uint32_t c = b < B ? P.cnt[b] : 0;
            if (c) {
                const unsigned long long g = atomicAdd(&gcur[b * SS_NX + grp], (unsigned long long)c);
                const bool o = g + c > (unsigned long long)SS_SUBCAP;
                over |= o;
                P.cnt[b] = o ? ~0u : (uint32_t)((uint64_t)b * SS_BSTRIDE + (uint64_t)grp * SS_SUBCAP + g);
            }
        }
        if (over) atomicOr(err, SORT_ERR_OVERSIZE);
        __syncthreads();
#pragma unroll
        for (int j = 0; j < SS_PK; j++) {
            const size_t i = tile0 + (size_t)j * SS_T + threadIdx.x;
            if (i < n) {
                const uint32_t base = P.cnt[br[j] >> 16];
                if (base != ~0u) part[base + (br[j] & 0xffffu)] = k[j];
            }
        }
        __syncthreads();   /* every place read before the next tile's counters */
        if constexpr (PF) {
#pragma unroll
            for (int j = 0; j < SS_PK; j++) k[j] = kn[j];
        }
    }
}

/* Longest probe sequence of a hash-set insert.  Linear probing stays far
 * below it up to a load of about 3/4; a set filled closer than that would
 * make each insert walk long runs of slots (quadratic in the keys). */
constexpr int SS_PROBES = 64;
constexpr int SS_PROBES_FAR = 512;

/* Drops the n distinct values at the front of the stage whose board (add +
 * value) fails the prune test; returns the count kept.  A pruned board is
 * rare, so the values are tested first and compacted only when one is
 * found: row by row, kept values moving left into rows already read. */
template <typename V>
__device__ __forceinline__ int ss_prune_distinct(int n, uint64_t add, const KeyPrune &kp, SsLds &L)
{
    V *T = ss_stage<V>(L);
    bool any = false;
    for (int i = threadIdx.x; i < n; i += SS_T) any |= cg_prune(add + (uint64_t)T[i], kp.stsl, kp.ltc, kp.sl) != 0;
    if (!__syncthreads_or(any)) return n;
    uint32_t kept = 0;
    for (int r0 = 0; r0 < n; r0 += SS_T) {
        const int i = r0 + (int)threadIdx.x;
        V x = 0;
        uint32_t k = 0;
        if (i < n) {
            x = T[i];
            k = cg_prune(add + (uint64_t)x, kp.stsl, kp.ltc, kp.sl) ? 0u : 1u;
        }
        uint32_t tot;
        const uint32_t o = block_excl_scan_dpp(k, &tot, L.wtmp);   /* its barrier ends the row's reads */
        if (k) T[kept + o] = x;
        kept += tot;
        __syncthreads();   /* the row's writes, and wtmp free again */
    }
    return (int)kept;
}

/* Deduplicates the bucket's keys (the SS_NX regions of L.segc[x] keys at s +
 * x * SS_SUBCAP) as values k - base of type V through an LDS hash set: the
 * stage as open-addressing slots with linear probing, all ones marking an
 * empty slot (never a value).  Keys are inserted as they are read, so only
 * distinct values are sorted: a move-candidate bucket holds each key about
 * 3.6 times.  The distinct values are compacted to the front of the stage
 * relative to the smallest of them (*vmin), and *nbits is the width of
 * their range.  Returns their count, or -1 when the set or one block's sort
 * (SS_T * K values) cannot hold them. */
template <typename V, int K>
__device__ __forceinline__ int ss_dedupe(const uint64_t *__restrict__ s, uint64_t base, SsLds &L, uint64_t *vmin,
                                         int *nbits, int probes, const KeyPrune &kp)
{
    constexpr int NS = (int)(sizeof(SsLds::stage) / sizeof(V));
    constexpr V EMPTY = (V)~(V)0;
    V *T = ss_stage<V>(L);
    for (int i = threadIdx.x; i < NS; i += SS_T) T[i] = EMPTY;
    if (threadIdx.x == 0) {
        L.vlo = ~0ULL;
        L.vhi = 0;
    }
    __syncthreads();
    bool full = false;
    V lo = EMPTY, hi = 0;
    for (int r = 0; r < SS_NX; r++)
        for (int i = threadIdx.x; i < (int)L.segc[r]; i += SS_T) {
            const V x = (V)(s[(size_t)r * SS_SUBCAP + i] - base);
            lo = x < lo ? x : lo;
            hi = x > hi ? x : hi;
            uint32_t h = (sizeof(V) == 4) ? (((uint32_t)x * 2654435761u) >> 18)
                                          : (uint32_t)(((uint64_t)x * 0x9E3779B97F4A7C15ULL) >> 51);   /* 13 bits */
            for (int probe = 0;; probe++) {
                if (probe == probes) {
                    /* the set is nearly full (mostly distinct keys), or
                     * every slot holds another value */
                    full = true;
                    break;
                }
                const V old = atomicCAS(&T[h], EMPTY, x);
                if (old == EMPTY || old == x) break;
                h = (h + 1) & (NS - 1);
            }
        }
    uint64_t l64 = lo, h64 = hi;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        const uint64_t a = __shfl_xor(l64, d, 64), b = __shfl_xor(h64, d, 64);
        l64 = a < l64 ? a : l64;
        h64 = b > h64 ? b : h64;
    }
    if (lane_id() == 0) {
        atomicMin((unsigned long long *)&L.vlo, (unsigned long long)l64);
        atomicMax((unsigned long long *)&L.vhi, (unsigned long long)h64);
    }
    if (__syncthreads_or(full)) return -1;
    const V vl = (V)L.vlo, vh = (V)L.vhi;
    constexpr int PER = NS / SS_T;
    V x[PER];
    uint32_t c = 0;
#pragma unroll
    for (int j = 0; j < PER; j++) {
        x[j] = T[threadIdx.x + j * SS_T];
        c += x[j] != EMPTY;
    }
    uint32_t all;
    uint32_t o = block_excl_scan_dpp(c, &all, L.wtmp);   /* its barrier ends the reads */
#pragma unroll
    for (int j = 0; j < PER; j++)
        if (x[j] != EMPTY) T[o++] = x[j] - vl;
    __syncthreads();
    int n = (int)all;
    if (n > SS_T * K) return -1;
    /* the range of every distinct value bounds the kept ones */
    *vmin = n ? (uint64_t)vl : 0;
    *nbits = (n > 1 && vh > vl) ? 64 - __clzll((long long)(uint64_t)(vh - vl)) : 0;
    if (kp.on) n = ss_prune_distinct<V>(n, base + (uint64_t)vl, kp, L);
    return n;
}

/* the value a bitonic stage keeps of a and its partner o: the smaller when
 * keep_lo */
template <typename V>
__device__ __forceinline__ V ss_pick(V a, V o, bool keep_lo)
{
    return keep_lo ? (a < o ? a : o) : (a < o ? o : a);
}

/* the stages of distance D < R, within each thread: pairs (r, r + D) */
template <typename V, int R, int D>
__device__ __forceinline__ void ss_intra(V (&v)[R], uint32_t first, uint32_t size)
{
#pragma unroll
    for (int r = 0; r < R; r++)
        if (!(r & D)) {
            const bool up = ((first + (uint32_t)r) & size) == 0;
            const V a = v[r], b = v[r + D];
            const V lo = a < b ? a : b, hi = a < b ? b : a;
            v[r] = up ? lo : hi;
            v[r + D] = up ? hi : lo;
        }
}

/* Bitonic sort of SS_T * R values held R per thread: value i = thread * R +
 * r is v[r].  Pairs less than R apart are exchanged within a thread, less
 * than 64 R apart between the lanes of a wave (shuffles), and the rest
 * through the stage in LDS (r-major, so every access is conflict free), one
 * barrier pair each: 10 of the network's stages whatever R is.  For
 * distances of R or more the pair side and the direction are the thread's.
 * The stage loops run at run time (unrolled, every stage's constants would
 * stay live and spill); only the exchanges within a thread are unrolled,
 * one instance per distance. */
template <typename V, int R>
__device__ __forceinline__ void ss_bitonic(V (&v)[R], SsLds &L)
{
    constexpr int LR = R == 1 ? 0 : R == 2 ? 1 : R == 4 ? 2 : R == 8 ? 3 : 4;
    static_assert((1 << LR) == R && SS_T == 1024, "R a power of two up to 16, 1024 threads");
    constexpr int LOGN = 10 + LR;
    V *T = ss_stage<V>(L);
    const uint32_t tid = threadIdx.x, first = tid * R;
    for (int ls = 1; ls <= LOGN; ls++) {
        const uint32_t size = 1u << ls;
        for (int ld = ls - 1; ld >= 0; ld--) {
            const uint32_t d = 1u << ld;
            if (ld >= LR + 6) {
                const uint32_t tx = d >> LR;
                const bool keep_lo = ((first & d) == 0) == ((first & size) == 0);
                __syncthreads();   /* the stage's last readers are done */
#pragma unroll
                for (int r = 0; r < R; r++) T[r * SS_T + tid] = v[r];
                __syncthreads();
#pragma unroll
                for (int r = 0; r < R; r++) v[r] = ss_pick(v[r], T[r * SS_T + (tid ^ tx)], keep_lo);
            } else if (ld >= LR) {
                const int lx = 1 << (ld - LR);
                const bool keep_lo = ((first & d) == 0) == ((first & size) == 0);
#pragma unroll
                for (int r = 0; r < R; r++) v[r] = ss_pick(v[r], (V)__shfl_xor(v[r], lx, 64), keep_lo);
            } else {
                switch (ld) {
                case 0: ss_intra<V, R, 1>(v, first, size); break;
                case 1: if constexpr (R > 2) ss_intra<V, R, 2>(v, first, size); break;
                case 2: if constexpr (R > 4) ss_intra<V, R, 4>(v, first, size); break;
                default: if constexpr (R > 8) ss_intra<V, R, 8>(v, first, size); break;
                }
            }
        }
    }
    __syncthreads();
}

/* The bucket of m <= SS_T * R keys (its SS_NX regions, L.segc[r] keys
 * each), as values k - base of type V, sorted by the bitonic network (the
 * positions past m hold the all-ones value, which no key's value reaches);
 * the first of each run of equal values is kept.  On return the stage's
 * first n entries hold the distinct values ascending (relative to base);
 * returns n. */
template <typename V, int R>
__device__ __forceinline__ int ss_bitonic_bucket(const uint64_t *__restrict__ s, uint64_t base, int m, SsLds &L,
                                                 const KeyPrune &kp)
{
    const uint32_t tid = threadIdx.x;
    uint32_t pre[SS_NX + 1];
    pre[0] = 0;
#pragma unroll
    for (int r = 0; r < SS_NX; r++) pre[r + 1] = pre[r] + L.segc[r];
    V v[R];
#pragma unroll
    for (int j = 0; j < R; j++) {
        const uint32_t i = (uint32_t)j * SS_T + tid;   /* loads in key order: coalesced */
        v[j] = (V)~(V)0;
        if (i < (uint32_t)m) {
            int r = 0;
#pragma unroll
            for (int q = 1; q < SS_NX; q++) r += i >= pre[q];
            v[j] = (V)(s[(size_t)r * SS_SUBCAP + (i - pre[r])] - base);
        }
    }
    ss_bitonic<V, R>(v, L);
    /* thread t holds positions tR .. tR + R - 1; the last value of the
     * thread before comes through LDS */
    uint64_t *last = reinterpret_cast<uint64_t *>(L.cnt);
    last[tid] = (uint64_t)v[R - 1];
    __syncthreads();
    const V before0 = tid ? (V)last[tid - 1] : (V)0;
    uint32_t keep = 0, c = 0;
#pragma unroll
    for (int r = 0; r < R; r++) {
        const uint32_t p = tid * R + (uint32_t)r;
        const V prev = r ? v[r - 1] : before0;
        if (p < (uint32_t)m && (p == 0 || v[r] != prev) &&
            !(kp.on && cg_prune(base + (uint64_t)v[r], kp.stsl, kp.ltc, kp.sl))) {
            keep |= 1u << r;
            c++;
        }
    }
    uint32_t all;
    uint32_t o = block_excl_scan_dpp(c, &all, L.wtmp);
    V *T = ss_stage<V>(L);
#pragma unroll
    for (int r = 0; r < R; r++)
        if ((keep >> r) & 1u) T[o++] = v[r];
    __syncthreads();
    return (int)all;
}

/* the bitonic bucket sort with the fewest rows per thread that hold m keys */
template <typename V>
__device__ __forceinline__ int ss_bitonic_any(const uint64_t *__restrict__ s, uint64_t base, int m, SsLds &L,
                                              const KeyPrune &kp)
{
    if (m <= SS_T) return ss_bitonic_bucket<V, 1>(s, base, m, L, kp);
    if (m <= 2 * SS_T) return ss_bitonic_bucket<V, 2>(s, base, m, L, kp);
    if (m <= 4 * SS_T) return ss_bitonic_bucket<V, 4>(s, base, m, L, kp);
    if (sizeof(V) == 8 || m <= 8 * SS_T) return ss_bitonic_bucket<V, 8>(s, base, m, L, kp);
    return ss_bitonic_bucket<V, sizeof(V) == 8 ? 8 : 16>(s, base, m, L, kp);
}

/* Sorts the n <= SS_T * R distinct values at the front of the stage
 * (ss_dedupe's output) by the bitonic network, the positions past n holding
 * the all-ones value (no distinct value reaches it: they lie below the hash
 * set's empty mark, or span less than 2^64 - 1); on return they are in
 * ascending order at the stage's first n entries.  Its cost depends on n
 * only, the LSD passes' on the bits the values span. */
template <typename V, int R>
__device__ __forceinline__ void ss_bitonic_distinct(int n, SsLds &L)
{
    V *T = ss_stage<V>(L);
    const uint32_t tid = threadIdx.x;
    V v[R];
    /* any placement sorts: r-major loads are free of bank conflicts */
#pragma unroll
    for (int r = 0; r < R; r++) {
        const uint32_t i = (uint32_t)r * SS_T + tid;
        v[r] = i < (uint32_t)n ? T[i] : (V)~(V)0;
    }
    /* the network's first stage exchange through LDS starts with a
     * barrier, and its last ends with one */
    ss_bitonic<V, R>(v, L);
#pragma unroll
    for (int r = 0; r < R; r++) {
        const uint32_t p = tid * R + (uint32_t)r;
        if (p < (uint32_t)n) T[p] = v[r];
    }
    __syncthreads();
}

template <typename V>
__device__ __forceinline__ void ss_bitonic_distinct_any(int n, SsLds &L)
{
    if (n <= SS_T) ss_bitonic_distinct<V, 1>(n, L);
    else if (n <= 2 * SS_T) ss_bitonic_distinct<V, 2>(n, L);
    else if (n <= 4 * SS_T) ss_bitonic_distinct<V, 4>(n, L);
    else if (sizeof(V) == 8 || n <= 8 * SS_T) ss_bitonic_distinct<V, 8>(n, L);
    else ss_bitonic_distinct<V, sizeof(V) == 8 ? 8 : 16>(n, L);
}

/* Sorts the n distinct values at the front of the stage (ss_dedupe's
 * output) over their low nb bits; on return they are in ascending order at
 * the stage's first n entries. */
template <typename V, int K>
__device__ __forceinline__ void ss_sort_distinct(int n, int nb, SsLds &L)
{
    const int w = threadIdx.x >> 6, lane = lane_id();
    const int chunk = ss_chunk<K>(n), rem = ss_rem(n, chunk);
    const V *T = ss_stage<V>(L);
    V v[K];
#pragma unroll
    for (int j = 0; j < K; j++) v[j] = j * 64 + lane < rem ? T[w * chunk + j * 64 + lane] : (V)0;
    __syncthreads();
    ss_block_sort<V, K>(v, n, chunk, nb, L);
}

/* One bucket per block, buckets taken in ticket order: bucket t holds the
 * counts[t * SS_NX + x] keys of region x at src + t * SS_BSTRIDE + x *
 * SS_SUBCAP (B == 1: the n1 keys at src).  An inner bucket's keys lie in
 * [spl[t - 1], spl[t]): when that range is below 2^32 - 1 the keys are taken
 * relative to spl[t - 1] as 32-bit values without a pass for their extremes;
 * else the smallest and largest key are found first and the values are
 * 32-bit when the keys span less than 2^32 - 1, 64-bit otherwise.  The keys
 * are deduplicated in LDS and the distinct values sorted; the bucket's
 * distinct count is published as soon as it is known, and the chained scan
 * of those counts over the buckets places the unique keys in dst (the last
 * bucket writes the total).  More than cap keys, or more distinct values
 * than the hash set holds, is an oversize bucket (published as empty). */
__global__ __launch_bounds__(SS_T, 8) void k_ss_local(const uint64_t *__restrict__ part,
                                                      const uint64_t *__restrict__ keys,
                                                      const unsigned long long *counts,
                                                      const uint64_t *__restrict__ spl,
                                                      const unsigned long long *dn, size_t nmax, SsShape sh,
                                                      uint64_t *__restrict__ dst, uint64_t *status,
                                                      unsigned *ticket, uint32_t epoch,
                                                      unsigned long long *total, unsigned *err, int cap,
                                                      int probes_direct, int bitonic, int dmax, int dnb,
                                                      KeyPrune kp)
{
    __shared__ SsLds L;
    const int w = threadIdx.x >> 6, lane = lane_id();
    const size_t n1 = ss_count(dn, nmax);
    const int B = (int)ss_nbuckets(n1, sh);
    /* the grid has a block for each bucket of nmax keys: tickets past B
     * (and every block but the first of a one-block sort) leave */
    if (threadIdx.x == 0) L.ticket = B > 1 ? atomicAdd(ticket, 1u) : blockIdx.x;
    __syncthreads();
    const size_t t = L.ticket;
    if (t >= (size_t)B) return;
    const uint64_t *s = B > 1 ? part + t * SS_BSTRIDE : keys;
    size_t m = 0;
    if (B > 1) {
        for (int x = 0; x < SS_NX; x++) m += counts[t * SS_NX + x];
        if (threadIdx.x < SS_NX) {
            const unsigned long long c = counts[t * SS_NX + threadIdx.x];
            L.segc[threadIdx.x] = (uint32_t)(c < (unsigned long long)SS_SUBCAP ? c : SS_SUBCAP);
        }
    } else {
        m = n1;
        if (threadIdx.x < SS_NX) L.segc[threadIdx.x] = threadIdx.x ? 0u : (uint32_t)n1;
    }
    uint64_t base = 0;
    bool wide = false, bounded = false;
    if (B > 1 && t > 0 && t + 1 < (size_t)B) {
        const uint64_t lo = spl[t - 1], up = spl[t];
        bounded = up > lo && up - lo - 1 < 0xffffffffULL;
        base = lo;
    }
    __syncthreads();
    bool over = m > (size_t)cap;
    if (!bounded) {
        uint64_t a = ~0ULL, b = 0;
        if (!over)
            for (int r = 0; r < SS_NX; r++)
                for (int i = threadIdx.x; i < (int)L.segc[r]; i += SS_T) {
                    const uint64_t k = s[(size_t)r * SS_SUBCAP + i];
                    a = k < a ? k : a;
                    b = k > b ? k : b;
                }
        uint64_t mx;
        ss_min_max(a, b, L, &base, &mx);
        /* 32-bit values stay below 0xFFFFFFFF (the hash set's empty mark);
         * a range of 2^64 - 1 would make the largest key's value the empty
         * mark of the 64-bit set */
        wide = m > 1 && mx > base && (mx - base) >= 0xffffffffULL;
        over |= m > 1 && mx > base && (mx - base) == ~0ULL;
    }
    if (over) {   /* oversize: published as empty */
        if (threadIdx.x == 0) atomicOr(err, SORT_ERR_OVERSIZE);
        __syncthreads();   /* every thread has read segc */
        if (threadIdx.x < SS_NX) L.segc[threadIdx.x] = 0;
        __syncthreads();
    }
    uint64_t vmin = 0;
    int nb = 0;
    /* The keys go through the LDS hash set and the distinct values are
     * sorted.  A bucket whose keys fit one block's registers (16384 32-bit
     * or 8192 64-bit values) gives up on a nearly full set (a longer probe
     * costs more than sorting every key) and is then sorted whole by the
     * bitonic network, equal runs dropped; a larger one probes further,
     * since its only other way is a repeat of the whole sort, but not
     * without bound (at a load near 1 every insert would walk the set).
     * With `bitonic` set, buckets that fit are sorted by the network
     * straight away. */
    const bool direct_ok = !over && m <= (size_t)SS_T * (wide ? SS_KS64 : SS_KS32);
    int n = -1;
    bool sorted = false;
    if (bitonic && direct_ok && m > 0) {
        n = wide ? ss_bitonic_any<uint64_t>(s, base, (int)m, L, kp) : ss_bitonic_any<uint32_t>(s, base, (int)m, L, kp);
        sorted = true;
    } else {
        const int probes = direct_ok ? probes_direct : SS_PROBES_FAR;
        n = wide ? ss_dedupe<uint64_t, SS_KS64>(s, base, L, &vmin, &nb, probes, kp)
                 : ss_dedupe<uint32_t, SS_KS32>(s, base, L, &vmin, &nb, probes, kp);
    }
    if (n < 0 && direct_ok) {
        /* the hash set filled up (mostly distinct keys): every key sorted
         * by the bitonic network, equal runs dropped */
        n = wide ? ss_bitonic_any<uint64_t>(s, base, (int)m, L, kp) : ss_bitonic_any<uint32_t>(s, base, (int)m, L, kp);
        vmin = 0;
        sorted = true;
    }
    if (n < 0) {   /* more distinct values than one block sorts: oversize */
        if (threadIdx.x == 0) atomicOr(err, SORT_ERR_OVERSIZE);
        n = 0;
    }
    uint64_t *my = status + t;
    if (B > 1 && threadIdx.x == 0) st_flag(my, t == 0 ? OS_INC : OS_AGG, epoch, (uint64_t)n);
    if (!sorted) {
        /* the distinct values: by the bitonic network when there are at
         * most dmax of them spanning at least dnb bits, else by 8-bit LSD
         * passes over their nb bits (n and nb are the block's) */
        if (n <= dmax && nb >= dnb) {
            if (wide) ss_bitonic_distinct_any<uint64_t>(n, L);
            else ss_bitonic_distinct_any<uint32_t>(n, L);
        } else if (wide) {
            ss_sort_distinct<uint64_t, SS_KS64>(n, nb, L);
        } else {
            ss_sort_distinct<uint32_t, SS_KS32>(n, nb, L);
        }
    }
    if (w == 0) {
        uint64_t excl = 0;
        if (B > 1 && t > 0) {
            long tt = (long)t - 1;
            while (tt >= 0) {
                const long ti = tt - lane;
                const uint64_t v = ti >= 0 ? ld_flag(status + ti) : 0;
                const uint32_t f = ti >= 0 ? fstate(v, epoch) : 2u;
                const uint64_t notready = __ballot(f == 0), inc = __ballot(f == 2);
                const int nr = notready ? __ffsll((long long)notready) - 1 : 64;
                const int ni = inc ? __ffsll((long long)inc) - 1 : 64;
                const int stop = nr < ni ? nr : ni;
                uint64_t part = (ti >= 0 && (lane < stop || (lane == ni && ni < nr))) ? (v & OS_VAL) : 0;
#pragma unroll
                for (int dd = 32; dd >= 1; dd >>= 1) part += __shfl_xor(part, dd, 64);
                excl += part;
                if (ni < nr) break;
                if (nr < 64) {
                    uint64_t wv = 0;
                    if (lane == 0) wv = wait_flag(status + (tt - nr), epoch, err);
                    wv = __shfl(wv, 0, 64);
                    excl += wv & OS_VAL;
                    if (fstate(wv, epoch) == 2) break;
                    tt = tt - nr - 1;
                } else {
                    tt -= 64;
                }
            }
            if (lane == 0) st_flag(my, OS_INC, epoch, excl + (uint64_t)n);
        }
        if (lane == 0) {
            L.excl = excl;
            if (t + 1 == (size_t)B) *total = excl + (uint64_t)n;
        }
    }
    __syncthreads();
    /* the sorted distinct values, in order, at the bucket's place */
    uint64_t *o = dst + L.excl;
    const uint64_t add = base + vmin;
    if (wide)
        for (int i = threadIdx.x; i < n; i += SS_T) o[i] = add + L.stage.v64[i];
    else
        for (int i = threadIdx.x; i < n; i += SS_T) o[i] = add + (uint64_t)L.stage.v32[i];
}

/* ------------------------------------------------------------------------ */
/* host side                                                                 */
/* ------------------------------------------------------------------------ */

static void snote(hipError_t e) { kernels_note_error(e); }

KeyXform make_xform(uint64_t vary)
{
    KeyXform x;
    for (int i = 0; i < 4; i++) x.lo[i] = 0, x.len[i] = 0, x.mask[i] = 0;
    x.nseg = 0;
    x.nbits = 0;
    if (!vary) return x;
    int lo_s[32], len_s[32], ns = 0;
    int b = 63;
    while (b >= 0) {   /* runs of set bits, most significant first */
        if (!((vary >> b) & 1)) {
            b--;
            continue;
        }
        int hi = b;
        while (b >= 0 && ((vary >> b) & 1)) b--;
        lo_s[ns] = b + 1;
        len_s[ns] = hi - b;
        ns++;
    }
    if (ns > 4) {
        /* too fragmented: one contiguous range from the lowest to the highest varying bit */
        int top = 63 - __builtin_clzll(vary), bot = __builtin_ctzll(vary);
        ns = 1;
        lo_s[0] = bot;
        len_s[0] = top - bot + 1;
    }
    for (int i = 0; i < ns; i++) {
        x.lo[i] = lo_s[i];
        x.len[i] = len_s[i];
        x.mask[i] = len_s[i] >= 64 ? ~0ULL : ((1ULL << len_s[i]) - 1);
        x.nbits += len_s[i];
    }
    x.nseg = ns;
    return x;
}

SortCode make_bit_code(uint64_t vary)
{
    SortCode c;
    memset(&c, 0, sizeof c);
    c.x = make_xform(vary);
    c.nbits = c.x.nbits;
    c.vmask = xform_inv(~0ULL, c.x);
    return c;
}

static int bits_for(int values)
{
    int b = 0;
    while ((1 << b) < values) b++;
    return b;
}

SortCode make_pair_code(uint64_t vary, uint64_t cell_max)
{
    SortCode c = make_bit_code(vary);
    if (!vary) return c;
    int nc = 0;
    uint64_t vm = 0;
    for (int cell = 0; cell < 16; cell++) {   /* most significant cell first */
        const int lo = 60 - 4 * cell;
        if ((vary >> lo) & 15) {
            const int base = (int)((cell_max >> lo) & 15) + 1;
            c.cell_lo[nc] = (uint8_t)lo;
            c.cbase[nc] = (uint8_t)(base < 2 ? 2 : base);
            nc++;
            vm |= 15ULL << lo;
        }
    }
    c.cells = 1;
    c.ncell = nc;
    c.npair = nc / 2;
    c.nbits = 0;
    for (int p = 0; p < c.npair; p++) {
        const int ba = c.cbase[2 * p], bb = c.cbase[2 * p + 1];
        c.pbits[p] = (uint8_t)bits_for(ba * bb);
        c.magic[p] = (65536u + (uint32_t)bb - 1) / (uint32_t)bb;
        c.nbits += c.pbits[p];
    }
    c.lbits = (nc & 1) ? bits_for(c.cbase[nc - 1]) : 0;
    c.nbits += c.lbits;
    c.vmask = vm;
    return c;
}

static int g_cell_code = 1;
static unsigned long long g_code_sorts = 0;
void sort_set_cell_code(int on) { g_cell_code = on ? 1 : 0; }
int sort_cell_code() { return g_cell_code; }
unsigned long long sort_code_count() { return g_code_sorts; }

/* ---- configurations ---- */

static const SortCfg k_cfgs[] = {
    {8, 32, 16}, {8, 32, 32}, {8, 16, 16}, {8, 16, 32}, {8, 8, 16},  {8, 8, 32},  {9, 16, 16},
    {9, 16, 32}, {9, 32, 16}, {9, 32, 32}, {9, 7, 16},  {9, 7, 32},  {10, 8, 16}, {10, 8, 32},
};
constexpr int N_CFGS = sizeof(k_cfgs) / sizeof(k_cfgs[0]);

int sort_num_configs() { return N_CFGS; }
SortCfg sort_config(int i) { return k_cfgs[i < 0 ? 0 : (i >= N_CFGS ? N_CFGS - 1 : i)]; }

static bool cfg_valid(const SortCfg &c)
{
    for (int i = 0; i < N_CFGS; i++)
        if (k_cfgs[i].bits == c.bits && k_cfgs[i].kpt == c.kpt && k_cfgs[i].lbw == c.lbw) return true;
    return false;
}

static SortCfg g_table[SORT_SIZE_CLASSES][65];
static bool g_table_init = false;
static SortCfg g_force = {0, 0, 0};

static void table_init()
{
    if (g_table_init) return;
    for (int c = 0; c < SORT_SIZE_CLASSES; c++)
        for (int b = 0; b <= 64; b++) g_table[c][b] = k_cfgs[0];
    g_table_init = true;
}

int sort_size_class(size_t n)
{
    if (n < ((size_t)1 << 20)) return 0;
    if (n < ((size_t)1 << 23)) return 1;
    return 2;
}

int sort_table_set(int cls, int nbits, SortCfg c)
{
    table_init();
    if (cls < 0 || cls >= SORT_SIZE_CLASSES || nbits < 0 || nbits > 64 || !cfg_valid(c)) return -1;
    g_table[cls][nbits] = c;
    return 0;
}

SortCfg sort_table_get(int cls, int nbits)
{
    table_init();
    if (cls < 0 || cls >= SORT_SIZE_CLASSES || nbits < 0 || nbits > 64) return k_cfgs[0];
    return g_table[cls][nbits];
}

void sort_table_reset()
{
    g_table_init = false;
    table_init();
}

int sort_force(int bits, int kpt, int lbw)
{
    g_force.bits = bits;
    g_force.kpt = kpt;
    g_force.lbw = lbw;
    return 0;
}

SortCfg sort_forced() { return g_force; }

/* the configuration a sort of n keys with nbits varying bits uses: the
 * table's entry with any forced field replaced; a forced combination that has
 * no kernel falls back to the nearest one with the forced digit width */
SortCfg sort_choose(size_t n, int nbits)
{
    table_init();
    SortCfg c = g_table[sort_size_class(n)][nbits < 0 ? 0 : (nbits > 64 ? 64 : nbits)];
    if (g_force.bits) c.bits = g_force.bits;
    if (g_force.kpt) c.kpt = g_force.kpt;
    if (g_force.lbw) c.lbw = g_force.lbw;
    if (cfg_valid(c)) return c;
    for (int i = 0; i < N_CFGS; i++)
        if (k_cfgs[i].bits == c.bits && k_cfgs[i].lbw == (c.lbw >= 32 ? 32 : 16)) return k_cfgs[i];
    return k_cfgs[0];
}

/* ---- look-back status array ---- */

/* one look-back status array per sort context: sorts of different contexts
 * may run concurrently (on different streams); sorts of one context run in
 * stream order */
struct SortCtx {
    uint64_t *status = nullptr;
    size_t cap = 0;   /* entries */
    uint32_t epoch = 0;
};
static SortCtx g_ctx[SORT_CONTEXTS];

static bool status_reserve(hipStream_t st, SortCtx &c, size_t entries)
{
    if (entries <= c.cap && c.status) return true;
    size_t cap = std::max(entries + entries / 4, (size_t)1 << 16);
    if (c.status) {
        snote(hipStreamSynchronize(st));
        snote(hipFree(c.status));
        c.status = nullptr;
        c.cap = 0;
    }
    hipError_t e = hipMalloc((void **)&c.status, cap * 8);
    if (e != hipSuccess) {
        snote(e);
        c.status = nullptr;
        return false;
    }
    c.cap = cap;
    snote(hipMemsetAsync(c.status, 0, cap * 8, st));
    c.epoch = 0;
    return true;
}

static uint32_t next_epoch(hipStream_t st, SortCtx &c)
{
    if (++c.epoch > OS_EPOCH_MAX) {
        snote(hipMemsetAsync(c.status, 0, c.cap * 8, st));
        c.epoch = 1;
    }
    return c.epoch;
}

void sort_release()
{
    for (SortCtx &c : g_ctx) {
        if (c.status) hipFree(c.status);
        c = SortCtx();
    }
}

/* ---- launches ---- */

struct OsScratch {
    unsigned *ctr;        /* OS_MAXPASS + 1 tile counters */
    uint64_t *cval;       /* the keys' constant bits */
    unsigned long long *ghist;   /* OS_MAXPASS x OS_MAXR */
    uint64_t *gbase;      /* OS_MAXPASS x OS_MAXR */
    /* sample sort: bucket ticket, then the counts of the buckets' regions
     * (cleared together), the splitters, their guide and its parameters */
    unsigned *ss_ticket;
    unsigned long long *ss_cnt;
    uint64_t *ss_spl;
    unsigned long long *ss_mnx;   /* the sample's ~min and max (in the cleared head) */
};

constexpr size_t OS_CLEAR_BYTES = 256 + (size_t)OS_MAXPASS * OS_MAXR * 8;
constexpr size_t SS_OFF = OS_CLEAR_BYTES + (size_t)OS_MAXPASS * OS_MAXR * 8;
constexpr size_t SS_CLEAR_BYTES = 256 + (size_t)SS_BMAX * SS_NX * 8;

size_t sort_scratch_bytes(size_t)
{
    return SS_OFF + SS_CLEAR_BYTES + (size_t)SS_BMAX * 8 + 256;
}

static OsScratch carve(void *scratch)
{
    char *p = (char *)scratch;
    OsScratch s;
    s.ctr = (unsigned *)p;
    s.cval = (uint64_t *)(p + 128);
    s.ghist = (unsigned long long *)(p + 256);
    s.gbase = (uint64_t *)(p + OS_CLEAR_BYTES);
    s.ss_ticket = (unsigned *)(p + SS_OFF);
    s.ss_cnt = (unsigned long long *)(p + SS_OFF + 256);
    s.ss_spl = (uint64_t *)(p + SS_OFF + SS_CLEAR_BYTES);
    s.ss_mnx = (unsigned long long *)(p + SS_OFF + 64);
    return s;
}

/* ---- sample sort: configuration ---- */

static int g_ss_on = 1;              /* sample sort for sizes it handles */
static int g_ss_bavg = 5000;         /* mean keys per bucket (move candidates) */
static int g_ss_cap = SS_KEYCAP;     /* keys of a bucket (lowered only to test the repeat) */
static int g_ss_bmax = SS_BMAX;      /* most buckets of a sort (SS_BSMALL, SS_BMID or SS_BMAX) */
static int g_ss_probes = SS_PROBES;  /* longest hash-set probe of a bucket the direct sort can take */
/* buckets that fit one block: bitonic sort of every key first (1), or the
 * hash set first with the bitonic sort when the set fills (0, the default:
 * C2 generated in 243-245 ms against 288-300 ms bitonic-first, since a
 * candidate bucket's 5000 keys hold only ~1400 distinct values) */
static int g_ss_bitonic = 0;
/* a bucket's distinct values sorted by the bitonic network when there are
 * at most g_ss_dmax of them (0: never) spanning at least g_ss_dnb bits (two
 * or more 8-bit LSD passes).  C2 bucket kernels, per step: 140.3-141.5 ms
 * with 2048 and 9, 142.6-143.3 ms with LSD passes only, 142.6-145.0 ms
 * with 4096 and 17, 147.2-147.6 ms with 16384 (every bucket). */
static int g_ss_dmax = 2048, g_ss_dnb = 9;
/* partition blocks: 0, one per 8192-key tile; else at most this many,
 * striding over the tiles (the splitters and guide loaded once a block).
 * C2 partitions per table: 23.4 ms at 512 blocks, 27.4 at 256, 25.7 at 768,
 * 25.8 one per tile; 64 batched small tables (16384 buckets): 7.4 ms at
 * 256 to 768, 7.6 at 1024, 11.2 one per tile. */
static int g_ss_part_blocks = 512;
/* the next tile's keys loaded before this tile's reservations, so their
 * latency overlaps the global atomics: C2 partitions 21.2 against 23.3 ms
 * per table, 64 batched tables 7.1 against 7.4 ms */
static int g_ss_part_pf = 1;
static unsigned long long g_ss_sorts = 0;

void sort_set_sample(int on, int bucket_avg, int cap)
{
    g_ss_on = on ? 1 : 0;
    if (bucket_avg > 0) g_ss_bavg = bucket_avg;
    g_ss_cap = (cap > 0 && cap <= SS_KEYCAP) ? cap : SS_KEYCAP;
}
void sort_set_sample_limits(int bmax, int probes)
{
    g_ss_bmax = (bmax == SS_BSMALL || bmax == SS_BMID) ? bmax : SS_BMAX;
    g_ss_probes = probes > 0 ? probes : SS_PROBES;
}
void sort_set_bucket_bitonic(int on) { g_ss_bitonic = on ? 1 : 0; }
void sort_set_part_blocks(int nb) { g_ss_part_blocks = nb > 0 ? nb : 0; }
int sort_part_blocks() { return g_ss_part_blocks; }
void sort_set_part_prefetch(int on) { g_ss_part_pf = on ? 1 : 0; }
int sort_part_prefetch() { return g_ss_part_pf; }
void sort_set_distinct_bitonic(int dmax, int dnb)
{
    g_ss_dmax = dmax < 0 ? 0 : (dmax > SS_LCAP ? SS_LCAP : dmax);
    g_ss_dnb = dnb < 0 ? 0 : (dnb > 64 ? 64 : dnb);
}
int sort_distinct_bitonic_max() { return g_ss_dmax; }
int sort_distinct_bitonic_bits() { return g_ss_dnb; }
int sort_bucket_bitonic() { return g_ss_bitonic; }
int sort_sample_bmax() { return g_ss_bmax; }
int sort_sample_probes() { return g_ss_probes; }
int sort_sample_on() { return g_ss_on; }
int sort_sample_bucket() { return g_ss_bavg; }
int sort_sample_cap() { return g_ss_cap; }
unsigned long long sort_sample_count() { return g_ss_sorts; }

/* Bucket shapes.  Move candidates hold each key about 3.6 times, so a
 * bucket of 5000 to 7000 keys has some 1400 to 2000 distinct values;
 * spawn-4 keys repeat far less, and a bucket of them whose keys span 2^32
 * or more must keep its distinct values within the 8192 of a 64-bit set:
 * 3000 to 3500 keys. */
static SsShape ss_shape(int mode)
{
    if (mode == SORT_SAMPLE_SPARSE) return SsShape{3000, 3500, g_ss_bmax};
    return SsShape{g_ss_bavg, g_ss_bavg > 7000 ? g_ss_bavg : 7000, g_ss_bmax};
}

/* The largest sort a shape takes: sh.bmax buckets of sh.bhi keys (1.15e8
 * keys at 16384 buckets of 7000).  More buckets, not larger ones: batched
 * candidate sets of 57M to 115M keys (their moves already deduplicated per
 * round, so the buckets hold mostly distinct keys spanning 2^32 or more) in
 * 8192 buckets of up to 14000 keys took the batched generation from 88 to
 * 290 ms against the LSD path; in 16384 buckets of 7000, from 75 to 59 ms
 * (64 small tables) and 667 to 524 ms (the specified draw). */
size_t sort_sample_limit(int mode)
{
    if (!g_ss_on || mode <= 0) return 0;
    const SsShape sh = ss_shape(mode);
    return (size_t)sh.bmax * (size_t)sh.bhi;
}

/* the largest bound sort_unique_dev takes: SS_BSMALL buckets of bhi keys */
size_t sort_sample_dev_limit()
{
    return g_ss_on ? (size_t)SS_BSMALL * (size_t)ss_shape(SORT_SAMPLE).bhi : 0;
}

/* buckets of a sample sort of n keys (1: one block), or 0 when the sort
 * takes the LSD path */
static size_t ss_buckets(size_t n, int mode = SORT_SAMPLE)
{
    if (n < 2 || n > sort_sample_limit(mode)) return 0;
    return ss_nbuckets(n, ss_shape(mode));
}

size_t sort_part_keys(size_t n, int mode)
{
    size_t B = ss_buckets(n, mode);
    return B > 1 ? B * (size_t)SS_BSTRIDE : 0;
}

static void launch_hist_scan(hipStream_t st, int bits, const uint64_t *keys, size_t n, const SortCode &c,
                             int npass, const OsScratch &s, unsigned *err)
{
    /* at most two blocks per CU (64 KiB of counters each): each block then
     * clears and flushes its counters once per 8192-key iteration at most */
    size_t hb = (n + 8191) / 8192;
    if (hb > 512) hb = 512;
    int h = prof_hook_begin(2, st);
    switch (bits) {
    case 9:
        hipLaunchKernelGGL(k_os_hist<9>, dim3((unsigned)hb), dim3(HIST_THREADS), 0, st, keys, n, c, npass,
                           s.ghist, s.cval, err);
        break;
    case 10:
        hipLaunchKernelGGL(k_os_hist<10>, dim3((unsigned)hb), dim3(HIST_THREADS), 0, st, keys, n, c, npass,
                           s.ghist, s.cval, err);
        break;
    default:
        hipLaunchKernelGGL(k_os_hist<8>, dim3((unsigned)hb), dim3(HIST_THREADS), 0, st, keys, n, c, npass,
                           s.ghist, s.cval, err);
    }
    snote(hipGetLastError());
    prof_hook_end(h, 8.0 * n, st);
    switch (bits) {
    case 9:
        hipLaunchKernelGGL(k_os_scan<9>, dim3(1), dim3(512), 0, st, s.ghist, npass, s.gbase);
        break;
    case 10:
        hipLaunchKernelGGL(k_os_scan<10>, dim3(1), dim3(1024), 0, st, s.ghist, npass, s.gbase);
        break;
    default:
        hipLaunchKernelGGL(k_os_scan<8>, dim3(1), dim3(256), 0, st, s.ghist, npass, s.gbase);
    }
    snote(hipGetLastError());
}

template <bool RAW>
static void launch_pass_t(hipStream_t st, SortCfg c, unsigned tiles, const uint64_t *src,
                          uint64_t *dst, size_t n, const SortCode &x, int shift, const uint64_t *gbase,
                          uint64_t *status, unsigned *ctr, uint32_t epoch, unsigned *err)
{
#define CG_PASS(B, K, W)                                                                         \
    hipLaunchKernelGGL((k_os_pass<B, K, W, RAW>), dim3(tiles), dim3(pass_threads(B, K)), 0, st, src, \
                       dst, n, x, shift, gbase, status, ctr, epoch, err)
    const bool wide = c.lbw >= 32;
    if (c.bits == 10) {
        if (wide) CG_PASS(10, 8, 32);
        else CG_PASS(10, 8, 16);
    } else if (c.bits == 9) {
        if (c.kpt == 7) {
            if (wide) CG_PASS(9, 7, 32);
            else CG_PASS(9, 7, 16);
        } else if (c.kpt == 32) {
            if (wide) CG_PASS(9, 32, 32);
            else CG_PASS(9, 32, 16);
        } else {
            if (wide) CG_PASS(9, 16, 32);
            else CG_PASS(9, 16, 16);
        }
    } else {
        if (c.kpt == 8) {
            if (wide) CG_PASS(8, 8, 32);
            else CG_PASS(8, 8, 16);
        } else if (c.kpt == 16) {
            if (wide) CG_PASS(8, 16, 32);
            else CG_PASS(8, 16, 16);
        } else {
            if (wide) CG_PASS(8, 32, 32);
            else CG_PASS(8, 32, 16);
        }
    }
#undef CG_PASS
    snote(hipGetLastError());
}

static void launch_pass(hipStream_t st, SortCfg c, unsigned tiles, const uint64_t *src,
                        uint64_t *dst, size_t n, const SortCode &x, int shift, const uint64_t *gbase,
                        uint64_t *status, unsigned *ctr, uint32_t epoch, unsigned *err, bool raw)
{
    if (raw) launch_pass_t<true>(st, c, tiles, src, dst, n, x, shift, gbase, status, ctr, epoch, err);
    else launch_pass_t<false>(st, c, tiles, src, dst, n, x, shift, gbase, status, ctr, epoch, err);
}

/* The sample sort of n keys into dst: *dn keys when dn is set (read on the
 * device; nmax bounds them and sizes the grids), else nmax.  Bmax =
 * ss_nbuckets(nmax) buckets at most, part holding Bmax * SS_BSTRIDE keys when
 * Bmax > 1; keys are only read. */
static void ss_sort_unique(hipStream_t st, const uint64_t *keys, const unsigned long long *dn, size_t nmax,
                           SsShape sh, size_t Bmax, uint64_t *part, const OsScratch &s, SortCtx &sc,
                           uint64_t *dst, unsigned long long *d_nuniq, unsigned *d_err, int *prof_unique,
                           KeyPrune kp)
{
    g_ss_sorts++;
    const uint32_t epoch = next_epoch(st, sc);
    if (Bmax > 1) {
        launch_fill_u64(st, (uint64_t *)s.ss_ticket, (256 + Bmax * SS_NX * 8) / 8, 0);
        int h = prof_hook_begin(8, st);
        /* 8 samples per bucket: SS_S up to SS_BSMALL buckets, SS_SMAX for
         * the largest sorts.  The sample, its sorted runs and (for a larger
         * sample) the ranks go to the start of part, which the partition
         * overwrites only after the splitters are made. */
        const int S = Bmax <= (size_t)SS_BSMALL ? SS_S
                      : Bmax <= 2 * (size_t)SS_BSMALL ? 2 * SS_S
                      : Bmax <= (size_t)SS_BMID ? 4 * SS_S : SS_SMAX;
        const int nrun = S / SS_RUN;
        uint32_t *runs = reinterpret_cast<uint32_t *>(part + S);
        hipLaunchKernelGGL(k_ss_gather, dim3(S / SS_T), dim3(SS_T), 0, st, keys, dn, nmax, sh, S, part,
                           s.ss_mnx);
        hipLaunchKernelGGL(k_ss_runs, dim3(nrun), dim3(SS_T), 0, st, (const uint64_t *)part,
                           (const unsigned long long *)s.ss_mnx, runs, dn, nmax, sh);
        if (S == SS_S) {
            hipLaunchKernelGGL(k_ss_select, dim3(SS_NRUN), dim3(SS_T), 0, st, (const uint32_t *)runs,
                               (const unsigned long long *)s.ss_mnx, dn, nmax, sh, s.ss_spl);
        } else {
            uint32_t *rank = runs + S;
            launch_fill_u64(st, reinterpret_cast<uint64_t *>(rank), (size_t)S / 2, 0);
            hipLaunchKernelGGL(k_ss_rank, dim3((unsigned)(nrun * (nrun / SS_RGRP))), dim3(SS_T), 0, st,
                               (const uint32_t *)runs, nrun, dn, nmax, sh, rank);
            hipLaunchKernelGGL(k_ss_pick, dim3((unsigned)(S / SS_T)), dim3(SS_T), 0, st, (const uint32_t *)runs,
                               (const uint32_t *)rank, (const unsigned long long *)s.ss_mnx, dn, nmax, sh,
                               S, s.ss_spl);
        }
        snote(hipGetLastError());
        prof_hook_end(h, 8.0 * S + 8.0 * (Bmax - 1), st);
        h = prof_hook_begin(9, st);
        size_t tiles = (nmax + SS_TILE - 1) / SS_TILE;
        if (g_ss_part_blocks > 0 && tiles > (size_t)g_ss_part_blocks) tiles = (size_t)g_ss_part_blocks;
#define CG_PART(BM, PF)                                                                                    \
    hipLaunchKernelGGL((k_ss_part<BM, PF>), dim3((unsigned)tiles), dim3(SS_T), 0, st, keys, dn, nmax, sh,     \
                       (const uint64_t *)s.ss_spl, (const unsigned long long *)s.ss_mnx, part, s.ss_cnt, d_err)
        if (Bmax <= (size_t)SS_BSMALL) {
            if (g_ss_part_pf) CG_PART(SS_BSMALL, true);
            else CG_PART(SS_BSMALL, false);
        } else if (Bmax <= (size_t)SS_BMID) {
            if (g_ss_part_pf) CG_PART(SS_BMID, true);
            else CG_PART(SS_BMID, false);
        } else {
            if (g_ss_part_pf) CG_PART(SS_BMAX, true);
            else CG_PART(SS_BMAX, false);
        }
#undef CG_PART
        snote(hipGetLastError());
        prof_hook_end(h, 16.0 * nmax, st);
    }
    *prof_unique = prof_hook_begin(10, st);
    hipLaunchKernelGGL(k_ss_local, dim3((unsigned)Bmax), dim3(SS_T), 0, st, (const uint64_t *)part, keys,
                       (const unsigned long long *)s.ss_cnt, (const uint64_t *)s.ss_spl, dn, nmax, sh, dst,
                       sc.status, s.ss_ticket, epoch, d_nuniq, d_err, g_ss_cap, g_ss_probes, g_ss_bitonic,
                       g_ss_dmax, g_ss_dnb, kp);
    snote(hipGetLastError());
    prof_hook_end(*prof_unique, 8.0 * nmax, st);
}

int sort_unique_dev(hipStream_t st, const uint64_t *keys, const unsigned long long *dn, size_t nmax,
                    void *scratch, uint64_t *dst, unsigned long long *d_nuniq, unsigned *d_err,
                    int *prof_unique, int ctx, uint64_t *part, size_t part_cap, KeyPrune kp)
{
    *prof_unique = -1;
    if (nmax > sort_sample_dev_limit()) return -1;
    SortCtx &sc = g_ctx[ctx < 0 || ctx >= SORT_CONTEXTS ? 0 : ctx];
    /* the grids, the partition variant and the sample size follow the
     * bound, which may be several times the count: bounds of at most
     * SS_BSMALL buckets' worth of keys keep the two-blocks-per-CU partition
     * and the 16384-key sample (larger ones measured slower when the count
     * needed neither) */
    SsShape sh = ss_shape(SORT_SAMPLE);
    sh.bmax = SS_BSMALL;
    const size_t Bmax = ss_nbuckets(nmax, sh);
    if (Bmax > 1 && (!part || part_cap < Bmax * (size_t)SS_BSTRIDE)) return -1;
    if (!status_reserve(st, sc, Bmax)) return -1;
    ss_sort_unique(st, keys, dn, nmax, sh, Bmax, part, carve(scratch), sc, dst, d_nuniq, d_err, prof_unique, kp);
    return 0;
}

int sort_reserve(hipStream_t st, int ctx, size_t entries, unsigned launches)
{
    SortCtx &sc = g_ctx[ctx < 0 || ctx >= SORT_CONTEXTS ? 0 : ctx];
    if (!status_reserve(st, sc, entries)) return -1;
    if (sc.epoch + launches > OS_EPOCH_MAX) {
        snote(hipMemsetAsync(sc.status, 0, sc.cap * 8, st));
        sc.epoch = 0;
    }
    return 0;
}

/* what sort_unique does for n keys: the sample sort with B buckets (B >= 1),
 * or the LSD path with code x, configuration c, npass passes over `tiles`
 * tiles and a compaction over `utiles` tiles */
struct SortPlan {
    size_t B = 0;
    SortCode x;
    SortCfg c = {0, 0, 0};
    int npass = 0;
    bool coded = false;
    size_t tiles = 0, utiles = 0;
    size_t entries = 0;      /* look-back words */
    unsigned launches = 0;   /* epochs */
};

static SortPlan sort_plan(size_t n, uint64_t vary, uint64_t cell_max, int sample)
{
    SortPlan P;
    const size_t B = ss_buckets(n, sample);
    if (B >= 1) {
        P.B = B;
        P.entries = B;
        P.launches = 1;
        return P;
    }
    P.x = make_bit_code(n > 1 ? vary : 0);
    P.c = sort_choose(n, P.x.nbits);
    P.npass = (P.x.nbits + P.c.bits - 1) / P.c.bits;
    if (n > 1 && cell_max != CELLS_UNBOUNDED && g_cell_code) {
        /* the cell-pair code when it saves a pass */
        SortCode pc = make_pair_code(vary, cell_max);
        SortCfg c2 = sort_choose(n, pc.nbits);
        int np2 = (pc.nbits + c2.bits - 1) / c2.bits;
        if (pc.cells && np2 < P.npass) {
            P.coded = true;
            P.x = pc;
            P.c = c2;
            P.npass = np2;
        }
    }
    const size_t tile_keys = (size_t)pass_threads(P.c.bits, P.c.kpt) * P.c.kpt;
    P.tiles = (n + tile_keys - 1) / tile_keys;
    P.utiles = (n + UQ_TILE - 1) / UQ_TILE;
    P.entries = std::max(P.npass ? pass_status_entries(P.tiles, P.c.bits) : 0, P.utiles);
    P.launches = (unsigned)P.npass + 1;
    return P;
}

int sort_prepare(hipStream_t st, int ctx, size_t n, uint64_t vary, uint64_t cell_max, int sample)
{
    if (n == 0) return 0;
    const SortPlan P = sort_plan(n, vary, cell_max, sample);
    return sort_reserve(st, ctx, P.entries, P.launches);
}

uint64_t *sort_unique(hipStream_t st, uint64_t *keys, uint64_t *tmp, size_t n, uint64_t vary,
                      void *scratch, uint64_t *dst, unsigned long long *d_nuniq, unsigned *d_err,
                      int *prof_unique, int ctx, uint64_t cell_max, uint64_t *part, size_t part_cap,
                      int sample, KeyPrune kp)
{
    *prof_unique = -1;
    SortCtx &sc = g_ctx[ctx < 0 || ctx >= SORT_CONTEXTS ? 0 : ctx];
    if (n == 0) {
        launch_fill_u64(st, (uint64_t *)d_nuniq, 1, 0);
        return keys;
    }
    OsScratch s = carve(scratch);
    const size_t Bs = ss_buckets(n, sample);
    const bool sample_ok = Bs == 1 || (Bs > 1 && part && part_cap >= Bs * (size_t)SS_BSTRIDE);
    const SortPlan P = sort_plan(n, vary, cell_max, sample_ok ? sample : 0);
    if (!status_reserve(st, sc, P.entries)) return keys;
    if (P.B >= 1) {
        ss_sort_unique(st, keys, nullptr, n, ss_shape(sample), P.B, part, s, sc, dst, d_nuniq, d_err,
                       prof_unique, kp);
        return keys;
    }
    if (P.coded) g_code_sorts++;
    const SortCode &x = P.x;
    const SortCfg &c = P.c;
    const int npass = P.npass;
    /* tile counters and histograms cleared together */
    launch_fill_u64(st, (uint64_t *)scratch, OS_CLEAR_BYTES / 8, 0);
    uint64_t *src = keys, *dst2 = tmp;
    if (npass > 0) {
        launch_hist_scan(st, c.bits, keys, n, x, npass, s, d_err);
        for (int p = 0; p < npass; p++) {
            int h = prof_hook_begin(3, st);
            launch_pass(st, c, (unsigned)P.tiles, src, dst2, n, x, c.bits * p,
                        s.gbase + ((size_t)p << c.bits), sc.status, s.ctr + p, next_epoch(st, sc), d_err,
                        p == 0);
            prof_hook_end(h, 16.0 * n, st);
            std::swap(src, dst2);
        }
    }
    *prof_unique = prof_hook_begin(4, st);
    if (kp.on)
        hipLaunchKernelGGL(k_os_unique<true>, dim3((unsigned)P.utiles), dim3(UQ_THREADS), 0, st, src, n, dst,
                           sc.status, s.ctr + OS_MAXPASS, next_epoch(st, sc), d_nuniq, d_err,
                           npass > 0 ? 1 : 0, x, s.cval, kp);
    else
        hipLaunchKernelGGL(k_os_unique<false>, dim3((unsigned)P.utiles), dim3(UQ_THREADS), 0, st, src, n, dst,
                           sc.status, s.ctr + OS_MAXPASS, next_epoch(st, sc), d_nuniq, d_err,
                           npass > 0 ? 1 : 0, x, s.cval, kp);
    snote(hipGetLastError());
    prof_hook_end(*prof_unique, 8.0 * n, st);
    return src;
}

} // namespace cg
