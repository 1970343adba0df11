/*
 * kernels.hip -- gfx950 kernels for the layered tablebase path.
 *
 * Everything here is integer/byte work bounded by HBM (or L2/MALL for the
 * lookups) except the FP64 spawn average; there is no dense contraction, so
 * no MFMA.  Design points:
 *   - wave64 throughout: ballots are 64-bit, tiles are multiples of 64 lanes;
 *   - the 128 KiB left-move row table is staged in LDS for the move kernel
 *     (right moves are the nibble-reversed left moves, the lock test is a
 *     bit trick on the pre/post rows, so no second table is needed);
 *   - LSD radix sort over the varying key bits only (canonical formation
 *     cells form a constant prefix), reduce-then-scan passes with a
 *     wave-ballot (match-any) stable ranking inside each 4096-key tile and a
 *     staged, coalesced scatter;
 *   - dedupe is a ballot/popcount stream compaction over the sorted keys;
 *   - every layer carries a prefix-bucket directory over its packed varying
 *     bits (about one bucket per key), so a lookup is one directory read and
 *     a one- or two-step search instead of a full binary search; the FP64 spawn
 *     sums are accumulated per board in ascending cell order (as the reference
 *     does), so the values are bit-identical to a sequential evaluation.
 */
#include <hip/hip_runtime.h>
#include "board.h"
#include "kernels.h"

namespace cg {

static hipError_t g_kerr = hipSuccess;
static void note_err(hipError_t e)
{
    if (e != hipSuccess && g_kerr == hipSuccess) g_kerr = e;
}
#define CGK(x) note_err(x)

hipError_t kernels_take_error()
{
    hipError_t e = g_kerr;
    g_kerr = hipSuccess;
    return e;
}

/* ------------------------------------------------------------------------ */
/* block / wave helpers                                                      */
/* ------------------------------------------------------------------------ */

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

__device__ __forceinline__ uint64_t lanemask_lt()
{
    int l = lane_id();
    return l ? (~0ULL >> (64 - l)) : 0ULL;
}

template <typename T>
__device__ __forceinline__ T wave_incl_scan(T v)
{
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        T u = __shfl_up(v, d, 64);
        if (lane_id() >= d) v += u;
    }
    return v;
}

/* Exclusive scan over the block (blockDim.x <= 1024).  `wtmp` holds 16 T. */
template <typename T>
__device__ __forceinline__ T block_excl_scan(T v, T *total, T *wtmp)
{
    int w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    T inc = wave_incl_scan(v);
    if (lane_id() == 63) wtmp[w] = inc;
    __syncthreads();
    if (threadIdx.x < 64) {
        T x = threadIdx.x < nw ? wtmp[threadIdx.x] : (T)0;
        T xi = wave_incl_scan(x);
        if (threadIdx.x < nw) wtmp[threadIdx.x] = xi - x;
        if (threadIdx.x == nw - 1) wtmp[16] = xi;
    }
    __syncthreads();
    T r = wtmp[w] + inc - v;
    *total = wtmp[16];
    __syncthreads();
    return r;
}

/* OR and AND of per-thread values over the block, one atomic pair per block */
__device__ __forceinline__ void block_or_and(uint64_t o, uint64_t a, unsigned long long *dst_or,
                                             unsigned long long *dst_and, uint64_t *ro, uint64_t *ra)
{
    __syncthreads();   /* ro/ra may still be read by a previous call */
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        o |= __shfl_xor(o, d, 64);
        a &= __shfl_xor(a, d, 64);
    }
    int nw = (blockDim.x + 63) >> 6;
    if (lane_id() == 0) ro[threadIdx.x >> 6] = o, ra[threadIdx.x >> 6] = a;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < nw; w++) o |= ro[w], a &= ra[w];
        o |= ro[0];
        a &= ra[0];
        if (o) atomicOr(dst_or, (unsigned long long)o);
        if (~a) atomicAnd(dst_and, (unsigned long long)a);
    }
}

/* ------------------------------------------------------------------------ */
/* move + lock + prune + canonicalise                                        */
/* ------------------------------------------------------------------------ */

constexpr int MV_THREADS = 1024;

__global__ __launch_bounds__(MV_THREADS) void k_move(const uint64_t *__restrict__ in, size_t n,
                                                     uint64_t *__restrict__ out,
                                                     unsigned long long *stats,
                                                     const uint16_t *__restrict__ glut,
                                                     MoveRules r)
{
    __shared__ uint16_t lut[65536];
    __shared__ uint32_t wtmp[17];
    __shared__ unsigned long long base_sh;
    __shared__ uint64_t ro[16], ra[16];
    uint64_t kor = 0, kand = ~0ULL;
    {
        const uint4 *g = reinterpret_cast<const uint4 *>(glut);
        uint4 *l = reinterpret_cast<uint4 *>(lut);
        for (int i = threadIdx.x; i < 65536 / 8; i += MV_THREADS) l[i] = g[i];
    }
    __syncthreads();
    const size_t stride = (size_t)gridDim.x * MV_THREADS;
    for (size_t b0 = (size_t)blockIdx.x * MV_THREADS; b0 < n; b0 += stride) {
        size_t i = b0 + threadIdx.x;
        uint64_t res[4];
        uint32_t k = 0;
        if (i < n) {
            uint64_t b = in[i];
            if (!(r.prune && cg_prune(b, r.stsl, r.ltc, r.sl))) {
                uint64_t y = cg_rot_cw(b);
#pragma unroll
                for (int d = 0; d < 4; d++) {
                    uint64_t t = d < 2 ? b : y;
                    if (cg_move_hori_tab(lut, &t, d == 1 || d == 2, r.ff) &&
                        !(r.prune && cg_prune(t, r.stsl, r.ltc, r.sl)))
                        res[k++] = cg_canon(t);
                }
            }
        }
        for (uint32_t j = 0; j < k; j++) kor |= res[j], kand &= res[j];
        uint32_t tot;
        uint32_t off = block_excl_scan<uint32_t>(k, &tot, wtmp);
        if (threadIdx.x == 0) base_sh = tot ? atomicAdd(&stats[0], (unsigned long long)tot) : 0;
        __syncthreads();
        unsigned long long base = base_sh;
        for (uint32_t j = 0; j < k; j++) out[base + off + j] = res[j];
        __syncthreads();
    }
    block_or_and(kor, kand, &stats[1], &stats[2], ro, ra);
}

void launch_move(hipStream_t st, const uint64_t *in, size_t n, uint64_t *out,
                 unsigned long long *stats, const uint16_t *lut, MoveRules r)
{
    if (!n) return;
    size_t blocks = (n + MV_THREADS - 1) / MV_THREADS;
    if (blocks > 512) blocks = 512;
    hipLaunchKernelGGL(k_move, dim3((unsigned)blocks), dim3(MV_THREADS), 0, st, in, n, out, stats,
                       lut, r);
}

/* ------------------------------------------------------------------------ */
/* spawn                                                                     */
/* ------------------------------------------------------------------------ */

constexpr int SP_THREADS = 256;

__global__ __launch_bounds__(SP_THREADS) void k_count_empties(const uint64_t *__restrict__ in,
                                                              size_t n, unsigned long long *count)
{
    __shared__ unsigned long long wtmp[17];
    unsigned long long s = 0;
    for (size_t i = (size_t)blockIdx.x * SP_THREADS + threadIdx.x; i < n;
         i += (size_t)gridDim.x * SP_THREADS)
        s += (unsigned long long)cg_empties(in[i]);
    unsigned long long tot;
    block_excl_scan<unsigned long long>(s, &tot, wtmp);
    if (threadIdx.x == 0 && tot) atomicAdd(count, tot);
}

void launch_count_empties(hipStream_t st, const uint64_t *in, size_t n, unsigned long long *count)
{
    if (!n) return;
    size_t blocks = (n + SP_THREADS * 8 - 1) / (SP_THREADS * 8);
    if (blocks > 2048) blocks = 2048;
    hipLaunchKernelGGL(k_count_empties, dim3((unsigned)blocks), dim3(SP_THREADS), 0, st, in, n,
                       count);
}

/* cells in ascending order x = 0..15 correspond to descending nibble position */
__device__ __forceinline__ int next_cell(uint64_t &z)
{
    int pos = 63 - __clzll(z);      /* bit 4*(15-x) */
    z &= ~(1ULL << pos);
    return 15 - (pos >> 2);
}

__global__ __launch_bounds__(SP_THREADS) void k_spawn(const uint64_t *__restrict__ in, size_t n,
                                                      uint64_t *__restrict__ out2,
                                                      uint64_t *__restrict__ out4,
                                                      unsigned long long *stats)
{
    __shared__ uint32_t wtmp[17];
    __shared__ unsigned long long base_sh;
    __shared__ uint64_t ro[16], ra[16];
    uint64_t or2 = 0, and2 = ~0ULL, or4 = 0, and4 = ~0ULL;
    const size_t stride = (size_t)gridDim.x * SP_THREADS;
    for (size_t b0 = (size_t)blockIdx.x * SP_THREADS; b0 < n; b0 += stride) {
        size_t i = b0 + threadIdx.x;
        uint64_t b = i < n ? in[i] : 0;
        uint64_t z = i < n ? cg_zmask(b) : 0;
        uint32_t e = (uint32_t)cg_popc64(z);
        uint32_t tot;
        uint32_t off = block_excl_scan<uint32_t>(e, &tot, wtmp);
        if (threadIdx.x == 0) base_sh = tot ? atomicAdd(&stats[0], (unsigned long long)tot) : 0;
        __syncthreads();
        size_t o = base_sh + off;
        while (z) {
            int x = next_cell(z);
            uint64_t c2 = cg_canon(cg_set_tile(b, x, 1)), c4 = cg_canon(cg_set_tile(b, x, 2));
            out2[o] = c2;
            out4[o] = c4;
            or2 |= c2;
            and2 &= c2;
            or4 |= c4;
            and4 &= c4;
            o++;
        }
        __syncthreads();
    }
    block_or_and(or2, and2, &stats[1], &stats[2], ro, ra);
    block_or_and(or4, and4, &stats[3], &stats[4], ro, ra);
}

void launch_spawn(hipStream_t st, const uint64_t *in, size_t n, uint64_t *out2, uint64_t *out4,
                  unsigned long long *stats)
{
    if (!n) return;
    size_t blocks = (n + SP_THREADS - 1) / SP_THREADS;
    if (blocks > 1024) blocks = 1024;   /* bounds the OR/AND atomics per launch */
    hipLaunchKernelGGL(k_spawn, dim3((unsigned)blocks), dim3(SP_THREADS), 0, st, in, n, out2, out4,
                       stats);
}

/* ------------------------------------------------------------------------ */
/* small fills (graph-safe replacements for memset)                          */
/* ------------------------------------------------------------------------ */

__global__ void k_fill_u64(uint64_t *p, size_t n, uint64_t v)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

void launch_fill_u64(hipStream_t st, uint64_t *p, size_t n, uint64_t v)
{
    if (!n) return;
    size_t blocks = (n + 255) / 256;
    if (blocks > 2048) blocks = 2048;
    hipLaunchKernelGGL(k_fill_u64, dim3((unsigned)blocks), dim3(256), 0, st, p, n, v);
}

/* producer statistics: [0] count, then (OR, AND) pairs; AND starts all-ones */
__global__ void k_init_stats(unsigned long long *s)
{
    if (threadIdx.x < 8) s[threadIdx.x] = (threadIdx.x == 2 || threadIdx.x == 4) ? ~0ULL : 0ULL;
}

void launch_init_stats(hipStream_t st, unsigned long long *stats)
{
    hipLaunchKernelGGL(k_init_stats, dim3(1), dim3(64), 0, st, stats);
}

/* OR / AND of a key array into stats[1], stats[2] */
__global__ __launch_bounds__(256) void k_or_and(const uint64_t *__restrict__ k, size_t n,
                                                unsigned long long *stats)
{
    __shared__ uint64_t ro[4], ra[4];
    uint64_t o = 0, a = ~0ULL;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        uint64_t v = k[i];
        o |= v;
        a &= v;
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        o |= __shfl_xor(o, d, 64);
        a &= __shfl_xor(a, d, 64);
    }
    if (lane_id() == 0) ro[threadIdx.x >> 6] = o, ra[threadIdx.x >> 6] = a;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 4; w++) o |= ro[w], a &= ra[w];
        o |= ro[0];
        a &= ra[0];
        atomicOr(&stats[1], (unsigned long long)o);
        atomicAnd(&stats[2], (unsigned long long)a);
    }
}

void launch_or_and(hipStream_t st, const uint64_t *k, size_t n, unsigned long long *stats)
{
    if (!n) return;
    size_t blocks = (n + 256 * 16 - 1) / (256 * 16);
    if (blocks > 2048) blocks = 2048;
    hipLaunchKernelGGL(k_or_and, dim3((unsigned)blocks), dim3(256), 0, st, k, n, stats);
}

/* ------------------------------------------------------------------------ */
/* onesweep LSD radix sort over the varying key bits + unique                */
/* ------------------------------------------------------------------------ */
/* The bits that differ between keys (OR & ~AND of the key set, produced by
 * the kernels that wrote the keys) are packed into a dense key on the fly
 * (an order-preserving bijection on the key set), so a DPDGap layer with 40
 * varying bits spread over bits 0..51 sorts in 5 passes of 8 bits.  One
 * histogram kernel computes every pass's digit counts in a single read; each
 * pass is one kernel whose tiles take their global offsets from a chained
 * (decoupled look-back) scan over per-tile digit counts.  A look-back wait
 * is bounded: if it ever exceeds its budget the kernel sets an error flag,
 * finishes, and the host reports a failure. */

constexpr int OS_THREADS = 256, OS_WAVES = OS_THREADS / 64;
constexpr int OS_KPT_MIN = 16;   /* keys per thread: 16 or 32 (tiles of 4096 or 8192 keys) */
constexpr int OS_TILE_MIN = OS_THREADS * OS_KPT_MIN;
static int g_sort_kpt = 32;
static int g_sort_lbw = 16;   /* look-back window of the radix passes: 16, 32 or 64 tiles */
constexpr int OS_BITS = 8, OS_R = 1 << OS_BITS;
constexpr int OS_MAXPASS = 8;
constexpr uint64_t OS_AGG = 1ULL << 62, OS_INC = 2ULL << 62, OS_VAL = (1ULL << 62) - 1;
constexpr int OS_SPIN_LIMIT = 1 << 22;

__device__ __forceinline__ uint64_t xform(uint64_t k, const KeyXform &x)
{
    /* branch-free: unused segments have mask 0 and shift 0 */
    uint64_t r = 0;
#pragma unroll
    for (int s = 0; s < 4; s++) {
        uint64_t part = (k >> x.lo[s]) & x.mask[s];
        r = (x.len[s] >= 64 ? 0ULL : (r << x.len[s])) | part;
    }
    return r;
}

__device__ __forceinline__ uint64_t ld_flag(const uint64_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void st_flag(uint64_t *p, uint64_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

/* waits until tile t's status is published; bounded */
__device__ __forceinline__ uint64_t wait_flag(const uint64_t *p, unsigned *err)
{
    uint64_t v;
    int spins = 0;
    while (((v = ld_flag(p)) >> 62) == 0) {
        if (++spins > OS_SPIN_LIMIT) {
            atomicOr(err, 1u);
            return OS_INC;
        }
        __builtin_amdgcn_s_sleep(2);
    }
    return v;
}

/* Exclusive prefix for one chained-scan lane: status of tile i at
 * status[i * stride].  The flags are coherent across XCDs, so every load is
 * a long round trip and the loads themselves cost bandwidth: the window
 * grows 1, 2, 4, ... up to LB_W predecessors per step (usually the nearest
 * one or two already carry an inclusive prefix), sums aggregates
 * nearest-first, and waits only on the first predecessor that has not
 * published yet. */
template <int LB_W>
__device__ __forceinline__ uint64_t lookback(const uint64_t *status, size_t stride, size_t tile,
                                             unsigned *err)
{
    uint64_t excl = 0;
    long t = (long)tile - 1;
    int win = 1;
    while (t >= 0) {
        uint64_t v[LB_W];
#pragma unroll
        for (int j = 0; j < LB_W; j++)
            v[j] = (j < win && t - j >= 0) ? ld_flag(status + (size_t)(t - j) * stride) : OS_INC;
        int j = 0;
        for (; j < win; j++) {
            uint64_t f = v[j] >> 62;
            if (f == 0) break;
            excl += v[j] & OS_VAL;
            if (f == 2) return excl;
        }
        if (j == win) {
            t -= win;
            win = win * 2 < LB_W ? win * 2 : LB_W;
            continue;
        }
        uint64_t w = wait_flag(status + (size_t)(t - j) * stride, err);
        excl += w & OS_VAL;
        if ((w >> 62) == 2) return excl;
        t = t - j - 1;
    }
    return excl;
}

__global__ __launch_bounds__(OS_THREADS) void k_os_hist(const uint64_t *__restrict__ keys, size_t n,
                                                        KeyXform x, int npass, uint32_t *ghist)
{
    __shared__ uint32_t h[OS_WAVES][OS_MAXPASS][OS_R];
    for (int i = threadIdx.x; i < OS_WAVES * OS_MAXPASS * OS_R; i += OS_THREADS)
        (&h[0][0][0])[i] = 0;
    __syncthreads();
    const int w = threadIdx.x >> 6;
    constexpr int HB = 8;   /* keys loaded per thread before counting */
    for (size_t b0 = (size_t)blockIdx.x * OS_THREADS * HB; b0 < n;
         b0 += (size_t)gridDim.x * OS_THREADS * HB) {
        uint64_t kk[HB];
#pragma unroll
        for (int j = 0; j < HB; j++) {
            size_t i = b0 + (size_t)j * OS_THREADS + threadIdx.x;
            kk[j] = i < n ? keys[i] : 0;
        }
#pragma unroll
        for (int j = 0; j < HB; j++) {
            size_t i = b0 + (size_t)j * OS_THREADS + threadIdx.x;
            if (i >= n) continue;
            uint64_t c = xform(kk[j], x);
            for (int p = 0; p < npass; p++)
                atomicAdd(&h[w][p][(c >> (OS_BITS * p)) & (OS_R - 1)], 1u);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < npass * OS_R; i += OS_THREADS) {
        int p = i / OS_R, d = i % OS_R;
        uint32_t t = h[0][p][d] + h[1][p][d] + h[2][p][d] + h[3][p][d];
        if (t) atomicAdd(&ghist[i], t);
    }
}

/* exclusive scan of every pass's 256 digit counts */
__global__ __launch_bounds__(OS_R) void k_os_scan(const uint32_t *ghist, int npass, uint64_t *gbase)
{
    __shared__ uint64_t wtmp[17];
    for (int p = 0; p < npass; p++) {
        uint64_t tot;
        uint64_t ex = block_excl_scan<uint64_t>(ghist[p * OS_R + threadIdx.x], &tot, wtmp);
        gbase[p * OS_R + threadIdx.x] = ex;
    }
}

template <int OS_KPT, int LB_W>
__global__ __launch_bounds__(OS_THREADS) void k_os_pass(const uint64_t *__restrict__ in,
                                                        uint64_t *__restrict__ out, size_t n,
                                                        KeyXform x, int shift,
                                                        const uint64_t *__restrict__ gbase,
                                                        uint64_t *status, unsigned *tile_ctr,
                                                        unsigned *err)
{
    constexpr int OS_TILE = OS_THREADS * OS_KPT, OS_WCHUNK = 64 * OS_KPT;
    __shared__ uint32_t wcnt[OS_WAVES * OS_R];
    __shared__ uint32_t wpre[OS_WAVES * OS_R];
    __shared__ uint32_t tstart[OS_R];
    __shared__ uint64_t obase[OS_R];
    __shared__ uint64_t stage[OS_TILE];
    __shared__ uint32_t wtmp[17];
    __shared__ unsigned tile_sh;

    const int w = threadIdx.x >> 6, lane = lane_id();
    if (threadIdx.x == 0) tile_sh = atomicAdd(tile_ctr, 1u);
    for (int i = threadIdx.x; i < OS_WAVES * OS_R; i += OS_THREADS) wcnt[i] = 0;
    __syncthreads();
    const size_t tile = tile_sh;
    const size_t base = tile * OS_TILE + (size_t)w * OS_WCHUNK;
    uint64_t key[OS_KPT];
    uint32_t rank[OS_KPT];
    const uint64_t lt = lanemask_lt();
#pragma unroll
    for (int j = 0; j < OS_KPT; j++) {
        size_t i = base + (size_t)j * 64 + lane;
        key[j] = i < n ? in[i] : 0;
    }
#pragma unroll
    for (int j = 0; j < OS_KPT; j++) {
        size_t i = base + (size_t)j * 64 + lane;
        bool valid = i < n;
        uint64_t kk = key[j];
        unsigned d = (unsigned)(xform(kk, x) >> shift) & (OS_R - 1);
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < OS_BITS; bit++) {
            bool s = (d >> bit) & 1u;
            uint64_t bb = __ballot(s);
            peers &= s ? bb : ~bb;
        }
        uint32_t before = (uint32_t)__popcll(peers & lt);
        uint32_t cnt = wcnt[w * OS_R + d];
        __builtin_amdgcn_wave_barrier();
        if (valid && before == 0) wcnt[w * OS_R + d] = cnt + (uint32_t)__popcll(peers);
        __builtin_amdgcn_wave_barrier();
        rank[j] = cnt + before;
    }
    __syncthreads();
    /* thread d: wave prefixes and the tile's count of digit d */
    const int d = threadIdx.x;
    uint32_t tot = 0;
#pragma unroll
    for (int ww = 0; ww < OS_WAVES; ww++) {
        wpre[ww * OS_R + d] = tot;
        tot += wcnt[ww * OS_R + d];
    }
    /* chained scan over tiles for digit d */
    uint64_t *my = status + tile * OS_R + d;
    uint64_t excl = 0;
    if (tile == 0) {
        st_flag(my, OS_INC | tot);
    } else {
        st_flag(my, OS_AGG | tot);
        excl = lookback<LB_W>(status + d, OS_R, tile, err);
        st_flag(my, OS_INC | (excl + tot));
    }
    obase[d] = gbase[d] + excl;
    uint32_t all;
    tstart[d] = block_excl_scan<uint32_t>(tot, &all, wtmp);   /* OS_THREADS == OS_R */
    __syncthreads();
#pragma unroll
    for (int j = 0; j < OS_KPT; j++) {
        size_t i = base + (size_t)j * 64 + lane;
        if (i < n) {
            unsigned dd = (unsigned)(xform(key[j], x) >> shift) & (OS_R - 1);
            stage[tstart[dd] + wpre[w * OS_R + dd] + rank[j]] = key[j];
        }
    }
    __syncthreads();
    size_t tile0 = tile * OS_TILE;
    uint32_t tn = (uint32_t)((n - tile0) < (size_t)OS_TILE ? (n - tile0) : (size_t)OS_TILE);
    for (uint32_t p = threadIdx.x; p < tn; p += OS_THREADS) {
        uint64_t kk = stage[p];
        unsigned dd = (unsigned)(xform(kk, x) >> shift) & (OS_R - 1);
        out[obase[dd] + (p - tstart[dd])] = kk;
    }
}

/* unique compaction of a sorted array with a chained scan over tiles;
 * the last tile writes the total to *total */
template <int OS_KPT>
__global__ __launch_bounds__(OS_THREADS) void k_os_unique(const uint64_t *__restrict__ k, size_t n,
                                                          uint64_t *__restrict__ dst,
                                                          uint64_t *status, unsigned *tile_ctr,
                                                          unsigned long long *total, unsigned *err)
{
    constexpr int OS_TILE = OS_THREADS * OS_KPT, OS_WCHUNK = 64 * OS_KPT;
    __shared__ unsigned tile_sh;
    __shared__ uint32_t wc[OS_WAVES];
    __shared__ uint64_t excl_sh;
    const int w = threadIdx.x >> 6, lane = lane_id();
    if (threadIdx.x == 0) tile_sh = atomicAdd(tile_ctr, 1u);
    __syncthreads();
    const size_t tile = tile_sh;
    const size_t base = tile * OS_TILE + (size_t)w * OS_WCHUNK;
    const uint64_t lt = lanemask_lt();
    uint32_t c = 0;
    uint64_t m[OS_KPT], kv[OS_KPT];
    /* one load per key; the predecessor comes from the neighbouring lane,
     * lane 0 of each row reads the key before the row */
#pragma unroll
    for (int j = 0; j < OS_KPT; j++) {
        size_t i = base + (size_t)j * 64 + lane;
        kv[j] = i < n ? k[i] : 0;
    }
    uint64_t prev0[OS_KPT];
#pragma unroll
    for (int j = 0; j < OS_KPT; j++) {
        size_t i0 = base + (size_t)j * 64;
        prev0[j] = (lane == 0 && i0 > 0 && i0 <= n) ? k[i0 - 1] : 0;
    }
#pragma unroll
    for (int j = 0; j < OS_KPT; j++) {
        size_t i = base + (size_t)j * 64 + lane;
        uint64_t pv = __shfl_up(kv[j], 1, 64);
        if (lane == 0) pv = prev0[j];
        bool f = i < n && (i == 0 || kv[j] != pv);
        m[j] = __ballot(f);
        c += (uint32_t)__popcll(m[j]);
    }
    if (lane == 0) wc[w] = c;
    __syncthreads();
    if (w == 0) {
        /* wave 0 publishes and looks back 64 tiles per step (lane i reads tile t - i) */
        uint32_t tot = wc[0] + wc[1] + wc[2] + wc[3];
        uint64_t *my = status + tile;
        uint64_t excl = 0;
        if (tile == 0) {
            if (lane == 0) st_flag(my, OS_INC | tot);
        } else {
            if (lane == 0) st_flag(my, OS_AGG | tot);
            long t = (long)tile - 1;
            while (t >= 0) {
                long ti = t - lane;
                uint64_t v = ti >= 0 ? ld_flag(status + ti) : OS_INC;
                uint64_t f = v >> 62;
                uint64_t notready = __ballot(f == 0), inc = __ballot(f == 2);
                /* nearest lane that is not ready, and nearest inclusive one */
                int nr = notready ? __ffsll((long long)notready) - 1 : 64;
                int ni = inc ? __ffsll((long long)inc) - 1 : 64;
                int stop = nr < ni ? nr : ni;          /* lanes before `stop` are aggregates */
                uint64_t part = (lane < stop || (lane == ni && ni < nr)) ? (v & OS_VAL) : 0;
#pragma unroll
                for (int dd = 32; dd >= 1; dd >>= 1) part += __shfl_xor(part, dd, 64);
                excl += part;
                if (ni < nr) break;                    /* reached an inclusive prefix */
                if (nr < 64) {
                    /* wait for the nearest unpublished tile, then continue past it */
                    uint64_t wv = 0;
                    if (lane == 0) wv = wait_flag(status + (t - nr), err);
                    wv = __shfl(wv, 0, 64);
                    excl += wv & OS_VAL;
                    if ((wv >> 62) == 2) break;
                    t = t - nr - 1;
                } else {
                    t -= 64;
                }
            }
            if (lane == 0) st_flag(my, OS_INC | (excl + tot));
        }
        if (lane == 0) {
            excl_sh = excl;
            if ((tile + 1) * OS_TILE >= n) *total = excl + tot;
        }
    }
    __syncthreads();
    uint64_t o = excl_sh;
    for (int ww = 0; ww < w; ww++) o += wc[ww];
#pragma unroll
    for (int j = 0; j < OS_KPT; j++) {
        size_t i = base + (size_t)j * 64 + lane;
        if ((m[j] >> lane) & 1ULL) dst[o + __popcll(m[j] & lt)] = kv[j];
        (void)i;
        o += __popcll(m[j]);
    }
}

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

size_t sort_scratch_bytes(size_t n)
{
    size_t tiles = (n + OS_TILE_MIN - 1) / OS_TILE_MIN + 1;
    return 4096 + OS_MAXPASS * OS_R * (4 + 8) + tiles * OS_R * 8 + tiles * 8 + 256;
}

struct OsScratch {
    unsigned *ctr;        /* OS_MAXPASS + 1 tile counters */
    uint32_t *ghist;
    uint64_t *gbase;
    uint64_t *status;
    uint64_t *ustatus;
};

static OsScratch carve(void *scratch, size_t n)
{
    size_t tiles = (n + OS_TILE_MIN - 1) / OS_TILE_MIN + 1;
    char *p = (char *)scratch;
    OsScratch s;
    s.ctr = (unsigned *)p;
    p += 256;
    s.ghist = (uint32_t *)p;
    p += OS_MAXPASS * OS_R * 4;
    s.gbase = (uint64_t *)p;
    p += OS_MAXPASS * OS_R * 8;
    s.ustatus = (uint64_t *)p;
    p += (tiles * 8 + 255) / 256 * 256;
    s.status = (uint64_t *)p;
    return s;
}

static void launch_pass(hipStream_t st, int kpt, int lbw, unsigned tiles, const uint64_t *src,
                        uint64_t *dst, size_t n, KeyXform x, int shift, const uint64_t *gbase,
                        uint64_t *status, unsigned *ctr, unsigned *err)
{
#define CG_PASS(K, W)                                                                          \
    hipLaunchKernelGGL((k_os_pass<K, W>), dim3(tiles), dim3(OS_THREADS), 0, st, src, dst, n, x, \
                       shift, gbase, status, ctr, err)
    if (kpt == 32) {
        if (lbw >= 64) CG_PASS(32, 64);
        else if (lbw >= 32) CG_PASS(32, 32);
        else CG_PASS(32, 16);
    } else {
        if (lbw >= 64) CG_PASS(16, 64);
        else if (lbw >= 32) CG_PASS(16, 32);
        else CG_PASS(16, 16);
    }
#undef CG_PASS
}

uint64_t *sort_unique(hipStream_t st, uint64_t *keys, uint64_t *tmp, size_t n, uint64_t vary,
                      void *scratch, uint64_t *dst, unsigned long long *d_nuniq, unsigned *d_err,
                      int *prof_unique)
{
    *prof_unique = -1;
    if (n == 0) {
        launch_fill_u64(st, (uint64_t *)d_nuniq, 1, 0);
        return keys;
    }
    OsScratch s = carve(scratch, n);
    const int kpt = g_sort_kpt == 32 ? 32 : 16;
    const size_t tile_keys = (size_t)OS_THREADS * kpt;
    const size_t tiles = (n + tile_keys - 1) / tile_keys;
    KeyXform x = make_xform(n > 1 ? vary : 0);
    int npass = (x.nbits + OS_BITS - 1) / OS_BITS;
    /* counters, error flag and histograms (256 + 8 KiB) cleared together */
    launch_fill_u64(st, (uint64_t *)scratch, (256 + OS_MAXPASS * OS_R * 4) / 8, 0);
    uint64_t *src = keys, *dst2 = tmp;
    if (npass > 0) {
        size_t hb = (n + OS_TILE_MIN - 1) / OS_TILE_MIN;
        if (hb > 2048) hb = 2048;
        int h = prof_hook_begin(2);
        hipLaunchKernelGGL(k_os_hist, dim3((unsigned)hb), dim3(OS_THREADS), 0, st, keys, n, x, npass,
                           s.ghist);
        prof_hook_end(h, 8.0 * n);
        hipLaunchKernelGGL(k_os_scan, dim3(1), dim3(OS_R), 0, st, s.ghist, npass, s.gbase);
        for (int p = 0; p < npass; p++) {
            launch_fill_u64(st, s.status, tiles * OS_R, 0);
            h = prof_hook_begin(3);
            launch_pass(st, kpt, g_sort_lbw, (unsigned)tiles, src, dst2, n, x, OS_BITS * p,
                        s.gbase + OS_R * p, s.status, s.ctr + p, d_err);
            prof_hook_end(h, 16.0 * n);
            uint64_t *t = src;
            src = dst2;
            dst2 = t;
        }
    }
    launch_fill_u64(st, s.ustatus, tiles, 0);
    *prof_unique = prof_hook_begin(4);
    if (kpt == 32)
        hipLaunchKernelGGL(k_os_unique<32>, dim3((unsigned)tiles), dim3(OS_THREADS), 0, st, src, n,
                           dst, s.ustatus, s.ctr + OS_MAXPASS, d_nuniq, d_err);
    else
        hipLaunchKernelGGL(k_os_unique<16>, dim3((unsigned)tiles), dim3(OS_THREADS), 0, st, src, n,
                           dst, s.ustatus, s.ctr + OS_MAXPASS, d_nuniq, d_err);
    prof_hook_end(*prof_unique, 8.0 * n);
    return src;
}

/* ------------------------------------------------------------------------ */
/* prefix-bucket directory + lookup                                          */
/* ------------------------------------------------------------------------ */

LayerIndexPlan plan_index(size_t n, uint64_t key_or, uint64_t key_and)
{
    LayerIndexPlan p;
    uint64_t vary = n > 1 ? (key_or & ~key_and) : 0;
    p.x = make_xform(vary);
    p.cmask = ~vary;
    p.cval = key_and & p.cmask;
    int lg = 0;
    while (((size_t)1 << (lg + 1)) <= n) lg++;
    int bits = lg < p.x.nbits ? lg : p.x.nbits;
    if (bits > 30) bits = 30;
    p.bits = bits;
    p.shift = p.x.nbits - bits;
    p.dir_entries = ((size_t)1 << bits) + 1;
    return p;
}

/* bucket counts: sorted keys make equal buckets contiguous, so each wave adds
 * one count per run of equal buckets */
__global__ __launch_bounds__(256) void k_dir_count(const uint64_t *__restrict__ keys, size_t n,
                                                   KeyXform x, int shift, uint32_t *cnt)
{
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    bool valid = i < n;
    uint64_t b = valid ? (xform(keys[i], x) >> shift) : ~0ULL;
    uint64_t bp = __shfl_up(b, 1, 64);
    int lane = lane_id();
    bool start = valid && (lane == 0 || bp != b);
    uint64_t starts = __ballot(start);
    uint64_t vmask = __ballot(valid);
    if (start) {
        uint64_t later = starts & ~((2ULL << lane) - 1ULL);   /* run starts after this lane */
        uint64_t endmask = later | ~vmask;
        endmask &= ~((2ULL << lane) - 1ULL);
        int end = endmask ? __ffsll((long long)endmask) - 1 : 64;
        atomicAdd(&cnt[b], (uint32_t)(end - lane));
    }
}

constexpr int DS_THREADS = 1024, DS_IPT = 4, DS_TILE = DS_THREADS * DS_IPT;

__global__ __launch_bounds__(DS_THREADS) void k_dscan_reduce(const uint32_t *__restrict__ in, size_t m,
                                                             uint32_t *sums)
{
    __shared__ uint32_t wtmp[17];
    size_t base = (size_t)blockIdx.x * DS_TILE;
    uint32_t s = 0;
#pragma unroll
    for (int j = 0; j < DS_IPT; j++) {
        size_t i = base + (size_t)j * DS_THREADS + threadIdx.x;
        if (i < m) s += in[i];
    }
    uint32_t tot;
    block_excl_scan<uint32_t>(s, &tot, wtmp);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(DS_THREADS) void k_dscan_top(uint32_t *sums, size_t nb)
{
    __shared__ uint32_t wtmp[17];
    uint32_t carry = 0;
    for (size_t c = 0; c < nb; c += DS_THREADS) {
        size_t i = c + threadIdx.x;
        uint32_t v = i < nb ? sums[i] : 0;
        uint32_t tot;
        uint32_t ex = block_excl_scan<uint32_t>(v, &tot, wtmp);
        if (i < nb) sums[i] = carry + ex;
        carry += tot;
    }
}

/* exclusive scan of cnt[0..m) into dir[0..m) plus dir[m] = total */
__global__ __launch_bounds__(DS_THREADS) void k_dscan_down(const uint32_t *__restrict__ in, size_t m,
                                                           const uint32_t *sums, uint32_t *out,
                                                           uint32_t total)
{
    __shared__ uint32_t wtmp[17];
    size_t base = (size_t)blockIdx.x * DS_TILE + (size_t)threadIdx.x * DS_IPT;
    uint32_t v[DS_IPT];
    uint32_t s = 0;
#pragma unroll
    for (int j = 0; j < DS_IPT; j++) {
        v[j] = base + j < m ? in[base + j] : 0;
        s += v[j];
    }
    uint32_t tot;
    uint32_t ex = block_excl_scan<uint32_t>(s, &tot, wtmp) + sums[blockIdx.x];
#pragma unroll
    for (int j = 0; j < DS_IPT; j++) {
        if (base + j < m) out[base + j] = ex;
        ex += v[j];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) out[m] = total;
}

size_t dir_scratch_bytes(size_t entries)
{
    return entries * 4 + ((entries + DS_TILE - 1) / DS_TILE + 1) * 4 + 1024;
}

void launch_build_dir(hipStream_t st, const uint64_t *keys, size_t n, const LayerIndexPlan &p,
                      uint32_t *dir, void *scratch)
{
    size_t m = p.dir_entries - 1;   /* buckets */
    uint32_t *cnt = (uint32_t *)scratch;
    uint32_t *sums = cnt + ((m + 2 + 63) / 64) * 64;
    launch_fill_u64(st, (uint64_t *)cnt, (m + 1) / 2 + 1, 0);
    if (n) {
        size_t blocks = (n + 255) / 256;
        hipLaunchKernelGGL(k_dir_count, dim3((unsigned)blocks), dim3(256), 0, st, keys, n, p.x,
                           p.shift, cnt);
    }
    size_t nb = (m + DS_TILE - 1) / DS_TILE;
    hipLaunchKernelGGL(k_dscan_reduce, dim3((unsigned)nb), dim3(DS_THREADS), 0, st, cnt, m, sums);
    hipLaunchKernelGGL(k_dscan_top, dim3(1), dim3(DS_THREADS), 0, st, sums, nb);
    hipLaunchKernelGGL(k_dscan_down, dim3((unsigned)nb), dim3(DS_THREADS), 0, st, cnt, m, sums, dir,
                       (uint32_t)n);
}

/* index of q in the layer, or -1: constant bits must match; the bucket's
 * range comes from the directory, then a lower-bound search inside it */
__device__ __forceinline__ long layer_find(uint64_t q, const LayerView &t)
{
    if (t.n == 0 || (q & t.cmask) != t.cval) return -1;
    uint64_t b = xform(q, t.x) >> t.shift;
    size_t lo = t.dir[b], hi = t.dir[b + 1];
    size_t e = hi;
    while (lo < hi) {
        size_t mid = (lo + hi) >> 1;
        if (t.keys[mid] < q) lo = mid + 1;
        else hi = mid;
    }
    return (lo < e && t.keys[lo] == q) ? (long)lo : -1;
}

constexpr int LK_THREADS = 256;

__global__ __launch_bounds__(LK_THREADS) void k_lookup(LayerView t, const uint64_t *__restrict__ q,
                                                       size_t nq, double *__restrict__ out, int canon)
{
    for (size_t i = (size_t)blockIdx.x * LK_THREADS + threadIdx.x; i < nq;
         i += (size_t)gridDim.x * LK_THREADS) {
        uint64_t k = q[i];
        if (canon) k = cg_canon(k);
        long j = layer_find(k, t);
        out[i] = j >= 0 ? t.vals[j] : 0.0;
    }
}

void launch_lookup(hipStream_t st, LayerView t, const uint64_t *q, size_t nq, double *out,
                   int canonicalize)
{
    if (!nq) return;
    size_t blocks = (nq + LK_THREADS - 1) / LK_THREADS;
    if (blocks > 16384) blocks = 16384;
    hipLaunchKernelGGL(k_lookup, dim3((unsigned)blocks), dim3(LK_THREADS), 0, st, t, q, nq, out,
                       canonicalize);
}

/* ------------------------------------------------------------------------ */
/* workload utilities: random boards, canonical form of an array             */
/* ------------------------------------------------------------------------ */

__device__ __forceinline__ uint64_t splitmix64(uint64_t x)
{
    x += 0x9e3779b97f4a7c15ULL;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
    return x ^ (x >> 31);
}

/* Board i: formation cells (nibbles set in fmask) are 0xF.  Skewed mode:
 * every free nibble is 0 with probability p0, otherwise k in 1..kmax with
 * P(k) proportional to 2^-k.  Uniform mode: every free nibble is uniform in
 * 0..kmax.  Random words are splitmix64(seed * 2^40 + 64 * i + j) for
 * j = 0, 1, ...: the empty test uses the low 32 bits of one word,
 * k = 1 + (number of trailing zeros of the high 32 bits), and a k above kmax
 * draws the next word (rejection); uniform values are (word >> 32) mod
 * (kmax + 1). */
__global__ void k_random_boards(uint64_t *out, size_t n, uint64_t seed, uint64_t fmask,
                                uint32_t p0_threshold, int kmax, int uniform)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        uint64_t ctr = (seed << 40) + 64 * (uint64_t)i;
        uint64_t b = 0;
        for (int x = 0; x < 16; x++) {
            int sh = 60 - 4 * x;
            if ((fmask >> sh) & 0xf) {
                b |= 0xfULL << sh;
                continue;
            }
            uint64_t r = splitmix64(ctr++);
            if (uniform) {
                b |= (uint64_t)((uint32_t)(r >> 32) % (uint32_t)(kmax + 1)) << sh;
                continue;
            }
            if ((uint32_t)r < p0_threshold) continue;
            int k;
            for (;;) {
                uint32_t hi = (uint32_t)(r >> 32);
                k = 1 + (hi ? __builtin_ctz(hi) : 32);
                if (k <= kmax) break;
                r = splitmix64(ctr++);
            }
            b |= (uint64_t)k << sh;
        }
        out[i] = b;
    }
}

void launch_random_boards(hipStream_t st, uint64_t *out, size_t n, uint64_t seed, uint64_t fmask,
                          double p_empty, int kmax, int uniform)
{
    if (!n) return;
    uint32_t thr = (uint32_t)(p_empty * 4294967296.0);
    hipLaunchKernelGGL(k_random_boards, dim3(4096), dim3(256), 0, st, out, n, seed, fmask, thr, kmax,
                       uniform);
}

__global__ void k_canon_inplace(uint64_t *a, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        a[i] = cg_canon(a[i]);
}

void launch_canon(hipStream_t st, uint64_t *a, size_t n)
{
    if (!n) return;
    hipLaunchKernelGGL(k_canon_inplace, dim3(4096), dim3(256), 0, st, a, n);
}

/* ------------------------------------------------------------------------ */
/* solve                                                                      */
/* ------------------------------------------------------------------------ */

constexpr int SV_THREADS = 256;

__device__ __forceinline__ bool win_satisfied(uint64_t b, const WinRules &w)
{
    if (w.nox && !w.score) return cg_has_nibble(b, (unsigned)w.nox);
    if (w.score) return false;
    for (int i = 0; i < w.n; i++)
        if ((b & w.mask[i]) == w.val[i]) return true;
    return false;
}

/* max over the 4 moves of the successor value; returns the first strict
 * maximum's direction in *arg (0 when every value is 0). */
__device__ __forceinline__ double best_successor(uint64_t c, const LayerView &t,
                                                 const uint16_t *lut, int ff, int *arg,
                                                 uint32_t &nlook, uint32_t &nhit)
{
    uint64_t y = cg_rot_cw(c);
    double best = 0.0;
    int bd = 0;
#pragma unroll
    for (int d = 0; d < 4; d++) {
        uint64_t m = d < 2 ? c : y;
        if (cg_move_hori_tab(lut, &m, d == 1 || d == 2, ff)) {
            long j = layer_find(cg_canon(m), t);
            nlook++;
            nhit += j >= 0;
            double v = j >= 0 ? t.vals[j] : 0.0;
            if (v > best) {
                best = v;
                bd = d;
            }
        }
    }
    *arg = bd;
    return best;
}

struct SolveSmem {
    uint64_t sb[SV_THREADS];
    uint8_t pb[16 * SV_THREADS], pc[16 * SV_THREADS];
    double m2s[SV_THREADS], m4s[SV_THREADS];
    uint8_t ags[SV_THREADS];
    uint32_t wtmp[17];
};

/* One 256-board tile of a layer: (board, empty cell) pairs are compacted per
 * tile so every lane works on a pair; the FP64 sums of each board run over
 * its cells in ascending order across the pair chunks. */
__device__ __forceinline__ void solve_tile(const uint64_t *__restrict__ M, size_t n, size_t tile,
                                           double *__restrict__ V, uint64_t *__restrict__ argmax,
                                           const LayerView &t2, const LayerView &t4,
                                           const WinRules &win, double p4,
                                           const uint16_t *__restrict__ lut, int ff,
                                           uint32_t &nlook, uint32_t &nhit, SolveSmem &sm)
{
    const double w2 = 1.0 - p4;
    size_t bi = tile * SV_THREADS + threadIdx.x;
    bool act = bi < n;
    uint64_t b = act ? M[bi] : 0;
    bool sat = act && win_satisfied(b, win);
    uint64_t z = (act && !sat) ? cg_zmask(b) : 0;
    uint32_t E = (uint32_t)cg_popc64(z);
    sm.sb[threadIdx.x] = b;
    uint32_t T;
    uint32_t off = block_excl_scan<uint32_t>(E, &T, sm.wtmp);   /* also a barrier */
    {
        uint64_t zz = z;
        uint32_t o = off;
        while (zz) {
            int x = next_cell(zz);
            sm.pb[o] = (uint8_t)threadIdx.x;
            sm.pc[o] = (uint8_t)x;
            o++;
        }
    }
    __syncthreads();
    double a2 = 0.0, a4 = 0.0;
    uint64_t am = 0;
    for (uint32_t c0 = 0; c0 < T; c0 += SV_THREADS) {
        uint32_t p = c0 + threadIdx.x;
        if (p < T) {
            uint64_t pbv = sm.sb[sm.pb[p]];
            int x = sm.pc[p];
            int g2, g4;
            sm.m2s[threadIdx.x] = best_successor(cg_set_tile(pbv, x, 1), t2, lut, ff, &g2, nlook, nhit);
            sm.m4s[threadIdx.x] = best_successor(cg_set_tile(pbv, x, 2), t4, lut, ff, &g4, nlook, nhit);
            sm.ags[threadIdx.x] = (uint8_t)(g2 | (g4 << 2));
        }
        __syncthreads();
        uint32_t q0 = off > c0 ? off : c0;
        uint32_t q1 = off + E < c0 + SV_THREADS ? off + E : c0 + SV_THREADS;
        for (uint32_t q = q0; q < q1; q++) {
            a2 += sm.m2s[q - c0];
            a4 += sm.m4s[q - c0];
            am |= (uint64_t)sm.ags[q - c0] << (4 * sm.pc[q]);
        }
        __syncthreads();
    }
    if (act) {
        double v;
        if (sat) {
            v = 1.0;
        } else {
            v = 0.0;
            if (E) v = (w2 * a2 / (double)E) + (p4 * a4 / (double)E);
            if (win.score && v == 0.0) v = (double)cg_tile_sum(b);
        }
        V[bi] = v;
        if (argmax) argmax[bi] = am;
    }
    __syncthreads();
}

__device__ __forceinline__ void add_stats(unsigned long long *stats, uint32_t nlook, uint32_t nhit)
{
    if (!stats) return;
    unsigned long long l = nlook, h = nhit;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        l += __shfl_xor(l, d, 64);
        h += __shfl_xor(h, d, 64);
    }
    if (lane_id() == 0 && l) {
        atomicAdd(&stats[0], l);
        atomicAdd(&stats[1], h);
    }
}

__global__ __launch_bounds__(SV_THREADS) void k_solve(const uint64_t *__restrict__ M, size_t n,
                                                      double *__restrict__ V,
                                                      uint64_t *__restrict__ argmax, LayerView t2,
                                                      LayerView t4, WinRules win, double p4,
                                                      const uint16_t *__restrict__ lut, int ff,
                                                      unsigned long long *stats)
{
    __shared__ SolveSmem sm;
    uint32_t nlook = 0, nhit = 0;
    for (size_t tile = blockIdx.x; tile * SV_THREADS < n; tile += gridDim.x)
        solve_tile(M, n, tile, V, argmax, t2, t4, win, p4, lut, ff, nlook, nhit, sm);
    add_stats(stats, nlook, nhit);
}

void launch_solve(hipStream_t st, const uint64_t *M, size_t n, double *V, uint64_t *argmax,
                  LayerView t2, LayerView t4, WinRules w, double p4, const uint16_t *lut, int ff,
                  unsigned long long *stats)
{
    if (!n) return;
    size_t tiles = (n + SV_THREADS - 1) / SV_THREADS;
    size_t blocks = tiles < 4096 ? tiles : 4096;
    hipLaunchKernelGGL(k_solve, dim3((unsigned)blocks), dim3(SV_THREADS), 0, st, M, n, V, argmax, t2,
                       t4, w, p4, lut, ff, stats);
}

/* ------------------------------------------------------------------------ */
/* adjoint                                                                    */
/* ------------------------------------------------------------------------ */

constexpr int AD_THREADS = 256;

__global__ __launch_bounds__(AD_THREADS) void k_adjoint(const uint64_t *__restrict__ M, size_t n,
                                                        const double *__restrict__ occ,
                                                        const uint64_t *__restrict__ argmax,
                                                        LayerView t2, double *occ2, LayerView t4,
                                                        double *occ4, WinRules win, double p4,
                                                        const uint16_t *__restrict__ lut, int ff,
                                                        double *D)
{
    __shared__ double red[AD_THREADS / 64];
    const double w2 = 1.0 - p4;
    double contrib = 0.0;
    for (size_t i = (size_t)blockIdx.x * AD_THREADS + threadIdx.x; i < n;
         i += (size_t)gridDim.x * AD_THREADS) {
        double o = occ[i];
        if (o == 0.0) continue;
        uint64_t b = M[i];
        if (win_satisfied(b, win)) continue;
        uint64_t z = cg_zmask(b);
        int E = cg_popc64(z);
        if (!E) continue;
        uint64_t am = argmax[i];
        double a2 = 0.0, a4 = 0.0;
        while (z) {
            int x = next_cell(z);
#pragma unroll
            for (int k = 0; k < 2; k++) {
                const LayerView &t = k ? t4 : t2;
                int d = (int)(am >> (4 * x + 2 * k)) & 3;
                uint64_t m = cg_set_tile(b, x, (unsigned)(k + 1));
                if (d >= 2) m = cg_rot_cw(m);
                if (!cg_move_hori_tab(lut, &m, d == 1 || d == 2, ff)) continue;
                long j = layer_find(cg_canon(m), t);
                if (j < 0) continue;
                double v = t.vals[j];
                if (k) a4 += v;
                else a2 += v;
                if (v > 0.0) atomicAdd(k ? &occ4[j] : &occ2[j], o * (k ? p4 : w2) / (double)E);
            }
        }
        contrib += o * (a4 - a2) / (double)E;
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) contrib += __shfl_xor(contrib, d, 64);
    if (lane_id() == 0) red[threadIdx.x >> 6] = contrib;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < AD_THREADS / 64; w++) s += red[w];
        if (s != 0.0) atomicAdd(D, s);
    }
}

void launch_adjoint(hipStream_t st, const uint64_t *M, size_t n, const double *occ,
                    const uint64_t *argmax, LayerView t2, double *occ2, LayerView t4, double *occ4,
                    WinRules w, double p4, const uint16_t *lut, int ff, double *D)
{
    if (!n) return;
    size_t blocks = (n + AD_THREADS - 1) / AD_THREADS;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_adjoint, dim3((unsigned)blocks), dim3(AD_THREADS), 0, st, M, n, occ, argmax,
                       t2, occ2, t4, occ4, w, p4, lut, ff, D);
}

__global__ void k_adjoint_init(const uint64_t *init, size_t n, LayerView t0, double *occ0,
                               const uint16_t *lut, int ff, double *P)
{
    for (size_t i = threadIdx.x; i < n; i += blockDim.x) {
        uint64_t b = init[i];
        double best = 0.0;
        long bj = -1;
        for (int d = 0; d < 4; d++) {
            uint64_t m = b;
            if (!cg_move_tab(lut, &m, d, ff)) continue;
            long j = layer_find(cg_canon(m), t0);
            double v = j >= 0 ? t0.vals[j] : 0.0;
            if (v > best) {
                best = v;
                bj = j;
            }
        }
        if (bj >= 0) atomicAdd(&occ0[bj], 1.0 / (double)n);
        if (best != 0.0) atomicAdd(P, best / (double)n);
    }
}

void launch_adjoint_init(hipStream_t st, const uint64_t *init, size_t n, LayerView t0, double *occ0,
                         const uint16_t *lut, int ff, double *P)
{
    if (!n) return;
    hipLaunchKernelGGL(k_adjoint_init, dim3(1), dim3(256), 0, st, init, n, t0, occ0, lut, ff, P);
}

/* ------------------------------------------------------------------------ */
/* batched tables: composite keys (table id << 48 | packed free cells)       */
/* ------------------------------------------------------------------------ */
/* A table's canonical boards keep their formation cells (all 0xF) at fixed
 * nibbles, so the free cells alone identify a board; packing them
 * (lowest tile index most significant) keeps the key order within a table,
 * and the table id above them keeps tables apart: one sort / dedupe serves
 * every table of a batch at once. */

__device__ __forceinline__ uint64_t benc(uint64_t b, const BatchTable &T, uint64_t tid, unsigned *err)
{
    if ((b & T.fnib) != T.fnib) atomicOr(err, 2u);   /* formation left its canonical place */
    uint64_t c = 0;
    for (int j = 0; j < T.nfree; j++) c = (c << 4) | ((b >> (60 - 4 * T.pos[j])) & 0xf);
    return (tid << BATCH_TID_SHIFT) | c;
}

__device__ __forceinline__ uint64_t bdec(uint64_t ck, const BatchTable &T)
{
    uint64_t c = ck & ((1ULL << BATCH_TID_SHIFT) - 1);
    uint64_t b = T.fnib;
    for (int j = T.nfree - 1; j >= 0; j--) {
        b |= (c & 0xf) << (60 - 4 * T.pos[j]);
        c >>= 4;
    }
    return b;
}

__global__ __launch_bounds__(MV_THREADS) void k_bmove(const uint64_t *__restrict__ in, size_t n,
                                                      uint64_t *__restrict__ out,
                                                      unsigned long long *stats,
                                                      const uint16_t *__restrict__ glut, MoveRules r,
                                                      const BatchTable *__restrict__ tabs, int step,
                                                      unsigned *err)
{
    __shared__ uint16_t lut[65536];
    __shared__ uint32_t wtmp[17];
    __shared__ unsigned long long base_sh;
    __shared__ uint64_t ro[16], ra[16];
    uint64_t kor = 0, kand = ~0ULL;
    {
        const uint4 *g = reinterpret_cast<const uint4 *>(glut);
        uint4 *l = reinterpret_cast<uint4 *>(lut);
        for (int i = threadIdx.x; i < 65536 / 8; i += MV_THREADS) l[i] = g[i];
    }
    __syncthreads();
    const size_t stride = (size_t)gridDim.x * MV_THREADS;
    for (size_t b0 = (size_t)blockIdx.x * MV_THREADS; b0 < n; b0 += stride) {
        size_t i = b0 + threadIdx.x;
        uint64_t res[4];
        uint32_t k = 0;
        if (i < n) {
            uint64_t ck = in[i];
            uint64_t tid = ck >> BATCH_TID_SHIFT;
            const BatchTable &T = tabs[tid];
            if (step < T.nsteps) {
                uint64_t b = bdec(ck, T);
                if (!(r.prune && cg_prune(b, r.stsl, r.ltc, r.sl))) {
                    uint64_t y = cg_rot_cw(b);
#pragma unroll
                    for (int d = 0; d < 4; d++) {
                        uint64_t t = d < 2 ? b : y;
                        if (cg_move_hori_tab(lut, &t, d == 1 || d == 2, r.ff) &&
                            !(r.prune && cg_prune(t, r.stsl, r.ltc, r.sl)))
                            res[k++] = benc(cg_canon(t), T, tid, err);
                    }
                }
            }
        }
        for (uint32_t j = 0; j < k; j++) kor |= res[j], kand &= res[j];
        uint32_t tot;
        uint32_t off = block_excl_scan<uint32_t>(k, &tot, wtmp);
        if (threadIdx.x == 0) base_sh = tot ? atomicAdd(&stats[0], (unsigned long long)tot) : 0;
        __syncthreads();
        unsigned long long base = base_sh;
        for (uint32_t j = 0; j < k; j++) out[base + off + j] = res[j];
        __syncthreads();
    }
    block_or_and(kor, kand, &stats[1], &stats[2], ro, ra);
}

__global__ __launch_bounds__(SP_THREADS) void k_bspawn(const uint64_t *__restrict__ in, size_t n,
                                                       uint64_t *__restrict__ out2,
                                                       uint64_t *__restrict__ out4,
                                                       unsigned long long *stats,
                                                       const BatchTable *__restrict__ tabs, int step,
                                                       unsigned *err)
{
    __shared__ uint32_t wtmp[17];
    __shared__ unsigned long long base_sh;
    __shared__ uint64_t ro[16], ra[16];
    uint64_t or2 = 0, and2 = ~0ULL, or4 = 0, and4 = ~0ULL;
    const size_t stride = (size_t)gridDim.x * SP_THREADS;
    for (size_t b0 = (size_t)blockIdx.x * SP_THREADS; b0 < n; b0 += stride) {
        size_t i = b0 + threadIdx.x;
        uint64_t b = 0, z = 0, tid = 0;
        if (i < n) {
            uint64_t ck = in[i];
            tid = ck >> BATCH_TID_SHIFT;
            /* the spawns feed step + 1 */
            if (step + 1 < tabs[tid].nsteps) {
                b = bdec(ck, tabs[tid]);
                z = cg_zmask(b);
            }
        }
        uint32_t e = (uint32_t)cg_popc64(z);
        uint32_t tot;
        uint32_t off = block_excl_scan<uint32_t>(e, &tot, wtmp);
        if (threadIdx.x == 0) base_sh = tot ? atomicAdd(&stats[0], (unsigned long long)tot) : 0;
        __syncthreads();
        size_t o = base_sh + off;
        while (z) {
            int x = next_cell(z);
            const BatchTable &T = tabs[tid];
            uint64_t c2 = benc(cg_canon(cg_set_tile(b, x, 1)), T, tid, err);
            uint64_t c4 = benc(cg_canon(cg_set_tile(b, x, 2)), T, tid, err);
            out2[o] = c2;
            out4[o] = c4;
            or2 |= c2;
            and2 &= c2;
            or4 |= c4;
            and4 &= c4;
            o++;
        }
        __syncthreads();
    }
    block_or_and(or2, and2, &stats[1], &stats[2], ro, ra);
    block_or_and(or4, and4, &stats[3], &stats[4], ro, ra);
}

/* composite -> boards; per-table count, OR and AND of the boards
 * (tstat[3 t + 0..2]); keys are sorted, so a wave whose lanes share one
 * table reduces first and adds once */
__global__ __launch_bounds__(256) void k_bdecode(const uint64_t *__restrict__ in, size_t n,
                                                 uint64_t *__restrict__ out,
                                                 const BatchTable *__restrict__ tabs,
                                                 unsigned long long *tstat)
{
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    bool valid = i < n;
    uint64_t tid = ~0ULL, b = 0;
    if (valid) {
        uint64_t ck = in[i];
        tid = ck >> BATCH_TID_SHIFT;
        b = bdec(ck, tabs[tid]);
        out[i] = b;
    }
    uint64_t t0 = __shfl(tid, 0, 64);
    bool same = __ballot(valid && tid == t0) == __ballot(valid) && __ballot(valid) == ~0ULL;
    if (same) {
        uint64_t o = b, a = b;
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) {
            o |= __shfl_xor(o, d, 64);
            a &= __shfl_xor(a, d, 64);
        }
        if (lane_id() == 0) {
            atomicAdd(&tstat[3 * t0], 64ULL);
            atomicOr(&tstat[3 * t0 + 1], (unsigned long long)o);
            atomicAnd(&tstat[3 * t0 + 2], (unsigned long long)a);
        }
    } else if (valid) {
        atomicAdd(&tstat[3 * tid], 1ULL);
        atomicOr(&tstat[3 * tid + 1], (unsigned long long)b);
        atomicAnd(&tstat[3 * tid + 2], (unsigned long long)b);
    }
}

/* per-table bucket counts into one concatenated array (table t's buckets
 * start at plan[t].dir_base); keys are a concatenation of sorted segments */
__global__ __launch_bounds__(256) void k_bdir_count(const uint64_t *__restrict__ keys, size_t n,
                                                    const size_t *__restrict__ seg, int ntab,
                                                    const BatchIndexPlan *__restrict__ plan,
                                                    uint32_t *cnt)
{
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    bool valid = i < n;
    uint64_t slot = ~0ULL;
    if (valid) {
        int lo = 0, hi = ntab;   /* last t with seg[t] <= i */
        while (hi - lo > 1) {
            int mid = (lo + hi) >> 1;
            if (seg[mid] <= i) lo = mid;
            else hi = mid;
        }
        const BatchIndexPlan &p = plan[lo];
        slot = p.dir_base + (xform(keys[i], p.x) >> p.shift);
    }
    uint64_t sp = __shfl_up(slot, 1, 64);
    int lane = lane_id();
    bool start = valid && (lane == 0 || sp != slot);
    uint64_t starts = __ballot(start);
    uint64_t vmask = __ballot(valid);
    if (start) {
        uint64_t endmask = (starts | ~vmask) & ~((2ULL << lane) - 1ULL);
        int end = endmask ? __ffsll((long long)endmask) - 1 : 64;
        atomicAdd(&cnt[slot], (uint32_t)(end - lane));
    }
}

void launch_bmove(hipStream_t st, const uint64_t *in, size_t n, uint64_t *out,
                  unsigned long long *stats, const uint16_t *lut, MoveRules r, const BatchTable *tabs,
                  int step, unsigned *err)
{
    if (!n) return;
    size_t blocks = (n + MV_THREADS - 1) / MV_THREADS;
    if (blocks > 512) blocks = 512;
    hipLaunchKernelGGL(k_bmove, dim3((unsigned)blocks), dim3(MV_THREADS), 0, st, in, n, out, stats,
                       lut, r, tabs, step, err);
}

void launch_bspawn(hipStream_t st, const uint64_t *in, size_t n, uint64_t *out2, uint64_t *out4,
                   unsigned long long *stats, const BatchTable *tabs, int step, unsigned *err)
{
    if (!n) return;
    size_t blocks = (n + SP_THREADS - 1) / SP_THREADS;
    if (blocks > 1024) blocks = 1024;
    hipLaunchKernelGGL(k_bspawn, dim3((unsigned)blocks), dim3(SP_THREADS), 0, st, in, n, out2, out4,
                       stats, tabs, step, err);
}

void launch_bdecode(hipStream_t st, const uint64_t *in, size_t n, uint64_t *out,
                    const BatchTable *tabs, unsigned long long *tstat)
{
    if (!n) return;
    hipLaunchKernelGGL(k_bdecode, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, in, n, out,
                       tabs, tstat);
}

void launch_bdir(hipStream_t st, const uint64_t *keys, size_t n, const size_t *seg, int ntab,
                 const BatchIndexPlan *plan, size_t total_entries, uint32_t *dir, void *scratch)
{
    size_t m = total_entries;
    uint32_t *cnt = (uint32_t *)scratch;
    uint32_t *sums = cnt + ((m + 2 + 63) / 64) * 64;
    launch_fill_u64(st, (uint64_t *)cnt, (m + 1) / 2 + 1, 0);
    if (n)
        hipLaunchKernelGGL(k_bdir_count, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, keys, n,
                           seg, ntab, plan, cnt);
    size_t nb = (m + DS_TILE - 1) / DS_TILE;
    hipLaunchKernelGGL(k_dscan_reduce, dim3((unsigned)nb), dim3(DS_THREADS), 0, st, cnt, m, sums);
    hipLaunchKernelGGL(k_dscan_top, dim3(1), dim3(DS_THREADS), 0, st, sums, nb);
    hipLaunchKernelGGL(k_dscan_down, dim3((unsigned)nb), dim3(DS_THREADS), 0, st, cnt, m, sums, dir,
                       (uint32_t)n);
}

/* batched solve: block b works on table t with desc[t].tile0 <= b <
 * desc[t+1].tile0 */
__global__ __launch_bounds__(SV_THREADS) void k_bsolve(const BSolveDesc *__restrict__ desc, int ntab,
                                                       double p4, const uint16_t *__restrict__ lut,
                                                       int ff)
{
    __shared__ SolveSmem sm;
    __shared__ int t_sh;
    if (threadIdx.x == 0) {
        int lo = 0, hi = ntab;
        while (hi - lo > 1) {
            int mid = (lo + hi) >> 1;
            if (desc[mid].tile0 <= blockIdx.x) lo = mid;
            else hi = mid;
        }
        t_sh = lo;
    }
    __syncthreads();
    const BSolveDesc &d = desc[t_sh];
    uint32_t nlook = 0, nhit = 0;
    size_t tile = blockIdx.x - d.tile0;
    if (tile * SV_THREADS < d.n)
        solve_tile(d.M, d.n, tile, d.V, d.arg, d.t2, d.t4, d.win, p4, lut, ff, nlook, nhit, sm);
}

void launch_bsolve(hipStream_t st, const BSolveDesc *desc, int ntab, size_t total_tiles, double p4,
                   const uint16_t *lut, int ff)
{
    if (!total_tiles) return;
    hipLaunchKernelGGL(k_bsolve, dim3((unsigned)total_tiles), dim3(SV_THREADS), 0, st, desc, ntab, p4,
                       lut, ff);
}

} // namespace cg

namespace cg {
void set_sort_kpt(int kpt) { g_sort_kpt = kpt == 32 ? 32 : 16; }
void set_sort_lbw(int w) { g_sort_lbw = w >= 64 ? 64 : (w >= 32 ? 32 : 16); }
int get_sort_lbw() { return g_sort_lbw; }
int get_sort_kpt() { return g_sort_kpt; }
}
