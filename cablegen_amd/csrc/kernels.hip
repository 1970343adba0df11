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
 *   - the solve keeps a splitter index of each successor layer in LDS and
 *     finishes every lookup with a short binary search in HBM; the FP64 spawn
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

/* ------------------------------------------------------------------------ */
/* move + lock + prune + canonicalise                                        */
/* ------------------------------------------------------------------------ */

constexpr int MV_THREADS = 1024;

__global__ __launch_bounds__(MV_THREADS) void k_move(const uint64_t *__restrict__ in, size_t n,
                                                     uint64_t *__restrict__ out,
                                                     unsigned long long *count,
                                                     const uint16_t *__restrict__ glut,
                                                     MoveRules r)
{
    __shared__ uint16_t lut[65536];
    __shared__ uint32_t wtmp[17];
    __shared__ unsigned long long base_sh;
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
        uint32_t tot;
        uint32_t off = block_excl_scan<uint32_t>(k, &tot, wtmp);
        if (threadIdx.x == 0) base_sh = tot ? atomicAdd(count, (unsigned long long)tot) : 0;
        __syncthreads();
        unsigned long long base = base_sh;
        for (uint32_t j = 0; j < k; j++) out[base + off + j] = res[j];
        __syncthreads();
    }
}

void launch_move(hipStream_t st, const uint64_t *in, size_t n, uint64_t *out,
                 unsigned long long *count, const uint16_t *lut, MoveRules r)
{
    if (!n) return;
    size_t blocks = (n + MV_THREADS - 1) / MV_THREADS;
    if (blocks > 512) blocks = 512;
    hipLaunchKernelGGL(k_move, dim3((unsigned)blocks), dim3(MV_THREADS), 0, st, in, n, out, count,
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
                                                      unsigned long long *count)
{
    __shared__ uint32_t wtmp[17];
    __shared__ unsigned long long base_sh;
    const size_t stride = (size_t)gridDim.x * SP_THREADS;
    for (size_t b0 = (size_t)blockIdx.x * SP_THREADS; b0 < n; b0 += stride) {
        size_t i = b0 + threadIdx.x;
        uint64_t b = i < n ? in[i] : 0;
        uint64_t z = i < n ? cg_zmask(b) : 0;
        uint32_t e = (uint32_t)cg_popc64(z);
        uint32_t tot;
        uint32_t off = block_excl_scan<uint32_t>(e, &tot, wtmp);
        if (threadIdx.x == 0) base_sh = tot ? atomicAdd(count, (unsigned long long)tot) : 0;
        __syncthreads();
        size_t o = base_sh + off;
        while (z) {
            int x = next_cell(z);
            out2[o] = cg_canon(cg_set_tile(b, x, 1));
            out4[o] = cg_canon(cg_set_tile(b, x, 2));
            o++;
        }
        __syncthreads();
    }
}

void launch_spawn(hipStream_t st, const uint64_t *in, size_t n, uint64_t *out2, uint64_t *out4,
                  unsigned long long *count)
{
    if (!n) return;
    size_t blocks = (n + SP_THREADS - 1) / SP_THREADS;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(k_spawn, dim3((unsigned)blocks), dim3(SP_THREADS), 0, st, in, n, out2, out4,
                       count);
}

/* ------------------------------------------------------------------------ */
/* device-wide exclusive scan (u32 -> u64), three kernels                    */
/* ------------------------------------------------------------------------ */

constexpr int SC_THREADS = 1024, SC_IPT = 4, SC_TILE = SC_THREADS * SC_IPT;

__global__ __launch_bounds__(SC_THREADS) void k_scan_reduce(const uint32_t *__restrict__ in,
                                                            size_t m, uint64_t *sums)
{
    __shared__ uint64_t wtmp[17];
    size_t base = (size_t)blockIdx.x * SC_TILE;
    uint64_t s = 0;
#pragma unroll
    for (int j = 0; j < SC_IPT; j++) {
        size_t i = base + (size_t)j * SC_THREADS + threadIdx.x;
        if (i < m) s += in[i];
    }
    uint64_t tot;
    block_excl_scan<uint64_t>(s, &tot, wtmp);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(SC_THREADS) void k_scan_top(uint64_t *sums, size_t nb, uint64_t *total)
{
    __shared__ uint64_t wtmp[17];
    uint64_t carry = 0;
    for (size_t c = 0; c < nb; c += SC_THREADS) {
        size_t i = c + threadIdx.x;
        uint64_t v = i < nb ? sums[i] : 0;
        uint64_t tot;
        uint64_t ex = block_excl_scan<uint64_t>(v, &tot, wtmp);
        if (i < nb) sums[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0 && total) *total = carry;
}

__global__ __launch_bounds__(SC_THREADS) void k_scan_down(const uint32_t *__restrict__ in, size_t m,
                                                          const uint64_t *sums, uint64_t *out)
{
    __shared__ uint64_t wtmp[17];
    size_t base = (size_t)blockIdx.x * SC_TILE + (size_t)threadIdx.x * SC_IPT;
    uint32_t v[SC_IPT];
    uint64_t s = 0;
#pragma unroll
    for (int j = 0; j < SC_IPT; j++) {
        v[j] = base + j < m ? in[base + j] : 0;
        s += v[j];
    }
    uint64_t tot;
    uint64_t ex = block_excl_scan<uint64_t>(s, &tot, wtmp) + sums[blockIdx.x];
#pragma unroll
    for (int j = 0; j < SC_IPT; j++) {
        if (base + j < m) out[base + j] = ex;
        ex += v[j];
    }
}

/* scratch layout for scan of m elements: sums[ceil(m/SC_TILE)] */
static size_t scan_scratch_elems(size_t m) { return (m + SC_TILE - 1) / SC_TILE + 1; }

static void device_scan(hipStream_t st, const uint32_t *in, size_t m, uint64_t *out, uint64_t *sums,
                        uint64_t *total)
{
    size_t nb = (m + SC_TILE - 1) / SC_TILE;
    if (!nb) return;
    hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)nb), dim3(SC_THREADS), 0, st, in, m, sums);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(SC_THREADS), 0, st, sums, nb, total);
    hipLaunchKernelGGL(k_scan_down, dim3((unsigned)nb), dim3(SC_THREADS), 0, st, in, m, sums, out);
}

/* ------------------------------------------------------------------------ */
/* LSD radix sort (keys only, u64) + unique                                  */
/* ------------------------------------------------------------------------ */

constexpr int RS_THREADS = 256, RS_WAVES = RS_THREADS / 64, RS_KPT = 16;
constexpr int RS_TILE = RS_THREADS * RS_KPT;       /* 4096 keys */
constexpr int RS_WCHUNK = 64 * RS_KPT;              /* 1024 keys per wave */
constexpr int RS_BITS = 8, RS_R = 1 << RS_BITS;

__global__ __launch_bounds__(RS_THREADS) void k_diff_or(const uint64_t *__restrict__ k, size_t n,
                                                        unsigned long long *acc)
{
    __shared__ uint64_t red[RS_WAVES];
    uint64_t k0 = k[0], m = 0;
    for (size_t i = (size_t)blockIdx.x * RS_THREADS + threadIdx.x; i < n;
         i += (size_t)gridDim.x * RS_THREADS)
        m |= k[i] ^ k0;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) m |= __shfl_xor(m, d, 64);
    if (lane_id() == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t t = 0;
        for (int w = 0; w < RS_WAVES; w++) t |= red[w];
        if (t) atomicOr(acc, (unsigned long long)t);
    }
}

__global__ __launch_bounds__(RS_THREADS) void k_rs_hist(const uint64_t *__restrict__ keys, size_t n,
                                                        int shift, uint32_t *counts, size_t nblocks)
{
    __shared__ uint32_t h[RS_R];
    for (int i = threadIdx.x; i < RS_R; i += RS_THREADS) h[i] = 0;
    __syncthreads();
    size_t base = (size_t)blockIdx.x * RS_TILE;
#pragma unroll 4
    for (int j = 0; j < RS_KPT; j++) {
        size_t i = base + (size_t)j * RS_THREADS + threadIdx.x;
        if (i < n) atomicAdd(&h[(unsigned)(keys[i] >> shift) & (RS_R - 1)], 1u);
    }
    __syncthreads();
    for (int d = threadIdx.x; d < RS_R; d += RS_THREADS)
        counts[(size_t)d * nblocks + blockIdx.x] = h[d];
}

/* offs[d * nblocks + b] = global output position of the first key of digit d
 * in tile b (exclusive scan over the digit-major count matrix). */
__global__ __launch_bounds__(RS_THREADS) void k_rs_scatter(const uint64_t *__restrict__ in,
                                                           uint64_t *__restrict__ out, size_t n,
                                                           int shift, const uint64_t *__restrict__ offs,
                                                           size_t nblocks)
{
    __shared__ uint32_t wcnt[RS_WAVES * RS_R];
    __shared__ uint32_t wpre[RS_WAVES * RS_R];
    __shared__ uint32_t tstart[RS_R];
    __shared__ uint64_t gbase[RS_R];
    __shared__ uint64_t stage[RS_TILE];
    __shared__ uint32_t wtmp[17];

    const int w = threadIdx.x >> 6, lane = lane_id();
    for (int i = threadIdx.x; i < RS_WAVES * RS_R; i += RS_THREADS) wcnt[i] = 0;
    __syncthreads();

    const size_t base = (size_t)blockIdx.x * RS_TILE + (size_t)w * RS_WCHUNK;
    uint64_t key[RS_KPT];
    uint32_t rank[RS_KPT];
    const uint64_t lt = lanemask_lt();
#pragma unroll
    for (int j = 0; j < RS_KPT; j++) {
        size_t i = base + (size_t)j * 64 + lane;
        bool valid = i < n;
        uint64_t kk = valid ? in[i] : 0;
        key[j] = kk;
        unsigned d = (unsigned)(kk >> shift) & (RS_R - 1);
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < RS_BITS; bit++) {
            bool s = (d >> bit) & 1u;
            uint64_t bb = __ballot(s);
            peers &= s ? bb : ~bb;
        }
        uint32_t before = (uint32_t)__popcll(peers & lt);
        uint32_t cnt = wcnt[w * RS_R + d];
        __builtin_amdgcn_wave_barrier();
        if (valid && before == 0) wcnt[w * RS_R + d] = cnt + (uint32_t)__popcll(peers);
        __builtin_amdgcn_wave_barrier();
        rank[j] = cnt + before;
    }
    __syncthreads();
    /* per digit: exclusive prefix over waves, tile total, tile start */
    uint32_t tot_d = 0;
    for (int d = threadIdx.x; d < RS_R; d += RS_THREADS) {
        uint32_t s = 0;
#pragma unroll
        for (int ww = 0; ww < RS_WAVES; ww++) {
            wpre[ww * RS_R + d] = s;
            s += wcnt[ww * RS_R + d];
        }
        tot_d = s;
        gbase[d] = offs[(size_t)d * nblocks + blockIdx.x];
    }
    uint32_t all;
    uint32_t st = block_excl_scan<uint32_t>(tot_d, &all, wtmp);   /* RS_THREADS == RS_R */
    tstart[threadIdx.x] = st;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < RS_KPT; j++) {
        size_t i = base + (size_t)j * 64 + lane;
        if (i < n) {
            unsigned d = (unsigned)(key[j] >> shift) & (RS_R - 1);
            stage[tstart[d] + wpre[w * RS_R + d] + rank[j]] = key[j];
        }
    }
    __syncthreads();
    size_t tile0 = (size_t)blockIdx.x * RS_TILE;
    uint32_t tn = (uint32_t)((n - tile0) < (size_t)RS_TILE ? (n - tile0) : (size_t)RS_TILE);
    for (uint32_t p = threadIdx.x; p < tn; p += RS_THREADS) {
        uint64_t kk = stage[p];
        unsigned d = (unsigned)(kk >> shift) & (RS_R - 1);
        out[gbase[d] + (p - tstart[d])] = kk;
    }
}

/* unique compaction over sorted keys: per 1024-key wave chunk */
__global__ __launch_bounds__(RS_THREADS) void k_uniq_count(const uint64_t *__restrict__ k, size_t n,
                                                           uint32_t *cc)
{
    const int w = threadIdx.x >> 6, lane = lane_id();
    size_t chunk = (size_t)blockIdx.x * RS_WAVES + w;
    size_t base = chunk * RS_WCHUNK;
    if (base >= n) return;
    uint32_t c = 0;
#pragma unroll 4
    for (int j = 0; j < RS_KPT; j++) {
        size_t i = base + (size_t)j * 64 + lane;
        bool f = i < n && (i == 0 || k[i] != k[i - 1]);
        c += (uint32_t)__popcll(__ballot(f));
    }
    if (lane == 0) cc[chunk] = c;
}

__global__ __launch_bounds__(RS_THREADS) void k_uniq_write(const uint64_t *__restrict__ k, size_t n,
                                                           const uint64_t *__restrict__ co,
                                                           uint64_t *__restrict__ dst)
{
    const int w = threadIdx.x >> 6, lane = lane_id();
    size_t chunk = (size_t)blockIdx.x * RS_WAVES + w;
    size_t base = chunk * RS_WCHUNK;
    if (base >= n) return;
    uint64_t o = co[chunk];
    const uint64_t lt = lanemask_lt();
#pragma unroll 4
    for (int j = 0; j < RS_KPT; j++) {
        size_t i = base + (size_t)j * 64 + lane;
        uint64_t kk = i < n ? k[i] : 0;
        bool f = i < n && (i == 0 || kk != k[i - 1]);
        uint64_t m = __ballot(f);
        if (f) dst[o + __popcll(m & lt)] = kk;
        o += __popcll(m);
    }
}

size_t sort_scratch_bytes(size_t n)
{
    size_t nb = (n + RS_TILE - 1) / RS_TILE;
    size_t m = nb * RS_R;                          /* count matrix */
    size_t chunks = (n + RS_WCHUNK - 1) / RS_WCHUNK;
    size_t mm = m > chunks ? m : chunks;
    return 64 + mm * 4 + mm * 8 + scan_scratch_elems(mm) * 8 + 64;
}

struct ScratchLayout {
    unsigned long long *diff;
    uint64_t *total;
    uint32_t *counts;
    uint64_t *offs;
    uint64_t *sums;
};

static ScratchLayout carve(void *scratch, size_t n)
{
    size_t nb = (n + RS_TILE - 1) / RS_TILE;
    size_t m = nb * RS_R;
    size_t chunks = (n + RS_WCHUNK - 1) / RS_WCHUNK;
    size_t mm = m > chunks ? m : chunks;
    char *p = (char *)scratch;
    ScratchLayout s;
    s.diff = (unsigned long long *)p;
    s.total = (uint64_t *)(p + 8);
    p += 64;
    s.counts = (uint32_t *)p;
    p += (mm * 4 + 7) / 8 * 8;
    s.offs = (uint64_t *)p;
    p += mm * 8;
    s.sums = (uint64_t *)p;
    return s;
}

SortState sort_unique_begin(hipStream_t st, uint64_t *keys, uint64_t *tmp, size_t n, void *scratch)
{
    SortState ss;
    ss.sorted = keys;
    ss.n = n;
    ss.n_unique = 0;
    ss.chunk_off = nullptr;
    if (n == 0) return ss;
    ScratchLayout s = carve(scratch, n);
    unsigned long long h_diff = 0;
    if (n > 1) {
        CGK(hipMemsetAsync(s.diff, 0, 8, st));
        size_t blocks = (n + RS_THREADS * 16 - 1) / (RS_THREADS * 16);
        if (blocks > 4096) blocks = 4096;
        hipLaunchKernelGGL(k_diff_or, dim3((unsigned)blocks), dim3(RS_THREADS), 0, st, keys, n,
                           s.diff);
        CGK(hipMemcpyAsync(&h_diff, s.diff, 8, hipMemcpyDeviceToHost, st));
        CGK(hipStreamSynchronize(st));
    }
    uint64_t *src = keys, *dst = tmp;
    if (h_diff) {
        int lo = __builtin_ctzll(h_diff), hi = 64 - __builtin_clzll(h_diff);
        size_t nb = (n + RS_TILE - 1) / RS_TILE;
        for (int shift = lo; shift < hi; shift += RS_BITS) {
            int h = prof_hook_begin(2);
            hipLaunchKernelGGL(k_rs_hist, dim3((unsigned)nb), dim3(RS_THREADS), 0, st, src, n,
                               shift, s.counts, nb);
            prof_hook_end(h, 8.0 * n);
            device_scan(st, s.counts, nb * RS_R, s.offs, s.sums, nullptr);
            h = prof_hook_begin(3);
            hipLaunchKernelGGL(k_rs_scatter, dim3((unsigned)nb), dim3(RS_THREADS), 0, st, src, dst,
                               n, shift, s.offs, nb);
            prof_hook_end(h, 16.0 * n);
            uint64_t *t = src;
            src = dst;
            dst = t;
        }
    }
    /* unique count */
    size_t chunks = (n + RS_WCHUNK - 1) / RS_WCHUNK;
    size_t ublocks = (chunks + RS_WAVES - 1) / RS_WAVES;
    hipLaunchKernelGGL(k_uniq_count, dim3((unsigned)ublocks), dim3(RS_THREADS), 0, st, src, n,
                       s.counts);
    device_scan(st, s.counts, chunks, s.offs, s.sums, s.total);
    uint64_t h_total = 0;
    CGK(hipMemcpyAsync(&h_total, s.total, 8, hipMemcpyDeviceToHost, st));
    CGK(hipStreamSynchronize(st));
    ss.sorted = src;
    ss.n_unique = (size_t)h_total;
    ss.chunk_off = s.offs;
    return ss;
}

void sort_unique_finish(hipStream_t st, const SortState &ss, uint64_t *dst)
{
    if (ss.n == 0) return;
    size_t chunks = (ss.n + RS_WCHUNK - 1) / RS_WCHUNK;
    size_t ublocks = (chunks + RS_WAVES - 1) / RS_WAVES;
    hipLaunchKernelGGL(k_uniq_write, dim3((unsigned)ublocks), dim3(RS_THREADS), 0, st, ss.sorted,
                       ss.n, ss.chunk_off, dst);
}

/* ------------------------------------------------------------------------ */
/* splitter index + lookup                                                    */
/* ------------------------------------------------------------------------ */

__global__ void k_build_split(const uint64_t *__restrict__ keys, size_t n, uint64_t *split, int ns)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ns) split[i] = keys[(size_t)i * n / (size_t)ns];
}

void launch_build_split(hipStream_t st, const uint64_t *keys, size_t n, uint64_t *split, int ns)
{
    if (!n || ns <= 0) return;
    hipLaunchKernelGGL(k_build_split, dim3((ns + 255) / 256), dim3(256), 0, st, keys, n, split, ns);
}

/* index of q in the layer, or -1: the last splitter <= q (LDS) bounds the
 * range, then a lower-bound binary search over that range in HBM. */
__device__ __forceinline__ long layer_find(uint64_t q, const LayerView &t, const uint64_t *spl)
{
    if (t.n == 0 || q < spl[0]) return -1;
    int lo = 0, hi = t.ns;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (spl[mid] <= q) lo = mid;
        else hi = mid;
    }
    size_t a = (size_t)lo * t.n / (size_t)t.ns;
    size_t e = (size_t)(lo + 1) * t.n / (size_t)t.ns;
    size_t b = e;
    while (a < b) {
        size_t mid = (a + b) >> 1;
        if (t.keys[mid] < q) a = mid + 1;
        else b = mid;
    }
    return (a < e && t.keys[a] == q) ? (long)a : -1;
}

constexpr int NS_MAX = 1024;
constexpr int LK_THREADS = 256;

__device__ __forceinline__ void load_split(uint64_t *dst, const LayerView &t)
{
    for (int i = threadIdx.x; i < t.ns; i += blockDim.x) dst[i] = t.split[i];
}

__global__ __launch_bounds__(LK_THREADS) void k_lookup(LayerView t, const uint64_t *__restrict__ q,
                                                       size_t nq, double *__restrict__ out, int canon)
{
    __shared__ uint64_t spl[NS_MAX];
    load_split(spl, t);
    __syncthreads();
    for (size_t i = (size_t)blockIdx.x * LK_THREADS + threadIdx.x; i < nq;
         i += (size_t)gridDim.x * LK_THREADS) {
        uint64_t k = q[i];
        if (canon) k = cg_canon(k);
        long j = layer_find(k, t, spl);
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
__device__ __forceinline__ double best_successor(uint64_t c, const LayerView &t, const uint64_t *spl,
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
            long j = layer_find(cg_canon(m), t, spl);
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

__global__ __launch_bounds__(SV_THREADS) void k_solve(const uint64_t *__restrict__ M, size_t n,
                                                      double *__restrict__ V,
                                                      uint64_t *__restrict__ argmax, LayerView t2,
                                                      LayerView t4, WinRules win, double p4,
                                                      const uint16_t *__restrict__ lut, int ff,
                                                      unsigned long long *stats)
{
    __shared__ uint64_t spl2[NS_MAX], spl4[NS_MAX];
    __shared__ uint64_t sb[SV_THREADS];
    __shared__ uint8_t pb[16 * SV_THREADS], pc[16 * SV_THREADS];
    __shared__ double m2s[SV_THREADS], m4s[SV_THREADS];
    __shared__ uint8_t ags[SV_THREADS];
    __shared__ uint32_t wtmp[17];
    load_split(spl2, t2);
    load_split(spl4, t4);
    const double w2 = 1.0 - p4;
    uint32_t nlook = 0, nhit = 0;
    for (size_t tile = blockIdx.x; tile * SV_THREADS < n; tile += gridDim.x) {
        size_t bi = tile * SV_THREADS + threadIdx.x;
        bool act = bi < n;
        uint64_t b = act ? M[bi] : 0;
        bool sat = act && win_satisfied(b, win);
        uint64_t z = (act && !sat) ? cg_zmask(b) : 0;
        uint32_t E = (uint32_t)cg_popc64(z);
        sb[threadIdx.x] = b;
        uint32_t T;
        uint32_t off = block_excl_scan<uint32_t>(E, &T, wtmp);   /* also a barrier */
        {
            uint64_t zz = z;
            uint32_t o = off;
            while (zz) {
                int x = next_cell(zz);
                pb[o] = (uint8_t)threadIdx.x;
                pc[o] = (uint8_t)x;
                o++;
            }
        }
        __syncthreads();
        double a2 = 0.0, a4 = 0.0;
        uint64_t am = 0;
        for (uint32_t c0 = 0; c0 < T; c0 += SV_THREADS) {
            uint32_t p = c0 + threadIdx.x;
            if (p < T) {
                uint64_t pbv = sb[pb[p]];
                int x = pc[p];
                int g2, g4;
                m2s[threadIdx.x] = best_successor(cg_set_tile(pbv, x, 1), t2, spl2, lut, ff, &g2,
                                                  nlook, nhit);
                m4s[threadIdx.x] = best_successor(cg_set_tile(pbv, x, 2), t4, spl4, lut, ff, &g4,
                                                  nlook, nhit);
                ags[threadIdx.x] = (uint8_t)(g2 | (g4 << 2));
            }
            __syncthreads();
            uint32_t q0 = off > c0 ? off : c0;
            uint32_t q1 = off + E < c0 + SV_THREADS ? off + E : c0 + SV_THREADS;
            for (uint32_t q = q0; q < q1; q++) {
                a2 += m2s[q - c0];
                a4 += m4s[q - c0];
                am |= (uint64_t)ags[q - c0] << (4 * pc[q]);
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
    if (stats) {
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
    __shared__ uint64_t spl2[NS_MAX], spl4[NS_MAX];
    __shared__ double red[AD_THREADS / 64];
    load_split(spl2, t2);
    load_split(spl4, t4);
    __syncthreads();
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
                const uint64_t *spl = k ? spl4 : spl2;
                int d = (int)(am >> (4 * x + 2 * k)) & 3;
                uint64_t m = cg_set_tile(b, x, (unsigned)(k + 1));
                if (d >= 2) m = cg_rot_cw(m);
                if (!cg_move_hori_tab(lut, &m, d == 1 || d == 2, ff)) continue;
                long j = layer_find(cg_canon(m), t, spl);
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
    __shared__ uint64_t spl[NS_MAX];
    load_split(spl, t0);
    __syncthreads();
    for (size_t i = threadIdx.x; i < n; i += blockDim.x) {
        uint64_t b = init[i];
        double best = 0.0;
        long bj = -1;
        for (int d = 0; d < 4; d++) {
            uint64_t m = b;
            if (!cg_move_tab(lut, &m, d, ff)) continue;
            long j = layer_find(cg_canon(m), t0, spl);
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

} // namespace cg
