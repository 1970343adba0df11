/*
 * kernels.hip -- gfx950 kernels for the layered tablebase path.
 *
 * Everything here is integer/byte work bounded by HBM (or L2/MALL for the
 * lookups) except the FP64 spawn average; there is no dense contraction, so
 * no MFMA.  Design points:
 *   - wave64 throughout: ballots are 64-bit, tiles are multiples of 64 lanes;
 *   - one 128 KiB left-move row table (a right move is the left move of the
 *     mirrored board, mirrored back; the lock test is a bit trick on the
 *     pre/post boards), so no second table is needed;
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
#include "devutil.h"

namespace cg {

static hipError_t g_kerr = hipSuccess;
static void note_err(hipError_t e)
{
    if (e != hipSuccess && g_kerr == hipSuccess) g_kerr = e;
}
#define CGK(x) note_err(x)

void kernels_note_error(hipError_t e) { note_err(e); }

hipError_t kernels_take_error()
{
    hipError_t e = g_kerr;
    g_kerr = hipSuccess;
    return e;
}

/* ------------------------------------------------------------------------ */
/* move + lock + prune + canonicalise                                        */
/* ------------------------------------------------------------------------ */

constexpr int MV_THREADS = 1024;
constexpr int MV_HASH = 8192;   /* hash slots per 2048-board round (<= 8192 candidates) */

/* Inserts c into the block's LDS hash set (open addressing with a 64-bit
 * compare-and-swap; 0 marks an empty slot, a moved board always holds a
 * tile); true on its first occurrence.  A set that holds at most MV_HASH
 * distinct keys always finds c or a free slot. */
__device__ __forceinline__ bool mv_first(unsigned long long *ht, uint64_t c)
{
    uint32_t sl = (uint32_t)((c * 0x9E3779B97F4A7C15ULL) >> (64 - 13));
    for (int probe = 0; probe < MV_HASH; probe++) {
        unsigned long long old = atomicCAS(&ht[sl], 0ULL, (unsigned long long)c);
        if (old == 0ULL) return true;
        if (old == c) return false;
        sl = (sl + 1) & (MV_HASH - 1);
    }
    return true;
}

/* The move primitive (cg_dev_move): every candidate is appended, and the
 * row table is staged in LDS. */
__global__ __launch_bounds__(MV_THREADS, 1) void k_move(
    const uint64_t *__restrict__ in, size_t n, const unsigned long long *dn,
    uint64_t *__restrict__ out, unsigned long long *stats, const uint16_t *__restrict__ glut,
    MoveRules r)
{
    __shared__ uint4 lds_lut[65536 / 8];   /* the row table, 128 KiB */
    __shared__ uint32_t wtmp[17];
    __shared__ unsigned long long base_sh;
    __shared__ uint64_t ro[16], ra[16];
    uint64_t kor = 0, kand = ~0ULL;
    {
        const uint4 *g = reinterpret_cast<const uint4 *>(glut);
        for (int i = threadIdx.x; i < 65536 / 8; i += MV_THREADS) lds_lut[i] = g[i];
        __syncthreads();
    }
    const uint16_t *lut = reinterpret_cast<const uint16_t *>(lds_lut);
    if (dn && *dn < n) n = (size_t)*dn;   /* count produced on the device */
    const size_t stride = (size_t)gridDim.x * MV_THREADS;
    for (size_t b0 = (size_t)blockIdx.x * MV_THREADS; b0 < n; b0 += stride) {
        const size_t i = b0 + threadIdx.x;
        uint64_t res[4];
        uint32_t k = 0;
        if (i < n) {
            const uint64_t b = in[i];
            if (!(r.prune && r.prune_inputs && cg_prune(b, r.stsl, r.ltc, r.sl))) {
                uint64_t y = cg_rot_cw(b);
#pragma unroll
                for (int d = 0; d < 4; d++) {
                    uint64_t t = d < 2 ? b : y;
                    if (cg_move_hori_tab(lut, &t, d == 1 || d == 2, r.ff) &&
                        !(r.prune && r.prune_moves && cg_prune(t, r.stsl, r.ltc, r.sl)))
                        res[k++] = cg_canon_fast(t, r.cs, d >= 2);
                }
            }
        }
#pragma unroll
        for (int j = 0; j < 4; j++)
            if (j < (int)k) kor |= res[j], kand &= res[j];
        uint32_t tot;
        uint32_t off = block_excl_scan<uint32_t>(k, &tot, wtmp);
        if (threadIdx.x == 0) base_sh = tot ? atomicAdd(&stats[0], (unsigned long long)tot) : 0;
        __syncthreads();
        unsigned long long o = base_sh + off;
#pragma unroll
        for (int j = 0; j < 4; j++)
            if (j < (int)k) out[o++] = res[j];
        __syncthreads();
    }
    block_or_and(kor, kand, &stats[1], &stats[2], ro, ra);
}

/* Appends the occupied slots of a block's hash set to out (one atomic on
 * stats[0] reserves the block's range) and folds them into kor / kand.  Wave
 * w reads slots (s * 16 + w) * 64 + lane in round s, so a ballot per round
 * gives each occupied slot its place in a run the wave writes contiguously.
 * Ends with a barrier: the set and wpre may be reused at once. */
__device__ __forceinline__ void hash_set_flush(const unsigned long long *ht, uint64_t *__restrict__ out,
                                               unsigned long long *stats, uint32_t *wpre,
                                               unsigned long long *base_sh, uint64_t &kor,
                                               uint64_t &kand)
{
    constexpr int NW = MV_THREADS / 64, ROUNDS = MV_HASH / MV_THREADS;
    const int w = threadIdx.x >> 6;
    uint64_t v[ROUNDS], m[ROUNDS];
    uint32_t wtot = 0;
#pragma unroll
    for (int s = 0; s < ROUNDS; s++) {
        v[s] = ht[(s * NW + w) * 64 + lane_id()];
        m[s] = __ballot(v[s] != 0ULL);
        wtot += (uint32_t)__popcll(m[s]);
    }
    if (lane_id() == 0) wpre[w] = wtot;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t acc = 0;
        for (int j = 0; j < NW; j++) {
            const uint32_t c = wpre[j];
            wpre[j] = acc;
            acc += c;
        }
        *base_sh = acc ? atomicAdd(&stats[0], (unsigned long long)acc) : 0;
    }
    __syncthreads();
    unsigned long long o = *base_sh + wpre[w];
    const uint64_t lt = lanemask_lt();
#pragma unroll
    for (int s = 0; s < ROUNDS; s++) {
        if (v[s]) {
            out[o + __popcll(m[s] & lt)] = v[s];
            kor |= v[s];
            kand &= v[s];
        }
        o += __popcll(m[s]);
    }
    __syncthreads();
}

/* The generation step's move kernel.  A block takes 2048 input boards per
 * round (two per thread) and inserts every candidate into its LDS hash set
 * (8192 slots: at most 8192 candidates, about 4,000 distinct), so each
 * candidate repeated within the 2048 boards is kept once; the set is then
 * appended with one output reservation per round (hash_set_flush). */
constexpr int MG_BPT = 2;   /* boards per thread per round */

/* PRUNE: some prune test runs here (a generation's first step, or with the
 * tests left in the move kernel); without it the kernel holds no prune code
 * and its parameters take no registers */
template <bool PRUNE>
__global__ __launch_bounds__(MV_THREADS, 8) void k_move_gen(
    const uint64_t *__restrict__ in, size_t n, const unsigned long long *dn,
    uint64_t *__restrict__ out, unsigned long long *stats, const uint16_t *__restrict__ lut,
    MoveRules r)
{
    __shared__ unsigned long long ht[MV_HASH];
    __shared__ uint32_t wpre[MV_THREADS / 64];
    __shared__ unsigned long long base_sh;
    __shared__ uint64_t ro[16], ra[16];
    uint64_t kor = 0, kand = ~0ULL;
    if (dn && *dn < n) n = (size_t)*dn;   /* count produced on the device */
    const size_t chunk = (size_t)MV_THREADS * MG_BPT;
    const size_t stride = (size_t)gridDim.x * chunk;
    for (size_t b0 = (size_t)blockIdx.x * chunk; b0 < n; b0 += stride) {
        for (int h = threadIdx.x; h < MV_HASH; h += MV_THREADS) ht[h] = 0ULL;
        __syncthreads();
#pragma unroll
        for (int k = 0; k < MG_BPT; k++) {
            const size_t i = b0 + (size_t)k * MV_THREADS + threadIdx.x;
            if (i >= n) continue;
            const uint64_t b = in[i];
            if (PRUNE && r.prune_inputs && cg_prune(b, r.stsl, r.ltc, r.sl)) continue;
            const uint64_t y = cg_rot_cw(b);
#pragma unroll
            for (int d = 0; d < 4; d++) {
                uint64_t t = d < 2 ? b : y;
                if (cg_move_hori_tab(lut, &t, d == 1 || d == 2, r.ff) &&
                    !(PRUNE && r.prune_moves && cg_prune(t, r.stsl, r.ltc, r.sl)))
                    mv_first(ht, cg_canon_fast(t, r.cs, d >= 2));
            }
        }
        __syncthreads();
        hash_set_flush(ht, out, stats, wpre, &base_sh, kor, kand);
    }
    block_or_and(kor, kand, &stats[1], &stats[2], ro, ra);
}

void launch_move(hipStream_t st, const uint64_t *in, size_t n, const unsigned long long *dn,
                 uint64_t *out, unsigned long long *stats, const uint16_t *lut, MoveRules r,
                 int dedupe)
{
    if (!n) return;
    size_t blocks = (n + MV_THREADS - 1) / MV_THREADS;
    if (blocks > 512) blocks = 512;
    if (dedupe && r.prune && (r.prune_inputs || r.prune_moves))
        hipLaunchKernelGGL(k_move_gen<true>, dim3((unsigned)blocks), dim3(MV_THREADS), 0, st, in, n, dn,
                           out, stats, lut, r);
    else if (dedupe)
        hipLaunchKernelGGL(k_move_gen<false>, dim3((unsigned)blocks), dim3(MV_THREADS), 0, st, in, n, dn,
                           out, stats, lut, r);
    else
        hipLaunchKernelGGL(k_move, dim3((unsigned)blocks), dim3(MV_THREADS), 0, st, in, n, dn,
                           out, stats, lut, r);
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

/* (board, empty cell) pairs of one 256-board chunk, in output order */
struct SpawnPairs {
    uint64_t sb[SP_THREADS];
    uint8_t pb[16 * SP_THREADS], pc[16 * SP_THREADS];
    uint8_t pf[SP_THREADS];   /* per board: bit 0 its 2-spawns fail the prune test, bit 1 its 4-spawns */
    uint32_t wtmp[17];
};

/* Each thread lists its board's empty cells (ascending) as pairs at its
 * exclusive-scan offset; then consecutive threads take consecutive pairs, so
 * the two output arrays are written with coalesced stores and every lane has
 * the same work whatever the number of empty cells.  A block first counts the
 * pairs of all its chunks and reserves its output range with one atomic (one
 * per chunk made the shared counter the bottleneck). */
__global__ __launch_bounds__(SP_THREADS) void k_spawn(const uint64_t *__restrict__ in, size_t n,
                                                      const unsigned long long *dn,
                                                      uint64_t *__restrict__ out2,
                                                      uint64_t *__restrict__ out4, size_t cap,
                                                      unsigned long long *stats, cg_canon_sel_t cs,
                                                      KeyPrune kp, int unsorted4)
{
    __shared__ SpawnPairs sp;
    __shared__ unsigned long long base_sh;
    __shared__ uint64_t ro[16], ra[16];
    uint64_t or4 = 0, and4 = ~0ULL;
    if (dn && *dn < n) n = (size_t)*dn;   /* count produced on the device */
    const size_t stride = (size_t)gridDim.x * SP_THREADS;
    const size_t first = (size_t)blockIdx.x * SP_THREADS + threadIdx.x;
    uint32_t mine = 0;
    for (size_t i = first; i < n; i += stride) mine += (uint32_t)cg_empties(in[i]);
    uint32_t all;
    block_excl_scan<uint32_t>(mine, &all, sp.wtmp);
    if (threadIdx.x == 0) base_sh = all ? atomicAdd(&stats[0], (unsigned long long)all) : 0;
    __syncthreads();
    size_t base = base_sh;
    /* the next chunk's board is loaded while this chunk is processed */
    size_t i = first;
    uint64_t nb = i < n ? in[i] : 0;
    for (size_t b0 = (size_t)blockIdx.x * SP_THREADS; b0 < n; b0 += stride, i += stride) {
        const uint64_t b = nb;
        nb = i + stride < n ? in[i + stride] : 0;
        uint64_t z = i < n ? cg_zmask(b) : 0;
        uint32_t e = (uint32_t)cg_popc64(z);
        if (kp.on) {
            /* a spawn fills one empty cell with a 2 (or a 4): the same
             * multiset whichever cell, so one test per board and value (at
             * the lowest empty cell) */
            const uint64_t lb = z & (0ULL - z);
            uint32_t f = 0;
            if (z && cg_prune(b | lb, kp.stsl, kp.ltc, kp.sl)) f |= 1u;
            if (z && unsorted4 && cg_prune(b | (lb << 1), kp.stsl, kp.ltc, kp.sl)) f |= 2u;
            sp.pf[threadIdx.x] = (uint8_t)f;
        }
        uint32_t tot;
        uint32_t off = block_excl_scan<uint32_t>(e, &tot, sp.wtmp);
        sp.sb[threadIdx.x] = b;
        while (z) {
            sp.pb[off] = (uint8_t)threadIdx.x;
            sp.pc[off] = (uint8_t)next_cell(z);
            off++;
        }
        __syncthreads();
        for (uint32_t p = threadIdx.x; p < tot; p += SP_THREADS) {
            uint64_t bb = sp.sb[sp.pb[p]];
            int x = sp.pc[p];
            uint64_t c2 = cg_canon_fast(cg_set_tile(bb, x, 1), cs, 0),
                     c4 = cg_canon_fast(cg_set_tile(bb, x, 2), cs, 0);
            if (kp.on) {
                const uint32_t f = sp.pf[sp.pb[p]];
                c2 = (f & 1u) ? 0ULL : c2;
                c4 = (f & 2u) ? 0ULL : c4;
            }
            size_t o = base + p;
            if (o < cap) {   /* past the capacity: only counted (the host retries) */
                out2[o] = c2;
                out4[o] = c4;
            }
            or4 |= c4;
            and4 &= c4;
        }
        base += tot;
        __syncthreads();
    }
    /* the OR/AND of the 4-spawns: the key bits of their sort */
    if (!unsorted4) block_or_and(or4, and4, &stats[3], &stats[4], ro, ra);
}

/* with no OR/AND atomics, more blocks: C2 spawn 25.2 ms per table with 512
 * blocks, 19.8 with 1024, 20.3 with 2048, 21.8 with 4096 */
static int g_spawn_blocks = 1024;
void spawn_set_blocks(int b) { g_spawn_blocks = b > 0 ? b : 1024; }
int spawn_blocks() { return g_spawn_blocks; }

void launch_spawn(hipStream_t st, const uint64_t *in, size_t n, const unsigned long long *dn,
                  uint64_t *out2, uint64_t *out4, size_t cap, unsigned long long *stats, cg_canon_sel_t cs,
                  KeyPrune kp, int unsorted4)
{
    if (!n) return;
    size_t blocks = (n + SP_THREADS - 1) / SP_THREADS;
    /* 512 blocks: the OR/AND atomics at the end of every block bound the
     * kernel (with two OR/AND pairs per block, C2 spawn took 22.4 ms per step
     * with 512 blocks, 25.1 with 1024, 38.7 with 2048, 29.7 with 256; with
     * the one pair now, 19.1 with 512 and 18.8 with 1024) */
    const size_t cap_blocks = unsorted4 ? (size_t)g_spawn_blocks : 512;
    if (blocks > cap_blocks) blocks = cap_blocks;
    hipLaunchKernelGGL(k_spawn, dim3((unsigned)blocks), dim3(SP_THREADS), 0, st, in, n, dn, out2,
                       out4, cap, stats, cs, kp, unsorted4);
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

__global__ void k_copy_and_u64(uint64_t *__restrict__ dst, const uint64_t *__restrict__ src,
                               size_t n, uint64_t mask)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i] & mask;
}

void launch_copy_and_u64(hipStream_t st, uint64_t *dst, const uint64_t *src, size_t n, uint64_t mask)
{
    if (!n) return;
    size_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_copy_and_u64, dim3((unsigned)blocks), dim3(256), 0, st, dst, src, n, mask);
}

/* per-table statistics of a batched layer, [3t] count, [3t+1] OR, [3t+2]
 * AND: 0, 0, all ones */
__global__ void k_tstat_init(unsigned long long *t, int ntab)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < 3 * ntab; i += gridDim.x * blockDim.x)
        t[i] = i % 3 == 2 ? ~0ULL : 0ULL;
}

void launch_tstat_init(hipStream_t st, unsigned long long *t, int ntab)
{
    if (ntab <= 0) return;
    const int blocks = (3 * ntab + 255) / 256;
    hipLaunchKernelGGL(k_tstat_init, dim3((unsigned)(blocks > 64 ? 64 : blocks)), dim3(256), 0, st, t, ntab);
}

__global__ void k_publish(const unsigned long long *cnt, int lo, int n, const unsigned *err,
                          unsigned long long *hpub, int nerr_at)
{
    int i = threadIdx.x;
    if (i < n) hpub[lo + i] = cnt[lo + i];
    if (i >= 62) hpub[nerr_at + (i - 62)] = err[i - 62];
}

void launch_publish(hipStream_t st, const unsigned long long *cnt, int lo, int n, const unsigned *err,
                    unsigned long long *hpub, int err_at)
{
    hipLaunchKernelGGL(k_publish, dim3(1), dim3(64), 0, st, cnt, lo, n, err, hpub, err_at);
}

/* producer statistics: [0] count, then (OR, AND) pairs; AND starts all-ones */
__global__ void k_init_stats(unsigned long long *s, unsigned long long *s2)
{
    /* slots 5..7 hold sort counts that later launches read on the device */
    if (threadIdx.x < 5) s[threadIdx.x] = (threadIdx.x == 2 || threadIdx.x == 4) ? ~0ULL : 0ULL;
    if (s2 && threadIdx.x >= 32 && threadIdx.x < 37) {
        const int i = threadIdx.x - 32;
        s2[i] = (i == 2 || i == 4) ? ~0ULL : 0ULL;
    }
}

void launch_init_stats(hipStream_t st, unsigned long long *stats, unsigned long long *stats2)
{
    hipLaunchKernelGGL(k_init_stats, dim3(1), dim3(64), 0, st, stats, stats2);
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
/* prefix-bucket directory + lookup                                          */
/* ------------------------------------------------------------------------ */

static int g_dir_delta = 0;
void set_dir_delta(int d) { g_dir_delta = d < -4 ? -4 : (d > 3 ? 3 : d); }
int get_dir_delta() { return g_dir_delta; }

LayerIndexPlan plan_index(size_t n, uint64_t key_or, uint64_t key_and)
{
    LayerIndexPlan p;
    uint64_t vary = n > 1 ? (key_or & ~key_and) : 0;
    p.x = make_xform(vary);
    p.cmask = ~vary;
    p.cval = key_and & p.cmask;
    int lg = 0;
    while (((size_t)1 << (lg + 1)) <= n) lg++;
    int bits = lg + g_dir_delta;
    if (bits < 0) bits = 0;
    if (bits > p.x.nbits) bits = p.x.nbits;
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

/* Exclusive scan of the bucket counts into the directory: tiles of 16384
 * counts, 16 consecutive per thread (four 16-byte loads and stores);
 * k_dscan_reduce sums each tile, and each block of k_dscan_down adds up the
 * tile sums before its own (at most DS_FOLD tiles; beyond that k_dscan_top
 * scans them first) -- two launches for a directory of up to 2^27 buckets. */
constexpr int DS_THREADS = 1024, DS_IPT = 16, DS_TILE = DS_THREADS * DS_IPT;
constexpr size_t DS_FOLD = 8192;

__device__ __forceinline__ void ds_load(const uint32_t *__restrict__ in, size_t base, size_t m,
                                        uint32_t (&v)[DS_IPT])
{
    if (base + DS_IPT <= m) {
        const uint4 *q = reinterpret_cast<const uint4 *>(in + base);
#pragma unroll
        for (int j = 0; j < DS_IPT / 4; j++) {
            const uint4 x = q[j];
            v[4 * j] = x.x;
            v[4 * j + 1] = x.y;
            v[4 * j + 2] = x.z;
            v[4 * j + 3] = x.w;
        }
    } else {
#pragma unroll
        for (int j = 0; j < DS_IPT; j++) v[j] = base + j < m ? in[base + j] : 0u;
    }
}

__global__ __launch_bounds__(DS_THREADS) void k_dscan_reduce(const uint32_t *__restrict__ in, size_t m,
                                                             uint32_t *sums)
{
    __shared__ uint32_t wtmp[17];
    uint32_t v[DS_IPT];
    ds_load(in, (size_t)blockIdx.x * DS_TILE + (size_t)threadIdx.x * DS_IPT, m, v);
    uint32_t s = 0;
#pragma unroll
    for (int j = 0; j < DS_IPT; j++) s += v[j];
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

/* exclusive scan of cnt[0..m) into dir[0..m) plus dir[m] = total; sums are
 * the tile sums (nb <= DS_FOLD) or their exclusive scan (nb > DS_FOLD) */
__global__ __launch_bounds__(DS_THREADS) void k_dscan_down(const uint32_t *__restrict__ in, size_t m,
                                                           const uint32_t *sums, size_t nb, uint32_t *out,
                                                           uint32_t total)
{
    __shared__ uint32_t wtmp[17], wtmp2[17];
    uint32_t pre;
    if (nb <= DS_FOLD) {
        uint32_t p = 0;
        for (size_t t = threadIdx.x; t < blockIdx.x; t += DS_THREADS) p += sums[t];
        block_excl_scan<uint32_t>(p, &pre, wtmp2);
    } else {
        pre = sums[blockIdx.x];
    }
    const size_t base = (size_t)blockIdx.x * DS_TILE + (size_t)threadIdx.x * DS_IPT;
    uint32_t v[DS_IPT];
    ds_load(in, base, m, v);
    uint32_t s = 0;
#pragma unroll
    for (int j = 0; j < DS_IPT; j++) s += v[j];
    uint32_t tot;
    uint32_t ex = block_excl_scan<uint32_t>(s, &tot, wtmp) + pre;
#pragma unroll
    for (int j = 0; j < DS_IPT; j++) {
        const uint32_t x = v[j];
        v[j] = ex;
        ex += x;
    }
    if (base + DS_IPT <= m) {
        uint4 *q = reinterpret_cast<uint4 *>(out + base);
#pragma unroll
        for (int j = 0; j < DS_IPT / 4; j++) q[j] = make_uint4(v[4 * j], v[4 * j + 1], v[4 * j + 2], v[4 * j + 3]);
    } else {
#pragma unroll
        for (int j = 0; j < DS_IPT; j++)
            if (base + j < m) out[base + j] = v[j];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) out[m] = total;
}

/* dir[0..m] from the counts cnt[0..m) (sums: scratch for the tile sums) */
static void launch_dscan(hipStream_t st, const uint32_t *cnt, size_t m, uint32_t *sums, uint32_t *dir,
                         uint32_t total)
{
    const size_t nb = (m + DS_TILE - 1) / DS_TILE;
    if (nb) hipLaunchKernelGGL(k_dscan_reduce, dim3((unsigned)nb), dim3(DS_THREADS), 0, st, cnt, m, sums);
    if (nb > DS_FOLD) hipLaunchKernelGGL(k_dscan_top, dim3(1), dim3(DS_THREADS), 0, st, sums, nb);
    /* m == 0: one block writes dir[0] = total */
    hipLaunchKernelGGL(k_dscan_down, dim3((unsigned)(nb ? nb : 1)), dim3(DS_THREADS), 0, st, cnt, m,
                       (const uint32_t *)sums, nb, dir, total);
}

size_t dir_scratch_bytes(size_t entries)
{
    return entries * 4 + ((entries + DS_TILE - 1) / DS_TILE + 1) * 4 + 1024;
}

void launch_build_dir(hipStream_t st, const uint64_t *keys, size_t n, const LayerIndexPlan &p,
                      uint32_t *dir, void *scratch)
{
    if (n >= LAYER_INDEX_MAX) {   /* 32-bit entries would wrap */
        kernels_note_error(hipErrorInvalidValue);
        return;
    }
    size_t m = p.dir_entries - 1;   /* buckets */
    uint32_t *cnt = (uint32_t *)scratch;
    uint32_t *sums = cnt + ((m + 2 + 63) / 64) * 64;
    launch_fill_u64(st, (uint64_t *)cnt, (m + 1) / 2 + 1, 0);
    if (n) {
        size_t blocks = (n + 255) / 256;
        hipLaunchKernelGGL(k_dir_count, dim3((unsigned)blocks), dim3(256), 0, st, keys, n, p.x,
                           p.shift, cnt);
    }
    launch_dscan(st, cnt, m, sums, dir, (uint32_t)n);
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

/* two zero words: the directory range [0, 0), key 0 and value 0.0 */
__device__ const uint64_t g_zero2[2] = { 0, 0 };

/* Lockstep lookups of N keys for the adjoint: slot s looks for q[s] in t2
 * (even s) or t4 (odd s), nothing when live[s] is false, as layer_find does
 * (constant bits checked by the caller, bucket range from the directory,
 * lower bound), with the loads of all N slots in flight in every round and
 * a zero word read where a slot has nothing to read -- the scheme of
 * best_successors.  Returns found[s] and the key's index pos[s]. */
template <int N>
__device__ __forceinline__ void find_lockstep(const uint64_t (&q)[N], const bool (&live)[N],
                                              const LayerView &t2, const LayerView &t4,
                                              uint32_t (&pos)[N], bool (&found)[N])
{
#define CG_SLOT_VIEW(s) (((s) & 1) ? t4 : t2)
    uint32_t hi[N];
#pragma unroll
    for (int s = 0; s < N; s++) {
        const LayerView &t = CG_SLOT_VIEW(s);
        const uint32_t *p = reinterpret_cast<const uint32_t *>(g_zero2);
        if (live[s]) p = t.dir + (xform(q[s], t.x) >> t.shift);
        pos[s] = p[0];
        hi[s] = p[1];
        found[s] = false;
    }
    for (;;) {
        uint64_t k[N];
        bool any = false;
#pragma unroll
        for (int s = 0; s < N; s++) {
            const LayerView &t = CG_SLOT_VIEW(s);
            k[s] = *(pos[s] < hi[s] ? t.keys + (pos[s] + ((hi[s] - pos[s]) >> 1)) : g_zero2);
        }
#pragma unroll
        for (int s = 0; s < N; s++) {
            if (pos[s] < hi[s]) {
                const uint32_t mid = pos[s] + ((hi[s] - pos[s]) >> 1);
                if (k[s] < q[s]) {
                    pos[s] = mid + 1;
                } else {
                    hi[s] = mid;
                    found[s] = k[s] == q[s];
                }
                any |= pos[s] < hi[s];
            }
        }
        if (!any) break;
    }
#undef CG_SLOT_VIEW
}

/* The eight successors of one (board, cell) pair -- the spawned 2 (c2, layer
 * t2) and the spawned 4 (c4, layer t4), four moves each -- are looked up in
 * lockstep, so each round (the directory, every step of the lower-bound
 * search, the value) has all eight loads in flight instead of one.  Each
 * lookup is layer_find's: constant bits, bucket range, lower bound; the
 * search remembers whether the last key at or above q equals q.  For each
 * spawn the result is the max over the 4 moves of the successor value and
 * the first strict maximum's direction (0 when every value is 0). */
__device__ __forceinline__ void best_successors(uint64_t c2, uint64_t c4, const LayerView &t2,
                                                const LayerView &t4, const uint16_t *lut, int ff,
                                                double &m2, double &m4, int &g2, int &g4,
                                                uint32_t &nlook, uint32_t &nhit)
{
    const uint64_t y2 = cg_rot_cw(c2), y4 = cg_rot_cw(c4);
    uint64_t q[8];
    uint32_t lo[8], hi[8];
    /* per-slot flags as bits of one word per lane (bit s: slot s) rather
     * than eight lane masks each */
    uint32_t live = 0, f = 0;
#pragma unroll
    for (int s = 0; s < 8; s++) {
        const int d = s & 3;
        const LayerView &t = s < 4 ? t2 : t4;
        uint64_t m = s < 4 ? (d < 2 ? c2 : y2) : (d < 2 ? c4 : y4);
        const bool ok = cg_move_hori_tab(lut, &m, d == 1 || d == 2, ff);
        nlook += ok;
        q[s] = cg_canon(m);
        live |= (uint32_t)(ok && t.n != 0 && (q[s] & t.cmask) == t.cval) << s;
        lo[s] = hi[s] = 0;
    }
    /* the loads are unconditional, from a zero word where a slot has nothing
     * to read, so all eight are issued before the first wait */
#pragma unroll
    for (int s = 0; s < 8; s++) {
        const LayerView &t = s < 4 ? t2 : t4;
        const uint32_t *p = reinterpret_cast<const uint32_t *>(g_zero2);
        if ((live >> s) & 1u) p = t.dir + (xform(q[s], t.x) >> t.shift);
        lo[s] = p[0];
        hi[s] = p[1];
    }
    for (;;) {
        uint64_t k[8];
        bool any = false;
#pragma unroll
        for (int s = 0; s < 8; s++) {
            const LayerView &t = s < 4 ? t2 : t4;
            k[s] = *(lo[s] < hi[s] ? t.keys + (lo[s] + ((hi[s] - lo[s]) >> 1)) : g_zero2);
        }
#pragma unroll
        for (int s = 0; s < 8; s++) {
            if (lo[s] < hi[s]) {
                const uint32_t mid = lo[s] + ((hi[s] - lo[s]) >> 1);
                if (k[s] < q[s]) {
                    lo[s] = mid + 1;
                } else {
                    hi[s] = mid;
                    f = (f & ~(1u << s)) | ((uint32_t)(k[s] == q[s]) << s);
                }
                any |= lo[s] < hi[s];
            }
        }
        if (!any) break;
    }
    double v[8];
#pragma unroll
    for (int s = 0; s < 8; s++) {
        const LayerView &t = s < 4 ? t2 : t4;
        const uint32_t fs = (f >> s) & 1u;
        v[s] = *(fs ? t.vals + lo[s] : reinterpret_cast<const double *>(g_zero2));
        nhit += fs;
    }
    double b2 = 0.0, b4 = 0.0;
    int d2 = 0, d4 = 0;
#pragma unroll
    for (int d = 0; d < 4; d++) {
        if (v[d] > b2) {
            b2 = v[d];
            d2 = d;
        }
        if (v[4 + d] > b4) {
            b4 = v[4 + d];
            d4 = d;
        }
    }
    m2 = b2;
    m4 = b4;
    g2 = d2;
    g4 = d4;
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
            best_successors(cg_set_tile(pbv, x, 1), cg_set_tile(pbv, x, 2), t2, t4, lut, ff,
                            sm.m2s[threadIdx.x], sm.m4s[threadIdx.x], g2, g4, nlook, nhit);
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

/* the block's lookup counts, one atomic pair per block (stats is null
 * unless the solve is being profiled) */
__device__ __forceinline__ void add_stats(unsigned long long *stats, uint32_t nlook, uint32_t nhit,
                                          unsigned long long *red)
{
    if (!stats) return;
    unsigned long long l = nlook, h = nhit;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        l += __shfl_xor(l, d, 64);
        h += __shfl_xor(h, d, 64);
    }
    if (lane_id() == 0) red[2 * (threadIdx.x >> 6)] = l, red[2 * (threadIdx.x >> 6) + 1] = h;
    __syncthreads();
    if (threadIdx.x == 0) {
        l = h = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) l += red[2 * w], h += red[2 * w + 1];
        if (l) {
            atomicAdd(&stats[0], l);
            atomicAdd(&stats[1], h);
        }
    }
}

/* MW: the waves per SIMD the registers are fitted to (0: the compiler's
 * choice, 4 waves); ST: the lookup counts are summed into stats (the
 * counters are dead code, and free their registers, without it) */
template <int MW, bool ST>
__global__ __launch_bounds__(SV_THREADS, MW > 0 ? MW : 1) void k_solve(const uint64_t *__restrict__ M, size_t n,
                                                      double *__restrict__ V,
                                                      uint64_t *__restrict__ argmax, LayerView t2,
                                                      LayerView t4, WinRules win, double p4,
                                                      const uint16_t *__restrict__ lut, int ff,
                                                      unsigned long long *stats)
{
    __shared__ SolveSmem sm;
    __shared__ unsigned long long red[2 * (SV_THREADS / 64)];
    uint32_t nlook = 0, nhit = 0;
    /* the grid sweeps the sorted layer front to back, so the lookups in
     * flight at any time come from one narrow window of boards */
    for (size_t tile = blockIdx.x; tile * SV_THREADS < n; tile += gridDim.x)
        solve_tile(M, n, tile, V, argmax, t2, t4, win, p4, lut, ff, nlook, nhit, sm);
    if constexpr (ST) add_stats(stats, nlook, nhit, red);
}

/* C2 solve with the per-slot flags as bit masks and without the lookup
 * counters: 71.4 ms fitted to 6 waves per SIMD (80 VGPRs, 10 spilled), 73.6
 * at 5 (96 VGPRs, none spilled), 72.7 at the compiler's 4 waves; with the
 * flags as lane masks and the counters, 75.7 / 73.6 / 83.5 ms at 4 / 5 / 6 */
static int g_solve_waves = 6;

void launch_solve(hipStream_t st, const uint64_t *M, size_t n, double *V, uint64_t *argmax,
                  LayerView t2, LayerView t4, WinRules w, double p4, const uint16_t *lut, int ff,
                  unsigned long long *stats)
{
    if (!n) return;
    /* 8192 blocks (32 per CU, 4 resident): on C2 the solve took 86.5 ms
     * with 2048, 79.3 with 4096, 75.9 with 8192 and 77.3 with 16384 or more */
    size_t tiles = (n + SV_THREADS - 1) / SV_THREADS;
    size_t blocks = tiles < 8192 ? tiles : 8192;
#define CG_SOLVE(MW)                                                                                      \
    do {                                                                                                  \
        if (stats)                                                                                        \
            hipLaunchKernelGGL((k_solve<MW, true>), dim3((unsigned)blocks), dim3(SV_THREADS), 0, st, M, n, V, \
                               argmax, t2, t4, w, p4, lut, ff, stats);                                    \
        else                                                                                              \
            hipLaunchKernelGGL((k_solve<MW, false>), dim3((unsigned)blocks), dim3(SV_THREADS), 0, st, M, n,   \
                               V, argmax, t2, t4, w, p4, lut, ff, stats);                                 \
    } while (0)
    if (g_solve_waves == 5) CG_SOLVE(5);
    else if (g_solve_waves == 6) CG_SOLVE(6);
    else CG_SOLVE(0);
#undef CG_SOLVE
}

void solve_set_waves(int w) { g_solve_waves = (w == 5 || w == 6) ? w : 0; }
int solve_waves() { return g_solve_waves; }

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
        /* two cells per round: slot s = 2 * (cell of the round) + (spawn 4?),
         * the four lookups in lockstep; the sums still run over the cells in
         * ascending order */
        while (z) {
            int xs[2];
            xs[0] = next_cell(z);
            const bool two = z != 0;
            xs[1] = two ? next_cell(z) : xs[0];
            uint64_t q[4];
            bool live[4];
#pragma unroll
            for (int s = 0; s < 4; s++) {
                const int k = s & 1, x = xs[s >> 1];
                const LayerView &t = k ? t4 : t2;
                const int d = (int)(am >> (4 * x + 2 * k)) & 3;
                uint64_t m = cg_set_tile(b, x, (unsigned)(k + 1));
                if (d >= 2) m = cg_rot_cw(m);
                const bool ok = cg_move_hori_tab(lut, &m, d == 1 || d == 2, ff) && (s < 2 || two);
                q[s] = cg_canon(m);
                live[s] = ok && t.n != 0 && (q[s] & t.cmask) == t.cval;
            }
            uint32_t pos[4];
            bool f[4];
            find_lockstep<4>(q, live, t2, t4, pos, f);
            double v[4];
#pragma unroll
            for (int s = 0; s < 4; s++) {
                const LayerView &t = (s & 1) ? t4 : t2;
                v[s] = *(f[s] ? t.vals + pos[s] : reinterpret_cast<const double *>(g_zero2));
            }
#pragma unroll
            for (int s = 0; s < 4; s++) {
                if (!f[s]) continue;
                const int k = s & 1;
                if (k) a4 += v[s];
                else a2 += v[s];
                if (v[s] > 0.0)
                    atomicAdd(k ? &occ4[pos[s]] : &occ2[pos[s]], o * (k ? p4 : w2) / (double)E);
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
                               const uint16_t *lut, int ff, double *P, int premove)
{
    for (size_t i = threadIdx.x; i < n; i += blockDim.x) {
        uint64_t b = init[i];
        double best = 0.0;
        long bj = -1;
        if (premove) {   /* the initial boards are layer 0's positions */
            bj = layer_find(cg_canon(b), t0);
            best = bj >= 0 ? t0.vals[bj] : 0.0;
        }
        for (int d = 0; d < 4 && !premove; d++) {
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
                         const uint16_t *lut, int ff, double *P, int premove)
{
    if (!n) return;
    hipLaunchKernelGGL(k_adjoint_init, dim3(1), dim3(256), 0, st, init, n, t0, occ0, lut, ff, P, premove);
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
    if (T.runs) {
        c = xform(b, T.fx);
    } else {
        for (int j = 0; j < T.nfree; j++) c = (c << 4) | ((b >> (60 - 4 * T.pos[j])) & 0xf);
    }
    return (tid << BATCH_TID_SHIFT) | c;
}

__device__ __forceinline__ uint64_t bdec(uint64_t ck, const BatchTable &T)
{
    uint64_t c = ck & ((1ULL << BATCH_TID_SHIFT) - 1);
    if (T.runs) return T.fnib | xform_inv(c, T.fx);
    uint64_t b = T.fnib;
    for (int j = T.nfree - 1; j >= 0; j--) {
        b |= (c & 0xf) << (60 - 4 * T.pos[j]);
        c >>= 4;
    }
    return b;
}

/* the batched move: decode, move, canonicalise, re-encode; as k_move_gen,
 * 2048 keys per round go through one LDS hash set, so duplicates among them
 * are appended once */
__global__ __launch_bounds__(MV_THREADS, 8) void k_bmove(const uint64_t *__restrict__ in, size_t n,
                                                         const unsigned long long *dn,
                                                         uint64_t *__restrict__ out,
                                                         unsigned long long *stats,
                                                         const uint16_t *__restrict__ lut, MoveRules r,
                                                         const BatchTable *__restrict__ tabs, int step,
                                                         unsigned *err)
{
    __shared__ unsigned long long ht[MV_HASH];
    __shared__ uint32_t wpre[MV_THREADS / 64];
    __shared__ unsigned long long base_sh;
    __shared__ uint64_t ro[16], ra[16];
    uint64_t kor = 0, kand = ~0ULL;
    if (dn && *dn < n) n = (size_t)*dn;
    const size_t chunk = (size_t)MV_THREADS * MG_BPT;
    const size_t stride = (size_t)gridDim.x * chunk;
    for (size_t b0 = (size_t)blockIdx.x * chunk; b0 < n; b0 += stride) {
        for (int h = threadIdx.x; h < MV_HASH; h += MV_THREADS) ht[h] = 0ULL;
        __syncthreads();
#pragma unroll
        for (int k = 0; k < MG_BPT; k++) {
            const size_t i = b0 + (size_t)k * MV_THREADS + threadIdx.x;
            if (i >= n) continue;
            const uint64_t ck = in[i];
            const uint64_t tid = ck >> BATCH_TID_SHIFT;
            const BatchTable &T = tabs[tid];
            if (step >= T.nsteps) continue;
            const uint64_t b = bdec(ck, T);
            if (r.prune && r.prune_inputs && cg_prune(b, r.stsl, r.ltc, r.sl)) continue;
            const uint64_t y = cg_rot_cw(b);
#pragma unroll
            for (int d = 0; d < 4; d++) {
                uint64_t t = d < 2 ? b : y;
                if (cg_move_hori_tab(lut, &t, d == 1 || d == 2, r.ff) &&
                    !(r.prune && r.prune_moves && cg_prune(t, r.stsl, r.ltc, r.sl)))
                    mv_first(ht, benc(cg_canon(t), T, tid, err));
            }
        }
        __syncthreads();
        hash_set_flush(ht, out, stats, wpre, &base_sh, kor, kand);
    }
    block_or_and(kor, kand, &stats[1], &stats[2], ro, ra);
}

__global__ __launch_bounds__(SP_THREADS) void k_bspawn(const uint64_t *__restrict__ in, size_t n,
                                                       const unsigned long long *dn,
                                                       uint64_t *__restrict__ out2,
                                                       uint64_t *__restrict__ out4, size_t cap,
                                                       unsigned long long *stats,
                                                       const BatchTable *__restrict__ tabs, int step,
                                                       unsigned *err)
{
    __shared__ SpawnPairs sp;
    __shared__ uint16_t tid_sh[SP_THREADS];   /* table ids < 2^(64 - BATCH_TID_SHIFT) */
    __shared__ unsigned long long base_sh;
    __shared__ uint64_t ro[16], ra[16];
    uint64_t or2 = 0, and2 = ~0ULL, or4 = 0, and4 = ~0ULL;
    if (dn && *dn < n) n = (size_t)*dn;
    const size_t stride = (size_t)gridDim.x * SP_THREADS;
    /* one output reservation per block, as in k_spawn; the spawns of a
     * board feed step + 1 of its table */
    uint32_t mine = 0;
    for (size_t i = (size_t)blockIdx.x * SP_THREADS + threadIdx.x; i < n; i += stride) {
        const uint64_t ck = in[i];
        const uint64_t t = ck >> BATCH_TID_SHIFT;
        if (step + 1 < tabs[t].nsteps) mine += (uint32_t)cg_empties(bdec(ck, tabs[t]));
    }
    uint32_t all;
    block_excl_scan<uint32_t>(mine, &all, sp.wtmp);
    if (threadIdx.x == 0) base_sh = all ? atomicAdd(&stats[0], (unsigned long long)all) : 0;
    __syncthreads();
    size_t base = base_sh;
    for (size_t b0 = (size_t)blockIdx.x * SP_THREADS; b0 < n; b0 += stride) {
        size_t i = b0 + threadIdx.x;
        uint64_t b = 0, z = 0, tid = 0;
        if (i < n) {
            uint64_t ck = in[i];
            tid = ck >> BATCH_TID_SHIFT;
            if (step + 1 < tabs[tid].nsteps) {
                b = bdec(ck, tabs[tid]);
                z = cg_zmask(b);
            }
        }
        uint32_t e = (uint32_t)cg_popc64(z);
        uint32_t tot;
        uint32_t off = block_excl_scan<uint32_t>(e, &tot, sp.wtmp);
        sp.sb[threadIdx.x] = b;
        tid_sh[threadIdx.x] = (uint16_t)tid;
        while (z) {
            sp.pb[off] = (uint8_t)threadIdx.x;
            sp.pc[off] = (uint8_t)next_cell(z);
            off++;
        }
        __syncthreads();
        for (uint32_t p = threadIdx.x; p < tot; p += SP_THREADS) {
            int q = sp.pb[p];
            uint64_t bb = sp.sb[q];
            uint64_t t = tid_sh[q];
            const BatchTable &T = tabs[t];
            int x = sp.pc[p];
            uint64_t c2 = benc(cg_canon(cg_set_tile(bb, x, 1)), T, t, err);
            uint64_t c4 = benc(cg_canon(cg_set_tile(bb, x, 2)), T, t, err);
            size_t o = base + p;
            if (o < cap) {
                out2[o] = c2;
                out4[o] = c4;
            }
            or2 |= c2;
            and2 &= c2;
            or4 |= c4;
            and4 &= c4;
        }
        base += tot;
        __syncthreads();
    }
    block_or_and(or2, and2, &stats[1], &stats[2], ro, ra);
    block_or_and(or4, and4, &stats[3], &stats[4], ro, ra);
}

/* composite -> boards; per-table count, OR and AND of the boards
 * (tstat[3 t + 0..2]).  The keys are sorted, so tables are contiguous: each
 * wave walks its own contiguous chunk of 64-key rows with a running (table,
 * count, OR, AND) accumulator and adds it to the table's counters only when
 * the table changes and at the end (a few thousand atomics per launch rather
 * than three per row on the same addresses); a row that straddles a table
 * boundary adds per key. */
__global__ __launch_bounds__(256) void k_bdecode(const uint64_t *__restrict__ in, size_t n,
                                                 const unsigned long long *dn,
                                                 uint64_t *__restrict__ out,
                                                 const BatchTable *__restrict__ tabs,
                                                 unsigned long long *tstat)
{
    const int lane = lane_id();
    if (dn && *dn < n) n = (size_t)*dn;
    const size_t rows = (n + 63) / 64;
    const size_t nwaves = (size_t)gridDim.x * 4;
    const size_t wv = ((size_t)blockIdx.x * 256 + threadIdx.x) >> 6;
    const size_t per = (rows + nwaves - 1) / nwaves;
    const size_t r0 = wv * per, r1 = r0 + per < rows ? r0 + per : rows;
    uint64_t cur = ~0ULL, cor = 0, cand = ~0ULL;   /* wave-uniform accumulator */
    unsigned long long ccnt = 0;
    for (size_t r = r0; r < r1; r++) {
        size_t i = r * 64 + lane;
        bool valid = i < n;
        uint64_t tid = ~0ULL, b = 0;
        if (valid) {
            uint64_t ck = in[i];
            tid = ck >> BATCH_TID_SHIFT;
            b = bdec(ck, tabs[tid]);
            out[i] = b;
        }
        uint64_t vm = __ballot(valid);
        uint64_t t0 = __shfl(tid, 0, 64);
        if (__ballot(valid && tid == t0) == vm) {
            uint64_t o = valid ? b : 0, a = valid ? b : ~0ULL;
#pragma unroll
            for (int d = 32; d >= 1; d >>= 1) {
                o |= __shfl_xor(o, d, 64);
                a &= __shfl_xor(a, d, 64);
            }
            if (t0 != cur) {
                if (cur != ~0ULL && lane == 0) {
                    atomicAdd(&tstat[3 * cur], ccnt);
                    atomicOr(&tstat[3 * cur + 1], (unsigned long long)cor);
                    atomicAnd(&tstat[3 * cur + 2], (unsigned long long)cand);
                }
                cur = t0;
                ccnt = 0;
                cor = 0;
                cand = ~0ULL;
            }
            ccnt += (unsigned long long)__popcll(vm);
            cor |= o;
            cand &= a;
        } else if (valid) {
            atomicAdd(&tstat[3 * tid], 1ULL);
            atomicOr(&tstat[3 * tid + 1], (unsigned long long)b);
            atomicAnd(&tstat[3 * tid + 2], (unsigned long long)b);
        }
    }
    if (cur != ~0ULL && lane == 0) {
        atomicAdd(&tstat[3 * cur], ccnt);
        atomicOr(&tstat[3 * cur + 1], (unsigned long long)cor);
        atomicAnd(&tstat[3 * cur + 2], (unsigned long long)cand);
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

void launch_bmove(hipStream_t st, const uint64_t *in, size_t n, const unsigned long long *dn,
                  uint64_t *out, unsigned long long *stats, const uint16_t *lut, MoveRules r,
                  const BatchTable *tabs, int step, unsigned *err)
{
    if (!n) return;
    size_t blocks = (n + MV_THREADS - 1) / MV_THREADS;
    if (blocks > 512) blocks = 512;
    hipLaunchKernelGGL(k_bmove, dim3((unsigned)blocks), dim3(MV_THREADS), 0, st, in, n, dn, out,
                       stats, lut, r, tabs, step, err);
}

void launch_bspawn(hipStream_t st, const uint64_t *in, size_t n, const unsigned long long *dn,
                   uint64_t *out2, uint64_t *out4, size_t cap, unsigned long long *stats,
                   const BatchTable *tabs, int step, unsigned *err)
{
    if (!n) return;
    size_t blocks = (n + SP_THREADS - 1) / SP_THREADS;
    if (blocks > 1024) blocks = 1024;
    hipLaunchKernelGGL(k_bspawn, dim3((unsigned)blocks), dim3(SP_THREADS), 0, st, in, n, dn, out2,
                       out4, cap, stats, tabs, step, err);
}

void launch_bdecode(hipStream_t st, const uint64_t *in, size_t n, const unsigned long long *dn,
                    uint64_t *out, const BatchTable *tabs, unsigned long long *tstat)
{
    if (!n) return;
    size_t blocks = (n + 255) / 256;
    if (blocks > 1024) blocks = 1024;
    hipLaunchKernelGGL(k_bdecode, dim3((unsigned)blocks), dim3(256), 0, st, in, n, dn, out, tabs,
                       tstat);
}

void launch_bdir(hipStream_t st, const uint64_t *keys, size_t n, const size_t *seg, int ntab,
                 const BatchIndexPlan *plan, size_t total_entries, uint32_t *dir, void *scratch)
{
    if (n >= LAYER_INDEX_MAX) {   /* 32-bit entries would wrap */
        kernels_note_error(hipErrorInvalidValue);
        return;
    }
    size_t m = total_entries;
    uint32_t *cnt = (uint32_t *)scratch;
    uint32_t *sums = cnt + ((m + 2 + 63) / 64) * 64;
    launch_fill_u64(st, (uint64_t *)cnt, (m + 1) / 2 + 1, 0);
    if (n)
        hipLaunchKernelGGL(k_bdir_count, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, keys, n,
                           seg, ntab, plan, cnt);
    launch_dscan(st, cnt, m, sums, dir, (uint32_t)n);
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

