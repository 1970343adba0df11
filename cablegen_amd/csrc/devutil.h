/*
 * devutil.h -- wave64 / block helpers and the key-bit packing shared by the
 * kernel files (device code only).
 */
#ifndef CABLEGEN_DEVUTIL_H
#define CABLEGEN_DEVUTIL_H
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "kernels.h"

namespace cg {

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

/* Exclusive scan over the block (blockDim.x <= 1024).  `wtmp` holds 17 T. */
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

/* Inclusive scan of a 32-bit value over the 64 lanes with DPP moves (no LDS
 * round trip): shifts by 1, 2, 4 and 8 inside each row of 16 lanes, then the
 * last lane of row 0 (2) added to row 1 (3) and lane 31 to rows 2 and 3.
 * Lanes whose DPP source lies outside the row read 0. */
__device__ __forceinline__ uint32_t wave_incl_scan_dpp(uint32_t x)
{
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xf, 0xf, false);   /* row_shr:1 */
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xf, 0xf, false);   /* row_shr:2 */
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xf, 0xf, false);   /* row_shr:4 */
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xf, 0xf, false);   /* row_shr:8 */
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xa, 0xf, false);   /* row_bcast:15 */
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xc, 0xf, false);   /* row_bcast:31 */
    return x;
}

/* Exclusive scan over a block of up to 16 waves with one barrier: each wave
 * scans with DPP, lane 63 publishes the wave total to wtmp[w], and every
 * wave scans the totals of the waves before it.  The caller separates two
 * uses of wtmp by a barrier. */
__device__ __forceinline__ uint32_t block_excl_scan_dpp(uint32_t v, uint32_t *total, uint32_t *wtmp)
{
    const int w = threadIdx.x >> 6, lane = lane_id(), nw = (blockDim.x + 63) >> 6;
    const uint32_t inc = wave_incl_scan_dpp(v);
    if (lane == 63) wtmp[w] = inc;
    __syncthreads();
    const uint32_t t = lane < nw ? wtmp[lane] : 0u;
    const uint32_t ts = wave_incl_scan_dpp(t);
    const uint32_t before = w ? (uint32_t)__shfl(ts, w - 1, 64) : 0u;
    *total = (uint32_t)__shfl(ts, nw - 1, 64);
    return before + inc - v;
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

/* packs the varying bits of k (KeyXform) into the low x.nbits bits, keeping
 * their order; branch-free: unused segments have mask 0 and length 0 */
__host__ __device__ __forceinline__ uint64_t xform(uint64_t k, const KeyXform &x)
{
    uint64_t r = 0;
#pragma unroll
    for (int s = 0; s < 4; s++) {
        uint64_t part = (k >> x.lo[s]) & x.mask[s];
        r = (x.len[s] >= 64 ? 0ULL : (r << x.len[s])) | part;
    }
    return r;
}

/* inverse of xform on the packed bits: scatters c back into the varying
 * positions (pdep-like); the other bits are zero */
__host__ __device__ __forceinline__ uint64_t xform_inv(uint64_t c, const KeyXform &x)
{
    uint64_t b = 0;
#pragma unroll
    for (int s = 3; s >= 0; s--) {
        b |= (c & x.mask[s]) << x.lo[s];
        c = x.len[s] >= 64 ? 0ULL : (c >> x.len[s]);
    }
    return b;
}

__host__ __device__ __forceinline__ uint32_t sc_cell(uint32_t hi, uint32_t lo, int at)
{
    return ((at >= 32 ? hi : lo) >> (at & 31)) & 15u;
}

/* the sort code of key k (SortCode); bad becomes nonzero when a cell of the
 * cell-pair code is outside its base */
__host__ __device__ __forceinline__ uint64_t sc_encode(uint64_t k, const SortCode &c, uint32_t &bad)
{
    if (!c.cells) return xform(k, c.x);
    const uint32_t hi = (uint32_t)(k >> 32), lo = (uint32_t)k;
    uint64_t v = 0;
#pragma unroll
    for (int p = 0; p < 8; p++) {
        if (p < c.npair) {
            const uint32_t ba = c.cbase[2 * p], bb = c.cbase[2 * p + 1];
            uint32_t a = sc_cell(hi, lo, c.cell_lo[2 * p]), b = sc_cell(hi, lo, c.cell_lo[2 * p + 1]);
            bad |= (uint32_t)(a >= ba) | (uint32_t)(b >= bb);
            v = (v << c.pbits[p]) | (a * bb + b);
        }
    }
    if (c.ncell & 1) {
        uint32_t a = sc_cell(hi, lo, c.cell_lo[c.ncell - 1]);
        bad |= (uint32_t)(a >= c.cbase[c.ncell - 1]);
        v = (v << c.lbits) | a;
    }
    return v;
}

/* inverse of sc_encode on the covered bits (the other bits are zero) */
__host__ __device__ __forceinline__ uint64_t sc_decode(uint64_t v, const SortCode &c)
{
    if (!c.cells) return xform_inv(v, c.x);
    uint64_t r = 0;
    if (c.ncell & 1) {
        r = (v & ((1ULL << c.lbits) - 1)) << c.cell_lo[c.ncell - 1];
        v >>= c.lbits;
    }
#pragma unroll
    for (int p = 7; p >= 0; p--) {
        if (p < c.npair) {
            uint32_t f = (uint32_t)v & ((1u << c.pbits[p]) - 1);
            v >>= c.pbits[p];
            uint32_t a = (f * c.magic[p]) >> 16, b = f - a * (uint32_t)c.cbase[2 * p + 1];
            r |= ((uint64_t)a << c.cell_lo[2 * p]) | ((uint64_t)b << c.cell_lo[2 * p + 1]);
        }
    }
    return r;
}

} // namespace cg
#endif
