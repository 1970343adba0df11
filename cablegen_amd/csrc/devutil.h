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
