// nb_device.h — wave64 / workgroup primitives shared by the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nb {

__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ uint64_t ballot64(bool p) { return __builtin_amdgcn_ballot_w64(p); }

__device__ __forceinline__ uint64_t lanemask_lt() {
    const int l = lane_id();
    return l == 0 ? 0ull : (~0ull >> (64 - l));
}

__device__ __forceinline__ int readfirst(int v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ uint64_t readfirst64(uint64_t v) {
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}

// Inclusive scan of a u32 across the 64 lanes of a wave.
// Set bits of a wave mask in the lanes below this one (v_mbcnt).
__device__ __forceinline__ uint32_t mask_below(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    const int l = lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t o = __shfl_up(v, d, 64);
        if (l >= d) v += o;
    }
    return v;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

// Exclusive scan over a workgroup of NT threads (NT multiple of 64, <= 1024).
// `smem` holds NT/64 + 1 words. Returns the exclusive prefix; *total gets the sum.
template <int NT>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* smem, uint32_t* total) {
    constexpr int NW = NT / 64;
    const int w = threadIdx.x >> 6, l = lane_id();
    const uint32_t inc = wave_incl_scan(v);
    if (l == 63) smem[w] = inc;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (int k = 0; k < NW; k++) {
            const uint32_t t = smem[k];
            smem[k] = run;
            run += t;
        }
        smem[NW] = run;
    }
    __syncthreads();
    const uint32_t r = smem[w] + inc - v;
    *total = smem[NW];
    __syncthreads();
    return r;
}

// Block-to-XCD remap: the hardware deals consecutive workgroups round-robin
// to the 8 XCDs; this returns a logical block index such that each XCD gets a
// contiguous run of logical blocks, so neighbouring (Morton-adjacent) work
// shares one L2.
__device__ __forceinline__ uint32_t xcd_remap(uint32_t b, uint32_t nblocks) {
    constexpr uint32_t X = 8;
    const uint32_t xcd = b % X, k = b / X;
    const uint32_t q = nblocks / X, r = nblocks % X;
    // XCD x receives q + (x < r) blocks.
    const uint32_t start = xcd * q + (xcd < r ? xcd : r);
    return start + k;
}

// Min/max of x and y over a workgroup of `threads` (a multiple of 64; at
// most 1024); the result is left in thread 0's arguments. Every thread of the
// workgroup must call it.
template <int threads>
__device__ __forceinline__ void box_block_reduce(double& xmin, double& xmax, double& ymin,
                                                 double& ymax) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        xmin = fmin(xmin, __shfl_xor(xmin, d, 64));
        xmax = fmax(xmax, __shfl_xor(xmax, d, 64));
        ymin = fmin(ymin, __shfl_xor(ymin, d, 64));
        ymax = fmax(ymax, __shfl_xor(ymax, d, 64));
    }
    if (threads == 64) return;
    __shared__ double s[4][threads / 64];
    const int w = threadIdx.x >> 6;
    if (lane_id() == 0) {
        s[0][w] = xmin;
        s[1][w] = xmax;
        s[2][w] = ymin;
        s[3][w] = ymax;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < threads / 64; k++) {
            xmin = fmin(xmin, s[0][k]);
            xmax = fmax(xmax, s[1][k]);
            ymin = fmin(ymin, s[2][k]);
            ymax = fmax(ymax, s[3][k]);
        }
    }
}

// The same, stored by thread 0 as part[4 b + 0..3] = xmin, xmax, ymin, ymax
// for b = blockIdx.x.
template <int threads>
__device__ __forceinline__ void box_block_store(double xmin, double xmax, double ymin,
                                                double ymax, double* __restrict__ part) {
    box_block_reduce<threads>(xmin, xmax, ymin, ymax);
    if (threadIdx.x == 0) {
        part[4 * blockIdx.x + 0] = xmin;
        part[4 * blockIdx.x + 1] = xmax;
        part[4 * blockIdx.x + 2] = ymin;
        part[4 * blockIdx.x + 3] = ymax;
    }
}

}  // namespace nb
