// nb_sort.hip — bounding box, Morton keys, LDS-staged LSD radix sort, gather.
//
// Semantics follow BarnesHutCuda.cu: the squared bounding box (:78-114,
// :648-680), 64-bit Morton keys with x in the odd bits (:54-69, :117-132), a
// stable ascending sort over all 64 key bits with the current index as payload
// (cub::DeviceRadixSort::SortPairs, :691-693) and the gather of every body
// array through the permutation (:135-147). The squaring runs on the GPU, so
// a step has no host round trip.
#include <float.h>

#include <algorithm>

#include "nb_device.h"
#include "nb_internal.h"

namespace nb {

// ---------------------------------------------------------------------------
// Bounding box: grid-stride min/max, wave shuffles, one partial per block,
// then a single-block kernel squares the box about the shorter axis.
// HBM traffic: 16 B/body read.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBboxThreads) void k_bbox_partial(const double2* __restrict__ pos,
                                                               uint32_t n,
                                                               double* __restrict__ part) {
    double xmin = INFINITY, xmax = -INFINITY, ymin = INFINITY, ymax = -INFINITY;
    for (uint32_t i = blockIdx.x * kBboxThreads + threadIdx.x; i < n;
         i += gridDim.x * kBboxThreads) {
        const double2 p = pos[i];
        xmin = fmin(xmin, p.x);
        xmax = fmax(xmax, p.x);
        ymin = fmin(ymin, p.y);
        ymax = fmax(ymax, p.y);
    }
    box_block_store<kBboxThreads>(xmin, xmax, ymin, ymax, part);
}

// Clears the radix sort's global histograms and tile counters and advances
// its epoch (see k_os_pass); done by one workgroup of the box kernel that
// starts every step, ahead of the Morton keys and the sort.
__device__ __forceinline__ void sort_glob_reset(uint32_t* __restrict__ g, int tid, int nthreads) {
    for (int k = tid; k < kSortGlobEpoch; k += nthreads) g[k] = 0u;
    if (tid == 0) g[kSortGlobEpoch] += 1u;
}

// Root size^2 of the opening test: side^2 (SIDE), or w^2 + h^2 = 2 side^2 on
// a square cell (DIAG; exact, as the reference's lengthSquared of (s, s)).
__device__ __forceinline__ double root_size_sq(double w, int diag) {
    const double s2 = w * w;
    return diag ? s2 + s2 : s2;
}

__global__ __launch_bounds__(kBboxThreads) void k_bbox_final(const double* __restrict__ part,
                                                             uint32_t nparts,
                                                             double* __restrict__ box,
                                                             uint32_t* __restrict__ meta,
                                                             int diag,
                                                             uint32_t* __restrict__ sort_glob) {
    sort_glob_reset(sort_glob, threadIdx.x, kBboxThreads);
    double xmin = INFINITY, xmax = -INFINITY, ymin = INFINITY, ymax = -INFINITY;
    for (uint32_t k = threadIdx.x; k < nparts; k += kBboxThreads) {
        xmin = fmin(xmin, part[4 * k + 0]);
        xmax = fmax(xmax, part[4 * k + 1]);
        ymin = fmin(ymin, part[4 * k + 2]);
        ymax = fmax(ymax, part[4 * k + 3]);
    }
    box_block_reduce<kBboxThreads>(xmin, xmax, ymin, ymax);
    if (threadIdx.x == 0) {
        // Square about the midpoint of the shorter axis (BarnesHutCuda.cu:665-677).
        const double xr = xmax - xmin, yr = ymax - ymin;
        if (xr > yr) {
            const double mid = (ymin + ymax) * 0.5;
            ymin = mid - xr * 0.5;
            ymax = mid + xr * 0.5;
        } else {
            const double mid = (xmin + xmax) * 0.5;
            xmin = mid - yr * 0.5;
            xmax = mid + yr * 0.5;
        }
        box[0] = xmin;
        box[1] = xmax;
        box[2] = ymin;
        box[3] = ymax;
        box[4] = root_size_sq(xmax - xmin, diag);  // root width_sq (BarnesHutCuda.cu:717)
        meta[META_MULTI_LEAVES] = 0;  // the tree build of this step queues from 0
        meta[META_QHEAD] = 0;         // the traversal work queue of this step starts empty
        meta[META_QTAIL] = 0;
        // flags stay set for the rest of the batch of steps
        if (!isfinite(xmin) || !isfinite(xmax) || !isfinite(ymin) || !isfinite(ymax))
            meta[META_FLAGS] |= (uint32_t)FLAG_NONFINITE;
    }
}

// Several systems: one workgroup per system reduces and squares its box.
__global__ __launch_bounds__(kBboxThreads) void k_bbox_sys(const double2* __restrict__ pos,
                                                           const uint32_t* __restrict__ sys_off,
                                                           double* __restrict__ box,
                                                           uint32_t* __restrict__ meta,
                                                           int diag,
                                                           uint32_t* __restrict__ sort_glob) {
    const uint32_t s = blockIdx.x;
    if (s == 0) sort_glob_reset(sort_glob, threadIdx.x, kBboxThreads);
    const uint32_t a = sys_off[s], e = sys_off[s + 1];
    double xmin = INFINITY, xmax = -INFINITY, ymin = INFINITY, ymax = -INFINITY;
    for (uint32_t i = a + threadIdx.x; i < e; i += kBboxThreads) {
        const double2 p = pos[i];
        xmin = fmin(xmin, p.x);
        xmax = fmax(xmax, p.x);
        ymin = fmin(ymin, p.y);
        ymax = fmax(ymax, p.y);
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        xmin = fmin(xmin, __shfl_xor(xmin, d, 64));
        xmax = fmax(xmax, __shfl_xor(xmax, d, 64));
        ymin = fmin(ymin, __shfl_xor(ymin, d, 64));
        ymax = fmax(ymax, __shfl_xor(ymax, d, 64));
    }
    __shared__ double sm[4][kBboxThreads / 64];
    const int w = threadIdx.x >> 6;
    if (lane_id() == 0) {
        sm[0][w] = xmin;
        sm[1][w] = xmax;
        sm[2][w] = ymin;
        sm[3][w] = ymax;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < kBboxThreads / 64; k++) {
            xmin = fmin(xmin, sm[0][k]);
            xmax = fmax(xmax, sm[1][k]);
            ymin = fmin(ymin, sm[2][k]);
            ymax = fmax(ymax, sm[3][k]);
        }
        const double xr = xmax - xmin, yr = ymax - ymin;
        if (xr > yr) {
            const double mid = (ymin + ymax) * 0.5;
            ymin = mid - xr * 0.5;
            ymax = mid + xr * 0.5;
        } else {
            const double mid = (xmin + xmax) * 0.5;
            xmin = mid - yr * 0.5;
            xmax = mid + yr * 0.5;
        }
        double* bx = box + (uint64_t)kBoxStride * s;
        bx[0] = xmin;
        bx[1] = xmax;
        bx[2] = ymin;
        bx[3] = ymax;
        bx[4] = root_size_sq(xmax - xmin, diag);
        if (s == 0) {
            meta[META_MULTI_LEAVES] = 0;
            meta[META_QHEAD] = 0;
            meta[META_QTAIL] = 0;
        }
        if (!isfinite(xmin) || !isfinite(xmax) || !isfinite(ymin) || !isfinite(ymax))
            atomicOr(&meta[META_FLAGS], (uint32_t)FLAG_NONFINITE);
    }
}

// ---------------------------------------------------------------------------
// Reference CPU box (get_universe_boundaries, Quadtree.cpp:9-31), scanned in
// the caller's body order as the reference does. The minimum is the true
// one; body i can raise the maximum only when it did not lower the minimum
// (the `else if`), i.e. when !(x_i < min(DBL_MAX, x_0 .. x_{i-1})), so body 0
// never does (unless its x is DBL_MAX). The box stays rectangular:
// {min, max - min}; a maximum no body raised stays -DBL_MAX, as there.
//   k_refbox_scatter  positions into the caller's order (4 B index + 16 B
//                     read + 16 B written per body)
//   k_refbox_min      per tile of kRefBoxTile bodies: its minima
//   k_refbox_max      per tile: the running minimum entering each body (the
//                     tiles before it, then an in-tile scan) and the maximum
//                     of the bodies that pass the test
//   k_refbox_final    one wave: the box and its root size^2 w^2 + h^2
// ---------------------------------------------------------------------------
constexpr int kRefThreads = 256;
constexpr int kRefItems = kRefBoxTile / kRefThreads;

__global__ __launch_bounds__(256) void k_refbox_scatter(const double2* __restrict__ pos,
                                                        const uint32_t* __restrict__ orig,
                                                        uint32_t n, double2* __restrict__ out) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    out[orig[i]] = pos[i];
}

__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v = fmin(v, __shfl_xor(v, d, 64));
    return v;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v = fmax(v, __shfl_xor(v, d, 64));
    return v;
}

__global__ __launch_bounds__(kRefThreads) void k_refbox_min(const double2* __restrict__ p,
                                                            uint32_t n,
                                                            double4* __restrict__ part) {
    double mx = DBL_MAX, my = DBL_MAX;
    const uint64_t base = (uint64_t)blockIdx.x * kRefBoxTile;
#pragma unroll
    for (int r = 0; r < kRefItems; r++) {
        const uint64_t o = base + (uint64_t)r * kRefThreads + threadIdx.x;
        if (o < n) {
            const double2 q = p[o];
            mx = fmin(mx, q.x);
            my = fmin(my, q.y);
        }
    }
    mx = wave_min(mx);
    my = wave_min(my);
    __shared__ double s[2][kRefThreads / 64];
    const int w = threadIdx.x >> 6;
    if (lane_id() == 0) {
        s[0][w] = mx;
        s[1][w] = my;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < kRefThreads / 64; k++) {
            mx = fmin(mx, s[0][k]);
            my = fmin(my, s[1][k]);
        }
        part[blockIdx.x] = make_double4(mx, my, -DBL_MAX, -DBL_MAX);
    }
}

// Exclusive min-scan of one double over the workgroup (thread order).
__device__ __forceinline__ double block_excl_min(double v, double* smem) {
    const int l = lane_id(), w = threadIdx.x >> 6;
    double inc = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const double o = __shfl_up(inc, d, 64);
        if (l >= d) inc = fmin(inc, o);
    }
    if (l == 63) smem[w] = inc;
    __syncthreads();
    double before = DBL_MAX;  // waves before this one
    for (int k = 0; k < w; k++) before = fmin(before, smem[k]);
    double ex = __shfl_up(inc, 1, 64);
    if (l == 0) ex = DBL_MAX;
    __syncthreads();  // smem may be reused by the caller
    return fmin(before, ex);
}

__global__ __launch_bounds__(kRefThreads) void k_refbox_max(const double2* __restrict__ p,
                                                            uint32_t n,
                                                            double4* __restrict__ part) {
    __shared__ double s[2][kRefThreads / 64];
    __shared__ double sr[4][kRefThreads / 64];
    const int t = threadIdx.x, w = t >> 6;
    // running minima entering this tile: the minima of every tile before it
    double in_x = DBL_MAX, in_y = DBL_MAX;
    for (uint32_t k = t; k < blockIdx.x; k += kRefThreads) {
        const double4 q = part[k];
        in_x = fmin(in_x, q.x);
        in_y = fmin(in_y, q.y);
    }
    in_x = wave_min(in_x);
    in_y = wave_min(in_y);
    if (lane_id() == 0) {
        s[0][w] = in_x;
        s[1][w] = in_y;
    }
    __syncthreads();
    for (int k = 0; k < kRefThreads / 64; k++) {
        in_x = fmin(in_x, s[0][k]);
        in_y = fmin(in_y, s[1][k]);
    }
    // thread t owns bodies base + t * kRefItems .. + kRefItems - 1 (caller order)
    const uint64_t base = (uint64_t)blockIdx.x * kRefBoxTile + (uint64_t)t * kRefItems;
    double2 q[kRefItems];
    double tx = DBL_MAX, ty = DBL_MAX;
#pragma unroll
    for (int r = 0; r < kRefItems; r++) {
        q[r] = base + r < n ? p[base + r] : make_double2(DBL_MAX, DBL_MAX);
        tx = fmin(tx, q[r].x);
        ty = fmin(ty, q[r].y);
    }
    double run_x = fmin(in_x, block_excl_min(tx, sr[0]));
    double run_y = fmin(in_y, block_excl_min(ty, sr[1]));
    double mx = -DBL_MAX, my = -DBL_MAX;
#pragma unroll
    for (int r = 0; r < kRefItems; r++) {
        if (base + r >= n) break;
        // if (x < min) min = x; else if (x > max) max = x;   (Quadtree.cpp:16-21)
        if (q[r].x < run_x)
            run_x = q[r].x;
        else
            mx = fmax(mx, q[r].x);
        if (q[r].y < run_y)
            run_y = q[r].y;
        else
            my = fmax(my, q[r].y);
    }
    mx = wave_max(mx);
    my = wave_max(my);
    if (lane_id() == 0) {
        sr[2][w] = mx;
        sr[3][w] = my;
    }
    __syncthreads();
    if (t == 0) {
        for (int k = 1; k < kRefThreads / 64; k++) {
            mx = fmax(mx, sr[2][k]);
            my = fmax(my, sr[3][k]);
        }
        // only this tile's maxima: other tiles read its minima meanwhile
        double* o = reinterpret_cast<double*>(&part[blockIdx.x]);
        o[2] = mx;
        o[3] = my;
    }
}

__global__ __launch_bounds__(64) void k_refbox_final(const double4* __restrict__ part,
                                                     uint32_t nparts, double* __restrict__ box,
                                                     uint32_t* __restrict__ meta,
                                                     uint32_t* __restrict__ sort_glob) {
    sort_glob_reset(sort_glob, threadIdx.x, 64);
    double xmin = DBL_MAX, ymin = DBL_MAX, xmax = -DBL_MAX, ymax = -DBL_MAX;
    for (uint32_t k = threadIdx.x; k < nparts; k += 64) {
        const double4 q = part[k];
        xmin = fmin(xmin, q.x);
        ymin = fmin(ymin, q.y);
        xmax = fmax(xmax, q.z);
        ymax = fmax(ymax, q.w);
    }
    xmin = wave_min(xmin);
    ymin = wave_min(ymin);
    xmax = wave_max(xmax);
    ymax = wave_max(ymax);
    if (threadIdx.x == 0) {
        box[0] = xmin;
        box[1] = xmax;
        box[2] = ymin;
        box[3] = ymax;
        const double w = xmax - xmin, h = ymax - ymin;  // Quadtree.cpp:30
        {
            // the root's size.lengthSquared() (BarnesHut.cpp:115), rounded as
            // written: no multiply-add
#pragma clang fp contract(off)
            const double w2 = w * w, h2 = h * h;
            box[4] = w2 + h2;
        }
        meta[META_MULTI_LEAVES] = 0;
        meta[META_QHEAD] = 0;
        meta[META_QTAIL] = 0;
        if (!isfinite(xmin) || !isfinite(ymin) || !isfinite(w) || !isfinite(h))
            meta[META_FLAGS] |= (uint32_t)FLAG_NONFINITE;
    }
}

bool box_from_drift(const nb_ctx* c) {
    return c->nsys == 1 && c->p.cells != NB_CELLS_REF_CPU &&
           c->p.algorithm == NB_ALGO_BARNES_HUT;
}

void ensure_box_parts(nb_ctx* c, int src) {
    if (!box_from_drift(c) || c->box_parts_ok || c->n == 0) return;
    hipLaunchKernelGGL(k_bbox_partial, dim3(c->bbox_blocks), dim3(kBboxThreads), 0, c->stream,
                       c->b.pos[src], (uint32_t)c->n, c->b.bbox_part);
    c->box_parts_ok = true;
}

void launch_bbox(nb_ctx* c, int src) {
    const int diag = c->p.mac == NB_MAC_DIAG;
    if (c->p.cells == NB_CELLS_REF_CPU) {
        const uint32_t n = (uint32_t)c->n;
        hipLaunchKernelGGL(k_refbox_scatter, dim3((n + 255) / 256), dim3(256), 0, c->stream,
                           c->b.pos[src], c->b.orig[src], n, c->b.ref_pos);
        hipLaunchKernelGGL(k_refbox_min, dim3(c->ref_blocks), dim3(kRefThreads), 0, c->stream,
                           c->b.ref_pos, n, c->b.ref_part);
        hipLaunchKernelGGL(k_refbox_max, dim3(c->ref_blocks), dim3(kRefThreads), 0, c->stream,
                           c->b.ref_pos, n, c->b.ref_part);
        hipLaunchKernelGGL(k_refbox_final, dim3(1), dim3(64), 0, c->stream, c->b.ref_part,
                           c->ref_blocks, c->b.box, c->b.meta, c->b.sort_glob);
        return;
    }
    if (c->nsys > 1) {
        hipLaunchKernelGGL(k_bbox_sys, dim3(c->nsys), dim3(kBboxThreads), 0, c->stream,
                           c->b.pos[src], c->b.sys_off, c->b.box, c->b.meta, diag,
                           c->b.sort_glob);
        return;
    }
    if (box_from_drift(c)) {
        // the previous drift wrote the partials (graph replays are preceded
        // by ensure_box_parts, so a capture records only the final reduction)
        if (!c->capturing) ensure_box_parts(c, src);
    } else {
        hipLaunchKernelGGL(k_bbox_partial, dim3(c->bbox_blocks), dim3(kBboxThreads), 0,
                           c->stream, c->b.pos[src], (uint32_t)c->n, c->b.bbox_part);
    }
    hipLaunchKernelGGL(k_bbox_final, dim3(1), dim3(kBboxThreads), 0, c->stream, c->b.bbox_part,
                       c->bbox_blocks, c->b.box, c->b.meta, diag, c->b.sort_glob);
}

// ---------------------------------------------------------------------------
// Radix-sort digits and their global histograms (see the sort below). The
// Morton kernels count the digits of every pass while they produce the keys
// (per workgroup in LDS, one row per workgroup), k_os_reduce adds the rows.
// ---------------------------------------------------------------------------
constexpr int kKeyThreads = 1024;  // threads of a Morton workgroup (4 wavefronts per SIMD)
constexpr int kKeyItems = 4;       // keys per thread of the Morton kernels
constexpr int kKeyTile = kKeyThreads * kKeyItems;

// Digit of pass p: key passes take bits [8p, 8p + 8) of the key; the system
// passes (p >= 8) take bits [8(p - 8), ...) of the system of the key's
// source position.
__device__ __forceinline__ uint32_t os_digit(uint64_t key, uint32_t sys, int pass) {
    return pass < 8 ? (uint32_t)(key >> (8 * pass)) & (kSortBins - 1)
                    : (sys >> (8 * (pass - 8))) & (kSortBins - 1);
}

// Lanes of this wave whose digit equals mine (valid lanes only).
__device__ __forceinline__ uint64_t digit_peers(uint32_t d, bool valid) {
    uint64_t peers = ballot64(valid);
#pragma unroll
    for (int b = 0; b < kSortBits; b++) {
        const bool bit = (d >> b) & 1u;
        const uint64_t m = ballot64(valid && bit);
        peers &= bit ? m : ~m;
    }
    return peers;
}

__device__ __forceinline__ void os_hist_clear(uint32_t (*h)[kSortBins]) {
    for (int k = threadIdx.x; k < kSortMaxPasses * kSortBins; k += blockDim.x) (&h[0][0])[k] = 0u;
}

// Counts one key per valid lane. The keys of a wave are 64 consecutive ones
// in the previous step's order, nearly sorted, so their high digits are
// mostly equal: such a digit is counted with one LDS add for the wave.
__device__ __forceinline__ void os_hist_add(uint32_t (*h)[kSortBins], uint64_t key, uint32_t sys,
                                            int npass, bool valid) {
    const uint64_t active = ballot64(valid);
    if (active == 0) return;
    const bool leader = (active & lanemask_lt()) == 0 && valid;
#pragma unroll
    for (int p = 0; p < kSortMaxPasses; p++) {
        if (p >= npass) break;
        const uint32_t d = os_digit(key, sys, p);
        // lane 0's digit (a scalar read); the wave counts at once when every
        // valid lane has it
        const uint32_t d0 = (uint32_t)readfirst((int)d);
        if (ballot64(valid && d == d0) == active) {
            if (leader) atomicAdd(&h[p][d0], (uint32_t)__popcll(active));
        } else if (valid) {
            atomicAdd(&h[p][d], 1u);
        }
    }
}

// Writes the workgroup's counts (all bins) to its row of `part`
// ([blocks][kSortMaxPasses * kSortBins]); k_os_reduce adds the rows up.
// (Global atomics from every workgroup would serialise on each bin.)
__device__ __forceinline__ void os_hist_flush(uint32_t (*h)[kSortBins], uint32_t* __restrict__ part,
                                              int npass) {
    __syncthreads();
    uint32_t* row = part + (uint64_t)blockIdx.x * (kSortMaxPasses * kSortBins);
    for (int k = threadIdx.x; k < npass * kSortBins; k += blockDim.x) row[k] = (&h[0][0])[k];
}

// Global digit histograms: workgroup (pass p, slice s) adds up the rows
// b = s, s + kOsSlices, ... of `part` for every digit (one thread per
// digit), then one atomic per digit into the cleared global counts.
constexpr int kOsSlices = 16;
__global__ __launch_bounds__(kSortBins) void k_os_reduce(const uint32_t* __restrict__ part,
                                                         uint32_t rows,
                                                         uint32_t* __restrict__ glob) {
    const int pass = blockIdx.x, d = threadIdx.x;
    uint32_t sum = 0;
    uint32_t b = blockIdx.y;
    for (; b + 7 * kOsSlices < rows; b += 8 * kOsSlices) {
        uint32_t v[8];
#pragma unroll
        for (int j = 0; j < 8; j++)
            v[j] = part[(uint64_t)(b + j * kOsSlices) * (kSortMaxPasses * kSortBins) +
                        pass * kSortBins + d];
#pragma unroll
        for (int j = 0; j < 8; j++) sum += v[j];
    }
    for (; b < rows; b += kOsSlices)
        sum += part[(uint64_t)b * (kSortMaxPasses * kSortBins) + pass * kSortBins + d];
    if (sum) atomicAdd(&glob[pass * kSortBins + d], sum);
}

// ---------------------------------------------------------------------------
// Morton keys. HBM traffic: 16 B read + 12 B written per body.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t spread_bits(uint32_t v) {
    uint64_t x = v;
    x = (x | (x << 16)) & 0x0000FFFF0000FFFFull;
    x = (x | (x << 8)) & 0x00FF00FF00FF00FFull;
    x = (x | (x << 4)) & 0x0F0F0F0F0F0F0F0Full;
    x = (x | (x << 2)) & 0x3333333333333333ull;
    x = (x | (x << 1)) & 0x5555555555555555ull;
    return x;
}

// floor(clamp(nrm * 2^32, 0, 2^32 - 1)); fmax maps NaN to 0.
__device__ __forceinline__ uint32_t quantise(double v, double lo, double inv) {
    const double nrm = (v - lo) * inv;
    const double s = fmin(fmax(nrm * 4294967296.0, 0.0), 4294967295.0);
    return (uint32_t)floor(s);
}

__global__ __launch_bounds__(kKeyThreads) void k_morton(const double2* __restrict__ pos, uint32_t n,
                                                const double* __restrict__ box,
                                                const uint32_t* __restrict__ sys_of,
                                                uint64_t* __restrict__ keys,
                                                uint32_t* __restrict__ vals, int npass,
                                                uint32_t* __restrict__ sort_part) {
    __shared__ uint32_t h[kSortMaxPasses][kSortBins];
    os_hist_clear(h);
    // every load of the workgroup's keys is issued before the first use
    double2 pp[kKeyItems];
    uint32_t ss[kKeyItems];
#pragma unroll
    for (int r = 0; r < kKeyItems; r++) {
        const uint32_t i = blockIdx.x * kKeyTile + r * kKeyThreads + threadIdx.x;
        pp[r] = i < n ? pos[i] : make_double2(0.0, 0.0);
        ss[r] = i < n ? sys_of[i] : 0u;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kKeyItems; r++) {
        const uint32_t i = blockIdx.x * kKeyTile + r * kKeyThreads + threadIdx.x;
        const bool valid = i < n;
        uint64_t key = 0;
        if (valid) {
            const double* bx = box + (uint64_t)kBoxStride * ss[r];
            const double x_min = bx[0], x_max = bx[1], y_min = bx[2], y_max = bx[3];
            const double inv_w = 1.0 / (x_max - x_min);
            const double inv_h = 1.0 / (y_max - y_min);
            const uint32_t ix = quantise(pp[r].x, x_min, inv_w);
            const uint32_t iy = quantise(pp[r].y, y_min, inv_h);
            key = (spread_bits(ix) << 1) | spread_bits(iy);
            keys[i] = key;
            vals[i] = i;
        }
        os_hist_add(h, key, ss[r], npass, valid);
    }
    os_hist_flush(h, sort_part, npass);
}

// Keys of the reference CPU quadtree (REF_CPU cells): the path of each body
// down fill_tree_recursive's splits (Quadtree.cpp:57-106), one base-4 digit
// per level for levels 0..29, x in the odd bits as for Morton keys (digit q =
// right << 1 | bottom, right = !(x < cx), bottom = !(y < cy)). The centre and
// the halved size of every level are computed as the reference computes them
// (Rect::getCenter = position + size / 2; children at the centre or the
// parent's position with size / 2), so a body's digits are its reference
// child slots; the 4 low key bits stay zero. HBM traffic: 16 B read + 12 B
// written per body (the descent is ~200 FP64 operations in registers).
__device__ __forceinline__ uint64_t ref_key(double2 p, const double* __restrict__ box) {
    double px = box[0], py = box[2];
    double sx = box[1] - box[0], sy = box[3] - box[2];
    uint64_t key = 0;
#pragma unroll 6
    for (int L = 0; L < kMaxLevel; L++) {
        const double cx = px + sx / 2.0, cy = py + sy / 2.0;
        const bool right = !(p.x < cx), bottom = !(p.y < cy);
        key |= (uint64_t)((right ? 2u : 0u) | (bottom ? 1u : 0u)) << (62 - 2 * L);
        px = right ? cx : px;
        py = bottom ? cy : py;
        sx = sx / 2.0;
        sy = sy / 2.0;
    }
    return key;
}

__global__ __launch_bounds__(kKeyThreads) void k_morton_ref(const double2* __restrict__ pos, uint32_t n,
                                                    const double* __restrict__ box,
                                                    uint64_t* __restrict__ keys,
                                                    uint32_t* __restrict__ vals,
                                                    uint32_t* __restrict__ sort_part) {
    __shared__ uint32_t h[kSortMaxPasses][kSortBins];
    os_hist_clear(h);
    double2 pp[kKeyItems];
#pragma unroll
    for (int r = 0; r < kKeyItems; r++) {
        const uint32_t i = blockIdx.x * kKeyTile + r * kKeyThreads + threadIdx.x;
        pp[r] = i < n ? pos[i] : make_double2(0.0, 0.0);
    }
    __syncthreads();
#pragma unroll 2
    for (int r = 0; r < kKeyItems; r++) {
        const uint32_t i = blockIdx.x * kKeyTile + r * kKeyThreads + threadIdx.x;
        const bool valid = i < n;
        uint64_t key = 0;
        if (valid) {
            key = ref_key(pp[r], box);
            keys[i] = key;
            vals[i] = i;
        }
        os_hist_add(h, key, 0u, 8, valid);
    }
    os_hist_flush(h, sort_part, 8);
}

void launch_morton(nb_ctx* c, int src) {
    const uint32_t n = (uint32_t)c->n;
    const uint32_t grid = (n + kKeyTile - 1) / kKeyTile;
    const int npass = 64 / kSortBits + (c->nsys > 1 ? 2 : 0);
    if (c->p.cells == NB_CELLS_REF_CPU)
        hipLaunchKernelGGL(k_morton_ref, dim3(grid), dim3(kKeyThreads), 0, c->stream, c->b.pos[src], n,
                           c->b.box, c->b.keys[0], c->b.vals[0], c->b.sort_part);
    else
        hipLaunchKernelGGL(k_morton, dim3(grid), dim3(kKeyThreads), 0, c->stream, c->b.pos[src], n,
                           c->b.box, c->b.sys_of, c->b.keys[0], c->b.vals[0], npass,
                           c->b.sort_part);
    hipLaunchKernelGGL(k_os_reduce, dim3(npass, kOsSlices), dim3(kSortBins), 0, c->stream,
                       c->b.sort_part, grid, c->b.sort_glob);
}

// ---------------------------------------------------------------------------
// LSD radix sort, 8-bit digits, 8 passes (keys ping-pong keys[0] -> keys[1]
// -> ... -> keys[0]), one launch per pass (a onesweep sort). The global
// digit histogram of every pass is counted by the Morton kernel as it makes
// the keys. Per pass:
//   k_os_pass  per tile of kSortTile keys: a stable rank of every key
//              among the tile's keys of its digit (8 ballots give a key's
//              same-digit lanes); the tile's digit counts are published, and
//              its exclusive prefix over the earlier tiles is read back from
//              what they and their groups published (a decoupled look-back,
//              see below); the tile is reordered by digit in LDS and written
//              out in runs of equal digits.
// Published words (u64): flag (1 = count, 2 = inclusive prefix) << 62 |
// tag << 32 | value, tag = (epoch * 16 + pass) mod 2^30 with an epoch that
// every sort advances, so words left by earlier sorts never match. The
// buffers' global histograms and tile counters are cleared by the box kernel
// that starts every step.
// HBM traffic per pass: 12 B read + 12 B written per key.
// ---------------------------------------------------------------------------
constexpr uint32_t kOsSpinLimit = 1u << 21;

__device__ __forceinline__ uint64_t os_word(uint32_t flag, uint32_t tag, uint32_t v) {
    return ((uint64_t)flag << 62) | ((uint64_t)tag << 32) | v;
}

// Exclusive prefix of column `d` over entries 0 .. idx-1 of a published
// array whose entries are counts (flag 1) or inclusive prefixes (flag 2):
// walks back from idx - 1, up to 32 entries per round trip, to the nearest
// inclusive prefix. Returns false if an entry never appears (the wait is
// cut at kOsSpinLimit rounds).
__device__ __forceinline__ bool os_lookback(const uint64_t* __restrict__ status, uint32_t idx,
                                            uint32_t d, uint32_t tag, uint32_t* excl_out) {
    constexpr int kAhead = 32;
    uint32_t excl = 0, spins = 0;
    int64_t k = (int64_t)idx - 1;
    while (k >= 0) {
        uint64_t v[kAhead];
#pragma unroll
        for (int j = 0; j < kAhead; j++)
            v[j] = k - j >= 0 ? __hip_atomic_load(status + (uint64_t)(k - j) * kSortBins + d,
                                                  __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                              : 0ull;
        bool done = false, stalled = false;
        int used = 0;
#pragma unroll
        for (int j = 0; j < kAhead; j++) {
            if (done || stalled || k - j < 0) continue;
            const uint32_t f = (uint32_t)(v[j] >> 62), vt = (uint32_t)(v[j] >> 32) & 0x3fffffffu;
            if (vt != tag || f == 0u) {
                stalled = true;  // retry from entry k - j
                continue;
            }
            excl += (uint32_t)v[j];
            used = j + 1;
            if (f == 2u) done = true;
        }
        if (done) break;
        k -= used;
        if (stalled) {
            if (++spins > kOsSpinLimit) return false;
            __builtin_amdgcn_s_sleep(1);
        }
    }
    *excl_out = excl;
    return true;
}

// Sum of the counts (flag 1) of entries [first, first + cnt) of column d,
// cnt <= kOsGroup, waiting (bounded) for entries not yet published.
constexpr int kOsGroup = 16;  // tiles per look-back group
__device__ __forceinline__ bool os_sum_counts(const uint64_t* __restrict__ status, uint32_t first,
                                              uint32_t cnt, uint32_t d, uint32_t tag,
                                              uint32_t* out) {
    uint32_t sum = 0, have = 0, spins = 0;
    while (have < cnt) {
        uint64_t v[kOsGroup];
#pragma unroll
        for (int j = 0; j < kOsGroup; j++)
            v[j] = have + j < cnt
                       ? __hip_atomic_load(status + (uint64_t)(first + have + j) * kSortBins + d,
                                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                       : 0ull;
        uint32_t got = 0;
        bool stop = false;
#pragma unroll
        for (int j = 0; j < kOsGroup; j++) {
            if (stop || have + j >= cnt) continue;
            const uint32_t f = (uint32_t)(v[j] >> 62), vt = (uint32_t)(v[j] >> 32) & 0x3fffffffu;
            if (vt != tag || f == 0u) {
                stop = true;
                continue;
            }
            sum += (uint32_t)v[j];
            got++;
        }
        have += got;
        if (have < cnt) {
            if (++spins > kOsSpinLimit) return false;
            __builtin_amdgcn_s_sleep(1);
        }
    }
    *out = sum;
    return true;
}

// One pass: 1024 threads (16 waves) per tile of kSortTile keys, tile =
// workgroup index. Workgroups are dispatched in index order (per XCD), so
// the lowest unfinished tile is always running and every wait below ends;
// the bounded spins only guard against that assumption failing, and then
// flag the step (FLAG_SORT) instead of hanging.
// Prefix of a tile: its own count is published at once; the counts of the
// earlier tiles of its group of kOsGroup are summed (one round trip), and
// the groups before it are looked back over (counts / inclusive prefixes
// published by each group's last tile): about two round trips after the
// counts appear, however many tiles there are.
constexpr int kOsThreads = 1024;
constexpr int kOsItems = kSortTile / kOsThreads;
static_assert(kOsThreads * kOsItems == kSortTile, "tile size");

template <bool SYS>
__global__ __launch_bounds__(kOsThreads) void k_os_pass(
    const uint64_t* __restrict__ kin, const uint32_t* __restrict__ vin, uint64_t* __restrict__ kout,
    uint32_t* __restrict__ vout, const uint32_t* __restrict__ sys_of, uint32_t n, int pass,
    const uint32_t* __restrict__ glob, uint64_t* __restrict__ status, uint64_t* __restrict__ gstatus,
    uint32_t* __restrict__ meta) {
    constexpr int NW = kOsThreads / 64;
    constexpr int kWaveKeys = kSortTile / NW;  // each wave ranks a contiguous part
    __shared__ uint64_t s_key[kSortTile];
    __shared__ uint32_t s_val[kSortTile];
    __shared__ uint8_t s_dig[kSortTile];
    __shared__ uint32_t wc[NW][kSortBins];   // per wave: keys of each digit, then its offset
    __shared__ uint32_t s_off[kSortBins];    // first slot of each digit in the tile
    __shared__ uint32_t s_dst[kSortBins];    // output position of each digit's first key
    __shared__ uint32_t s_scan[NW + 1];
    const int t = threadIdx.x, w = t >> 6, l = lane_id();
    for (int k = t; k < NW * kSortBins; k += kOsThreads) (&wc[0][0])[k] = 0u;
    __syncthreads();
    const uint32_t tile = blockIdx.x;
    const uint64_t base = (uint64_t)tile * kSortTile;
    const uint32_t nvalid = (uint32_t)min((uint64_t)kSortTile, (uint64_t)n - base);
    // Wave w ranks keys [w * kWaveKeys, (w + 1) * kWaveKeys) of the tile in
    // rounds of 64 consecutive keys, against its own digit counters (LDS
    // accesses of one wave are ordered: no barrier): stable within the wave.
    uint64_t key[kOsItems];
    uint32_t val[kOsItems], rank[kOsItems];
    uint8_t dig[kOsItems];
#pragma unroll
    for (int r = 0; r < kOsItems; r++) {
        const uint32_t k = (uint32_t)(w * kWaveKeys + r * 64 + l);
        const bool valid = k < nvalid;
        key[r] = valid ? kin[base + k] : 0ull;
        val[r] = valid ? vin[base + k] : 0u;
    }
#pragma unroll
    for (int r = 0; r < kOsItems; r++) {
        const uint32_t k = (uint32_t)(w * kWaveKeys + r * 64 + l);
        const bool valid = k < nvalid;
        const uint32_t d = valid ? os_digit(key[r], SYS ? sys_of[val[r]] : 0u, pass) : 0u;
        dig[r] = (uint8_t)d;
        const uint64_t peers = digit_peers(d, valid);
        const uint64_t below = peers & lanemask_lt();
        const uint32_t before = valid ? wc[w][d] : 0u;
        rank[r] = before + (uint32_t)__popcll(below);
        if (valid && below == 0) wc[w][d] = before + (uint32_t)__popcll(peers);
    }
    __syncthreads();
    // threads t < 256 own digit t: its count in the tile and each wave's offset
    const bool owner = t < kSortBins;
    uint32_t cnt = 0;
    if (owner) {
#pragma unroll
        for (int k = 0; k < NW; k++) {
            const uint32_t c = wc[k][t];
            wc[k][t] = cnt;
            cnt += c;
        }
    }
    uint32_t total;
    const uint32_t toff = block_excl_scan<kOsThreads>(cnt, s_scan, &total);
    if (owner) s_off[t] = toff;
    // the tile's exclusive prefix over the earlier tiles, digit by digit: the
    // counts are published first, and the scan of the global counts and the
    // reordering of the tile in LDS (neither needs the prefix) run while
    // they travel to the later tiles
    const uint32_t tag = (glob[kSortGlobEpoch] * 16u + (uint32_t)pass) & 0x3fffffffu;
    const uint32_t g = tile / kOsGroup, j = tile % kOsGroup;
    if (owner)
        __hip_atomic_store(status + (uint64_t)tile * kSortBins + t, os_word(1u, tag, cnt),
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint32_t gtotal;
    const uint32_t gex =
        block_excl_scan<kOsThreads>(owner ? glob[pass * kSortBins + t] : 0u, s_scan, &gtotal);
    // the tile in digit order in LDS (s_off and the wave offsets in wc are
    // visible after the scan's barriers)
#pragma unroll
    for (int r = 0; r < kOsItems; r++) {
        const uint32_t k = (uint32_t)(w * kWaveKeys + r * 64 + l);
        if (k < nvalid) {
            const uint32_t slot = s_off[dig[r]] + wc[w][dig[r]] + rank[r];
            s_key[slot] = key[r];
            s_val[slot] = val[r];
            s_dig[slot] = dig[r];
        }
    }
    if (owner) {
        uint32_t within = 0, groups = 0;
        bool ok = os_sum_counts(status, g * kOsGroup, j, (uint32_t)t, tag, &within);
        uint64_t* gmine = gstatus + (uint64_t)g * kSortBins + t;
        const bool last = j == kOsGroup - 1;
        if (last)  // the group's count (its inclusive prefix for group 0)
            __hip_atomic_store(gmine, os_word(g == 0 ? 2u : 1u, tag, within + cnt),
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (g > 0) ok = ok && os_lookback(gstatus, g, (uint32_t)t, tag, &groups);
        if (last && g > 0)
            __hip_atomic_store(gmine, os_word(2u, tag, groups + within + cnt), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        if (!ok) atomicOr(&meta[META_FLAGS], (uint32_t)FLAG_SORT);  // never expected
        s_dst[t] = gex + groups + within;
    }
    __syncthreads();
    // runs of equal digits written out
#pragma unroll
    for (int r = 0; r < kOsItems; r++) {
        const uint32_t slot = (uint32_t)r * kOsThreads + t;
        if (slot < nvalid) {
            const uint32_t d = s_dig[slot];
            const uint32_t dst = s_dst[d] + (slot - s_off[d]);
            if (dst < n) {  // always, unless a wait was cut short (FLAG_SORT)
                kout[dst] = s_key[slot];
                vout[dst] = s_val[slot];
            }
        }
    }
}

// Several systems: two more stable passes on the 16-bit system number (of
// the key's source position) put the systems back in place, each sorted by
// its own keys. An even pass count leaves the result in keys[0] / vals[0].
void launch_radix_sort(nb_ctx* c) {
    const uint32_t n = (uint32_t)c->n;
    const uint32_t tiles = c->sort_tiles;
    const int key_passes = 64 / kSortBits, npass = key_passes + (c->nsys > 1 ? 2 : 0);
    // the digit histograms were counted by the Morton kernel
    for (int pass = 0; pass < npass; pass++) {
        const int in = pass & 1, out = in ^ 1;
        uint64_t* gstatus = c->b.sort_status + (uint64_t)tiles * kSortBins;
        if (pass >= key_passes)
            hipLaunchKernelGGL(k_os_pass<true>, dim3(tiles), dim3(kOsThreads), 0, c->stream,
                               c->b.keys[in], c->b.vals[in], c->b.keys[out], c->b.vals[out],
                               c->b.sys_of, n, pass, c->b.sort_glob, c->b.sort_status, gstatus,
                               c->b.meta);
        else
            hipLaunchKernelGGL(k_os_pass<false>, dim3(tiles), dim3(kOsThreads), 0, c->stream,
                               c->b.keys[in], c->b.vals[in], c->b.keys[out], c->b.vals[out],
                               c->b.sys_of, n, pass, c->b.sort_glob, c->b.sort_status, gstatus,
                               c->b.meta);
    }
}

// ---------------------------------------------------------------------------
// Gather every body array through the permutation (BarnesHutCuda.cu:135-147),
// and, in the same launch, the first stage of the tree build on the sorted
// keys: each pair's delta and internal-node count (tree_pair).
// HBM traffic: 4 B perm + 44 B read + 44 B written per body, plus the pair's.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_gather(
    const uint32_t* __restrict__ perm, const double2* __restrict__ pin,
    const double2* __restrict__ vin, const double* __restrict__ min_,
    const uint32_t* __restrict__ oin, double2* __restrict__ pout, double2* __restrict__ vout,
    double* __restrict__ mout, uint32_t* __restrict__ oout, uint32_t n,
    const uint64_t* __restrict__ keys, const uint32_t* __restrict__ sys_of,
    int8_t* __restrict__ delta, uint32_t* __restrict__ cnt) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    if (i + 1 < n) tree_pair(i, keys, sys_of, delta, cnt);
    uint32_t s = perm[i];
    if (s >= n) s = i;  // only after a stalled sort (FLAG_SORT), whose step is void
    pout[i] = pin[s];
    vout[i] = vin[s];
    mout[i] = min_[s];
    oout[i] = oin[s];
}

void launch_gather(nb_ctx* c, int src) {
    const uint32_t n = (uint32_t)c->n;
    const int dst = src ^ 1;
    hipLaunchKernelGGL(k_gather, dim3((n + 255) / 256), dim3(256), 0, c->stream, c->b.vals[0],
                       c->b.pos[src], c->b.vel[src], c->b.mass[src], c->b.orig[src],
                       c->b.pos[dst], c->b.vel[dst], c->b.mass[dst], c->b.orig[dst], n,
                       c->b.keys[0], c->b.sys_of, c->b.delta, c->b.cnt);
}

// ---------------------------------------------------------------------------
// Un-permute the state into the caller's body order for a snapshot
// (scatter_to_original_order, BarnesHutCuda.cu:498-506): out[0..n) holds the
// positions and out[n..2n) the velocities. HBM traffic: 4 B index + 32 B read
// + 32 B written per body.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_unpermute(const uint32_t* __restrict__ orig,
                                                   const double2* __restrict__ pos,
                                                   const double2* __restrict__ vel,
                                                   double2* __restrict__ out, uint32_t n) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const uint32_t j = orig[i];
    out[j] = pos[i];
    out[(uint64_t)n + j] = vel[i];
}

void launch_unpermute(nb_ctx* c, double2* out) {
    const uint32_t n = (uint32_t)c->n;
    hipLaunchKernelGGL(k_unpermute, dim3((n + 255) / 256), dim3(256), 0, c->stream,
                       c->b.orig[c->cur], c->b.pos[c->cur], c->b.vel[c->cur], out, n);
}

// The same into two arrays (positions, velocities): the blocking readback
// un-permutes into the idle copy of the body arrays.
__global__ __launch_bounds__(256) void k_unpermute2(const uint32_t* __restrict__ orig,
                                                    const double2* __restrict__ pos,
                                                    const double2* __restrict__ vel,
                                                    double2* __restrict__ pout,
                                                    double2* __restrict__ vout, uint32_t n) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const uint32_t j = orig[i];
    pout[j] = pos[i];
    vout[j] = vel[i];
}

void launch_unpermute2(nb_ctx* c, double2* pos_out, double2* vel_out) {
    const uint32_t n = (uint32_t)c->n;
    hipLaunchKernelGGL(k_unpermute2, dim3((n + 255) / 256), dim3(256), 0, c->stream,
                       c->b.orig[c->cur], c->b.pos[c->cur], c->b.vel[c->cur], pos_out, vel_out,
                       n);
}

// ---------------------------------------------------------------------------
// Exclusive in-place scan of u32, one launch: each workgroup scans a tile of
// kScanTile values, publishes its total, and reads the prefix of the earlier
// tiles back from what they published (a decoupled look-back over one
// column, 64 tiles per round trip, as the sort's). Published words (u64):
// flag (1 = tile total, 2 = inclusive prefix) << 62 | tag << 32 | value, with
// a tag that every scan advances (the last tile bumps it when it is done:
// by then every tile of the scan has read it), so words of earlier scans
// never match. Workgroups start in index order, so the lowest unfinished tile
// always runs and every wait ends; the bounded spins flag FLAG_SORT instead
// of hanging if that ever failed.
// HBM traffic: 4 B read + 4 B written per value.
// ---------------------------------------------------------------------------
constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;  // consecutive values per thread
constexpr int kScanTile = kScanThreads * kScanItems;

__global__ __launch_bounds__(kScanThreads) void k_scan(uint32_t* __restrict__ data,
                                                       uint32_t count,
                                                       uint64_t* __restrict__ status,
                                                       uint32_t* __restrict__ epoch,
                                                       uint32_t* __restrict__ total_out,
                                                       uint32_t* __restrict__ meta) {
    __shared__ uint32_t s[kScanThreads / 64 + 1];
    __shared__ uint32_t s_prefix;
    const uint32_t tile = blockIdx.x;
    const uint32_t tag = *epoch & 0x3fffffffu;
    const uint64_t base = (uint64_t)tile * kScanTile + (uint64_t)threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    if (base + kScanItems <= count) {
        const uint4* src = reinterpret_cast<const uint4*>(data + base);
#pragma unroll
        for (int r = 0; r < kScanItems / 4; r++) {
            const uint4 x = src[r];
            v[4 * r] = x.x;
            v[4 * r + 1] = x.y;
            v[4 * r + 2] = x.z;
            v[4 * r + 3] = x.w;
        }
    } else {
#pragma unroll
        for (int r = 0; r < kScanItems; r++) v[r] = base + r < count ? data[base + r] : 0u;
    }
    uint32_t tsum = 0;
#pragma unroll
    for (int r = 0; r < kScanItems; r++) tsum += v[r];
    uint32_t agg;
    const uint32_t ex = block_excl_scan<kScanThreads>(tsum, s, &agg);
    if (threadIdx.x < 64) {
        const int l = lane_id();
        if (l == 0)
            __hip_atomic_store(status + tile, os_word(tile == 0 ? 2u : 1u, tag, agg),
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        uint32_t excl = 0;
        bool ok = true;
        int64_t k = (int64_t)tile - 1;  // lane l reads tile k - l
        uint32_t spins = 0;
        while (k >= 0) {
            const uint64_t w = k - l >= 0 ? __hip_atomic_load(status + (k - l), __ATOMIC_RELAXED,
                                                              __HIP_MEMORY_SCOPE_AGENT)
                                          : os_word(2u, tag, 0u);  // nothing before tile 0
            const uint32_t f = (uint32_t)(w >> 62), wt = (uint32_t)(w >> 32) & 0x3fffffffu;
            const bool ready = wt == tag && f != 0u;
            const uint64_t notready = __ballot(!ready), incl = __ballot(ready && f == 2u);
            const int first_nr = notready ? __ffsll((unsigned long long)notready) - 1 : 64;
            const int first_in = incl ? __ffsll((unsigned long long)incl) - 1 : 64;
            const int take = first_in < first_nr ? first_in + 1 : first_nr;  // lanes 0..take-1
            excl += wave_sum(l < take ? (uint32_t)w : 0u);
            if (first_in < first_nr) break;
            k -= take;
            if (take == 0) {
                if (++spins > kOsSpinLimit) {
                    ok = false;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        if (l == 0) {
            if (tile > 0)
                __hip_atomic_store(status + tile, os_word(2u, tag, excl + agg), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            if (!ok) atomicOr(&meta[META_FLAGS], (uint32_t)FLAG_SORT);  // never expected
            s_prefix = excl;
        }
    }
    __syncthreads();
    uint32_t run = s_prefix + ex;
    if (base + kScanItems <= count) {
        uint4* dst = reinterpret_cast<uint4*>(data + base);
#pragma unroll
        for (int r = 0; r < kScanItems / 4; r++) {
            uint4 x;
            x.x = run;
            run += v[4 * r];
            x.y = run;
            run += v[4 * r + 1];
            x.z = run;
            run += v[4 * r + 2];
            x.w = run;
            run += v[4 * r + 3];
            dst[r] = x;
        }
    } else {
#pragma unroll
        for (int r = 0; r < kScanItems; r++) {
            if (base + r < count) data[base + r] = run;
            run += v[r];
        }
    }
    if (tile == gridDim.x - 1 && threadIdx.x == 0) {
        if (total_out) *total_out = s_prefix + agg;
        *epoch += 1u;
    }
}

void launch_scan_u32(nb_ctx* c, uint32_t* data, uint32_t count, uint32_t* total_out) {
    const uint32_t tiles = (count + kScanTile - 1) / kScanTile;
    hipLaunchKernelGGL(k_scan, dim3(tiles ? tiles : 1), dim3(kScanThreads), 0, c->stream, data,
                       count, c->b.scan_status, c->b.sort_glob + kSortGlobScanEpoch, total_out,
                       c->b.meta);
}

}  // namespace nb
