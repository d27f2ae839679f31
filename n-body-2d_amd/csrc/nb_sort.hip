// nb_sort.hip — bounding box, Morton keys, LDS-staged LSD radix sort, gather.
//
// Semantics follow BarnesHutCuda.cu: the squared bounding box (:78-114,
// :648-680), 64-bit Morton keys with x in the odd bits (:54-69, :117-132), a
// stable ascending sort over all 64 key bits with the current index as payload
// (cub::DeviceRadixSort::SortPairs, :691-693) and the gather of every body
// array through the permutation (:135-147). The squaring runs on the GPU, so
// a step has no host round trip.
#include <float.h>

#include "nb_device.h"
#include "nb_internal.h"

namespace nb {

// ---------------------------------------------------------------------------
// Bounding box: grid-stride min/max, wave shuffles, one partial per block,
// then a single-block kernel squares the box about the shorter axis.
// HBM traffic: 16 B/body read.
// ---------------------------------------------------------------------------
constexpr int kBboxThreads = 256;

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
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        xmin = fmin(xmin, __shfl_xor(xmin, d, 64));
        xmax = fmax(xmax, __shfl_xor(xmax, d, 64));
        ymin = fmin(ymin, __shfl_xor(ymin, d, 64));
        ymax = fmax(ymax, __shfl_xor(ymax, d, 64));
    }
    __shared__ double s[4][kBboxThreads / 64];
    const int w = threadIdx.x >> 6;
    if (lane_id() == 0) {
        s[0][w] = xmin;
        s[1][w] = xmax;
        s[2][w] = ymin;
        s[3][w] = ymax;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < kBboxThreads / 64; k++) {
            xmin = fmin(xmin, s[0][k]);
            xmax = fmax(xmax, s[1][k]);
            ymin = fmin(ymin, s[2][k]);
            ymax = fmax(ymax, s[3][k]);
        }
        part[4 * blockIdx.x + 0] = xmin;
        part[4 * blockIdx.x + 1] = xmax;
        part[4 * blockIdx.x + 2] = ymin;
        part[4 * blockIdx.x + 3] = ymax;
    }
}

// Root size^2 of the opening test: side^2 (SIDE), or w^2 + h^2 = 2 side^2 on
// a square cell (DIAG; exact, as the reference's lengthSquared of (s, s)).
__device__ __forceinline__ double root_size_sq(double w, int diag) {
    const double s2 = w * w;
    return diag ? s2 + s2 : s2;
}

__global__ __launch_bounds__(64) void k_bbox_final(const double* __restrict__ part, uint32_t nparts,
                                                   double* __restrict__ box,
                                                   uint32_t* __restrict__ meta, int diag) {
    double xmin = INFINITY, xmax = -INFINITY, ymin = INFINITY, ymax = -INFINITY;
    for (uint32_t k = threadIdx.x; k < nparts; k += 64) {
        xmin = fmin(xmin, part[4 * k + 0]);
        xmax = fmax(xmax, part[4 * k + 1]);
        ymin = fmin(ymin, part[4 * k + 2]);
        ymax = fmax(ymax, part[4 * k + 3]);
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        xmin = fmin(xmin, __shfl_xor(xmin, d, 64));
        xmax = fmax(xmax, __shfl_xor(xmax, d, 64));
        ymin = fmin(ymin, __shfl_xor(ymin, d, 64));
        ymax = fmax(ymax, __shfl_xor(ymax, d, 64));
    }
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
                                                           int diag) {
    const uint32_t s = blockIdx.x;
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
                                                     uint32_t* __restrict__ meta) {
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
                           c->ref_blocks, c->b.box, c->b.meta);
        return;
    }
    if (c->nsys > 1) {
        hipLaunchKernelGGL(k_bbox_sys, dim3(c->nsys), dim3(kBboxThreads), 0, c->stream,
                           c->b.pos[src], c->b.sys_off, c->b.box, c->b.meta, diag);
        return;
    }
    hipLaunchKernelGGL(k_bbox_partial, dim3(c->bbox_blocks), dim3(kBboxThreads), 0, c->stream,
                       c->b.pos[src], (uint32_t)c->n, c->b.bbox_part);
    hipLaunchKernelGGL(k_bbox_final, dim3(1), dim3(64), 0, c->stream, c->b.bbox_part,
                       c->bbox_blocks, c->b.box, c->b.meta, diag);
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

__global__ __launch_bounds__(256) void k_morton(const double2* __restrict__ pos, uint32_t n,
                                                const double* __restrict__ box,
                                                const uint32_t* __restrict__ sys_of,
                                                uint64_t* __restrict__ keys,
                                                uint32_t* __restrict__ vals) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const double* bx = box + (uint64_t)kBoxStride * sys_of[i];
    const double x_min = bx[0], x_max = bx[1], y_min = bx[2], y_max = bx[3];
    const double inv_w = 1.0 / (x_max - x_min);
    const double inv_h = 1.0 / (y_max - y_min);
    const double2 p = pos[i];
    const uint32_t ix = quantise(p.x, x_min, inv_w);
    const uint32_t iy = quantise(p.y, y_min, inv_h);
    keys[i] = (spread_bits(ix) << 1) | spread_bits(iy);
    vals[i] = i;
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
__global__ __launch_bounds__(256) void k_morton_ref(const double2* __restrict__ pos, uint32_t n,
                                                    const double* __restrict__ box,
                                                    uint64_t* __restrict__ keys,
                                                    uint32_t* __restrict__ vals) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    double px = box[0], py = box[2];
    double sx = box[1] - box[0], sy = box[3] - box[2];
    const double2 p = pos[i];
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
    keys[i] = key;
    vals[i] = i;
}

void launch_morton(nb_ctx* c, int src) {
    const uint32_t n = (uint32_t)c->n;
    if (c->p.cells == NB_CELLS_REF_CPU) {
        hipLaunchKernelGGL(k_morton_ref, dim3((n + 255) / 256), dim3(256), 0, c->stream,
                           c->b.pos[src], n, c->b.box, c->b.keys[0], c->b.vals[0]);
        return;
    }
    hipLaunchKernelGGL(k_morton, dim3((n + 255) / 256), dim3(256), 0, c->stream, c->b.pos[src], n,
                       c->b.box, c->b.sys_of, c->b.keys[0], c->b.vals[0]);
}

// ---------------------------------------------------------------------------
// LSD radix sort, 8-bit digits, 8 passes (keys ping-pong keys[0] -> keys[1]
// -> ... -> keys[0]). Per pass:
//   hist    per-tile digit histogram (wave-aggregated LDS adds)
//   scan    one workgroup per digit scans that digit's column across tiles
//   scatter stable rank inside the tile (8 ballots per key give the lanes of
//           the wave holding the same digit), then a coalesced-by-digit store
// HBM traffic per pass: 8 B (hist) + 12 B read + 12 B written per key.
// ---------------------------------------------------------------------------

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

// SYS: the digit is taken from the system of the key's source position
// (sys_of[val]) instead of the key, for the passes that keep systems apart.
template <bool SYS>
__device__ __forceinline__ uint32_t sort_digit(uint64_t key, uint32_t val, int shift,
                                               const uint32_t* __restrict__ sys_of) {
    if (SYS) return (sys_of[val] >> shift) & (kSortBins - 1);
    return (uint32_t)(key >> shift) & (kSortBins - 1);
}

template <bool SYS>
__global__ __launch_bounds__(kSortThreads) void k_rs_hist(const uint64_t* __restrict__ keys,
                                                          const uint32_t* __restrict__ vals,
                                                          const uint32_t* __restrict__ sys_of,
                                                          uint32_t n, int shift,
                                                          uint32_t* __restrict__ hist,
                                                          uint32_t tiles) {
    __shared__ uint32_t h[kSortBins];
    h[threadIdx.x] = 0;
    __syncthreads();
    const uint64_t base = (uint64_t)blockIdx.x * kSortTile;
#pragma unroll 4
    for (int r = 0; r < kSortItems; r++) {
        const uint64_t idx = base + (uint64_t)r * kSortThreads + threadIdx.x;
        const bool valid = idx < n;
        const uint32_t d =
            valid ? sort_digit<SYS>(SYS ? 0ull : keys[idx], SYS ? vals[idx] : 0u, shift, sys_of)
                  : 0u;
        const uint64_t peers = digit_peers(d, valid);
        if (valid && (peers & lanemask_lt()) == 0) atomicAdd(&h[d], (uint32_t)__popcll(peers));
    }
    __syncthreads();
    hist[threadIdx.x * tiles + blockIdx.x] = h[threadIdx.x];
}

__global__ __launch_bounds__(kSortThreads) void k_rs_scan(uint32_t* __restrict__ hist,
                                                          uint32_t tiles,
                                                          uint32_t* __restrict__ tot) {
    __shared__ uint32_t s[kSortThreads / 64 + 1];
    uint32_t* col = hist + (uint64_t)blockIdx.x * tiles;
    uint32_t carry = 0;
    for (uint32_t b0 = 0; b0 < tiles; b0 += kSortThreads) {
        const uint32_t k = b0 + threadIdx.x;
        const uint32_t v = k < tiles ? col[k] : 0u;
        uint32_t sum;
        const uint32_t ex = block_excl_scan<kSortThreads>(v, s, &sum);
        if (k < tiles) col[k] = carry + ex;
        carry += sum;
    }
    if (threadIdx.x == 0) tot[blockIdx.x] = carry;
}

template <bool SYS>
__global__ __launch_bounds__(kSortThreads) void k_rs_scatter(
    const uint64_t* __restrict__ kin, const uint32_t* __restrict__ vin, uint64_t* __restrict__ kout,
    uint32_t* __restrict__ vout, const uint32_t* __restrict__ sys_of, uint32_t n, int shift,
    const uint32_t* __restrict__ hist, const uint32_t* __restrict__ tot, uint32_t tiles) {
    constexpr int NW = kSortThreads / 64;
    __shared__ uint32_t run[kSortBins];      // next output position of each digit
    __shared__ uint32_t wcnt[NW][kSortBins]; // keys of each digit per wave, this round
    __shared__ uint32_t s[NW + 1];
    const int t = threadIdx.x, w = t >> 6;
    {
        uint32_t sum;
        const uint32_t ex = block_excl_scan<kSortThreads>(tot[t], s, &sum);
        run[t] = ex + hist[(uint64_t)t * tiles + blockIdx.x];
#pragma unroll
        for (int k = 0; k < NW; k++) wcnt[k][t] = 0;
    }
    const uint64_t base = (uint64_t)blockIdx.x * kSortTile;
    uint64_t key[kSortItems];
    uint32_t val[kSortItems];
#pragma unroll
    for (int r = 0; r < kSortItems; r++) {
        const uint64_t idx = base + (uint64_t)r * kSortThreads + t;
        key[r] = idx < n ? kin[idx] : 0ull;
        val[r] = idx < n ? vin[idx] : 0u;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kSortItems; r++) {
        const uint64_t idx = base + (uint64_t)r * kSortThreads + t;
        const bool valid = idx < n;
        const uint32_t d = valid ? sort_digit<SYS>(key[r], val[r], shift, sys_of) : 0u;
        const uint64_t peers = digit_peers(d, valid);
        const uint64_t below = peers & lanemask_lt();
        if (valid && below == 0) wcnt[w][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t dst = run[d] + (uint32_t)__popcll(below);
            for (int k = 0; k < w; k++) dst += wcnt[k][d];
            kout[dst] = key[r];
            vout[dst] = val[r];
        }
        __syncthreads();
        uint32_t add = 0;
#pragma unroll
        for (int k = 0; k < NW; k++) {
            add += wcnt[k][t];
            wcnt[k][t] = 0;
        }
        run[t] += add;
        __syncthreads();
    }
}

// Several systems: two more stable passes on the 16-bit system number (of
// the key's source position) put the systems back in place, each sorted by
// its own keys. An even pass count leaves the result in keys[0] / vals[0].
void launch_radix_sort(nb_ctx* c) {
    const uint32_t n = (uint32_t)c->n;
    const uint32_t tiles = c->sort_tiles;
    const int key_passes = 64 / kSortBits, sys_passes = c->nsys > 1 ? 2 : 0;
    for (int pass = 0; pass < key_passes + sys_passes; pass++) {
        const bool sys = pass >= key_passes;
        const int shift = (sys ? pass - key_passes : pass) * kSortBits;
        const int in = pass & 1, out = in ^ 1;
        if (sys)
            hipLaunchKernelGGL(k_rs_hist<true>, dim3(tiles), dim3(kSortThreads), 0, c->stream,
                               c->b.keys[in], c->b.vals[in], c->b.sys_of, n, shift,
                               c->b.sort_hist, tiles);
        else
            hipLaunchKernelGGL(k_rs_hist<false>, dim3(tiles), dim3(kSortThreads), 0, c->stream,
                               c->b.keys[in], c->b.vals[in], c->b.sys_of, n, shift,
                               c->b.sort_hist, tiles);
        hipLaunchKernelGGL(k_rs_scan, dim3(kSortBins), dim3(kSortThreads), 0, c->stream,
                           c->b.sort_hist, tiles, c->b.sort_tot);
        if (sys)
            hipLaunchKernelGGL(k_rs_scatter<true>, dim3(tiles), dim3(kSortThreads), 0, c->stream,
                               c->b.keys[in], c->b.vals[in], c->b.keys[out], c->b.vals[out],
                               c->b.sys_of, n, shift, c->b.sort_hist, c->b.sort_tot, tiles);
        else
            hipLaunchKernelGGL(k_rs_scatter<false>, dim3(tiles), dim3(kSortThreads), 0,
                               c->stream, c->b.keys[in], c->b.vals[in], c->b.keys[out],
                               c->b.vals[out], c->b.sys_of, n, shift, c->b.sort_hist,
                               c->b.sort_tot, tiles);
    }
}

// ---------------------------------------------------------------------------
// Gather every body array through the permutation (BarnesHutCuda.cu:135-147).
// HBM traffic: 4 B perm + 44 B read + 44 B written per body.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_gather(
    const uint32_t* __restrict__ perm, const double2* __restrict__ pin,
    const double2* __restrict__ vin, const double* __restrict__ min_,
    const uint32_t* __restrict__ oin, double2* __restrict__ pout, double2* __restrict__ vout,
    double* __restrict__ mout, uint32_t* __restrict__ oout, uint32_t n) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const uint32_t s = perm[i];
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
                       c->b.pos[dst], c->b.vel[dst], c->b.mass[dst], c->b.orig[dst], n);
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

// ---------------------------------------------------------------------------
// Exclusive in-place scan of u32 (reduce / top-level scan / downsweep).
// ---------------------------------------------------------------------------
constexpr int kScanThreads = 256;
constexpr int kScanItems = 4;
constexpr int kScanTile = kScanThreads * kScanItems;

__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const uint32_t* __restrict__ data,
                                                              uint32_t count,
                                                              uint32_t* __restrict__ sums) {
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
    uint32_t v = 0;
#pragma unroll
    for (int r = 0; r < kScanItems; r++) {
        const uint64_t idx = base + (uint64_t)r * kScanThreads + threadIdx.x;
        if (idx < count) v += data[idx];
    }
    v = wave_sum(v);
    __shared__ uint32_t s[kScanThreads / 64];
    if (lane_id() == 0) s[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int k = 0; k < kScanThreads / 64; k++) t += s[k];
        sums[blockIdx.x] = t;
    }
}

__global__ __launch_bounds__(1024) void k_scan_top(uint32_t* __restrict__ sums, uint32_t nb,
                                                   uint32_t* __restrict__ total) {
    __shared__ uint32_t s[1024 / 64 + 1];
    uint32_t carry = 0;
    for (uint32_t b0 = 0; b0 < nb; b0 += 1024) {
        const uint32_t k = b0 + threadIdx.x;
        const uint32_t v = k < nb ? sums[k] : 0u;
        uint32_t sum;
        const uint32_t ex = block_excl_scan<1024>(v, s, &sum);
        if (k < nb) sums[k] = carry + ex;
        carry += sum;
    }
    if (threadIdx.x == 0) *total = carry;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_down(uint32_t* __restrict__ data,
                                                            uint32_t count,
                                                            const uint32_t* __restrict__ sums) {
    __shared__ uint32_t s[kScanThreads / 64 + 1];
    // Blocked layout: thread t owns items t*kScanItems .. +kScanItems-1.
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint32_t tsum = 0;
#pragma unroll
    for (int r = 0; r < kScanItems; r++) {
        v[r] = (base + r) < count ? data[base + r] : 0u;
        tsum += v[r];
    }
    uint32_t total;
    uint32_t run = sums[blockIdx.x] + block_excl_scan<kScanThreads>(tsum, s, &total);
#pragma unroll
    for (int r = 0; r < kScanItems; r++) {
        if ((base + r) < count) data[base + r] = run;
        run += v[r];
    }
}

void launch_scan_u32(nb_ctx* c, uint32_t* data, uint32_t count, uint32_t* total_out) {
    const uint32_t nb = (count + kScanTile - 1) / kScanTile;
    const uint32_t g = nb ? nb : 1;
    hipLaunchKernelGGL(k_scan_reduce, dim3(g), dim3(kScanThreads), 0, c->stream, data, count,
                       c->b.scan_tmp);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, c->stream, c->b.scan_tmp, g,
                       total_out ? total_out : c->b.scan_tmp + g);
    hipLaunchKernelGGL(k_scan_down, dim3(g), dim3(kScanThreads), 0, c->stream, data, count,
                       c->b.scan_tmp);
}

}  // namespace nb
