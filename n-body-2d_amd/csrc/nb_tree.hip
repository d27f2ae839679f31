// nb_tree.hip — bottom-up quadtree build from sorted Morton keys, then a
// level-synchronous mass / centre-of-mass / second-moment reduction.
//
// Same tree as the reference's level-by-level build (BarnesHutCuda.cu:153-240,
// 713-777): a node at level L is a 2L-bit key prefix; it is internal when it
// holds >= 2 bodies and L < 30, and an internal node owns four child slots;
// its occupancy mask marks the non-empty ones, and every reader reaches a
// child through that mask, so empty slots are never written (40 % of the
// slots of a disk: the tree's largest write). The reference discovers the
// tree top-down with one launch and one host round trip per level; here it
// is read directly off the sorted keys:
//   delta_i = number of leading equal base-4 digits of keys i and i+1 (<= 30)
//   pair i owns the internal nodes at levels (delta_{i-1}, min(delta_i, 29)]
//   (those are exactly the internal nodes whose first body is i)
// so an exclusive scan of the per-pair counts numbers every internal node in
// pre-order (sorted by first body, then level), and internal node p owns child
// slots S+4p..S+4p+3 (S = number of systems; slot s is the root of system s).
// No atomics, no host synchronisation, deterministic ids. With several
// systems, delta = -1 between the last body of one system and the first of
// the next, so every system gets its own root and its own subtree.
#include "nb_device.h"
#include "nb_internal.h"

namespace nb {


// Writes a whole 64-byte record with four 16-byte stores, so every record
// reaches memory as one full line (a record written in parts costs HBM
// read-modify-writes of partial lines).
__device__ __forceinline__ void store_rec(NodeRec* __restrict__ r, double cx, double cy, double m,
                                          double wsq, int32_t c0, uint32_t mask, uint32_t bs,
                                          uint32_t be, float ixx, float ixy, float iyy) {
    uint4* w = reinterpret_cast<uint4*>(r);
    w[0] = make_uint4(__double2loint(cx), __double2hiint(cx), __double2loint(cy),
                      __double2hiint(cy));
    w[1] = make_uint4(__double2loint(m), __double2hiint(m), __double2loint(wsq),
                      __double2hiint(wsq));
    w[2] = make_uint4((uint32_t)c0, mask, bs, be);
    w[3] = make_uint4(__float_as_uint(ixx), __float_as_uint(ixy), __float_as_uint(iyy),
                      __float_as_uint((float)m));
}

// Builds internal node p = (first body a, level L): its body range, the four
// child ranges (binary searches on the digit, lower_bound_quadrant of
// BarnesHutCuda.cu:153-163) and the child slots. Leaves get their whole
// record here; multi-body leaves (only possible at level 30) are queued for
// leaf_multi (run by the k_level_hist launch). An internal child only learns
// its slot (inode_slot): its record is written once, whole, by the COM
// reduction, from the range and wsq that its own thread keeps in inode_info.
__device__ void build_node(uint32_t p, const uint64_t* __restrict__ keys, uint32_t nsys,
                           const uint32_t* __restrict__ sys_of,
                           const uint32_t* __restrict__ sys_off,
                           const double2* __restrict__ pos, const double* __restrict__ mass,
                           const int8_t* __restrict__ delta, const uint32_t* __restrict__ off,
                           const double* __restrict__ box, const uint32_t* __restrict__ node_a,
                           NodeRec* __restrict__ nodes, uint32_t* __restrict__ inode_slot,
                           const uint8_t* __restrict__ inode_level,
                           uint8_t* __restrict__ inode_mask, uint4* __restrict__ inode_info,
                           uint32_t* __restrict__ multi_list, uint32_t* __restrict__ meta,
                           float4* __restrict__ mom3) {
    const uint32_t a = node_a[p];
    const int L = inode_level[p];
    const uint32_t sys = sys_of[a];
    const uint32_t n = sys_off[sys + 1];  // end of this node's system
    const double root_wsq = box[(uint64_t)kBoxStride * sys + 4];
    uint32_t b;
    if (L == 0) {
        b = n - 1;
    } else {
        const int sh = 64 - 2 * L;
        const uint64_t P = keys[a] >> sh;
        uint32_t lo = a + 1, step = 1;  // keys[a+1] shares the prefix (delta_a >= L)
        for (;;) {
            const uint64_t probe = (uint64_t)lo + step;
            if (probe >= n || (keys[probe] >> sh) != P) break;
            lo = (uint32_t)probe;
            step <<= 1;
        }
        uint32_t hi = (uint64_t)lo + step < n ? lo + step : n;
        while (hi - lo > 1) {
            const uint32_t mid = lo + (hi - lo) / 2;
            if ((keys[mid] >> sh) == P)
                lo = mid;
            else
                hi = mid;
        }
        b = lo;
    }
    const int s = 62 - 2 * L;
    uint32_t bq[5];
    bq[0] = a;
    bq[4] = b + 1;
#pragma unroll
    for (int q = 1; q < 4; q++) {
        uint32_t lo = bq[q - 1], hi = b + 1;
        while (lo < hi) {
            const uint32_t mid = lo + (hi - lo) / 2;
            if (((keys[mid] >> s) & 3ull) < (uint64_t)q)
                lo = mid + 1;
            else
                hi = mid;
        }
        bq[q] = lo;
    }
    const double cwsq = ldexp(root_wsq, -2 * (L + 1));
    uint32_t mask = 0;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const uint32_t cs = bq[q], ce1 = bq[q + 1];
        NodeRec* r = &nodes[nsys + 4ull * p + q];
        if (cs >= ce1) continue;  // empty: bit q of the mask stays clear
        mask |= 1u << q;
        const uint32_t ce = ce1 - 1;
        if (cs == ce) {
            const double2 x = pos[cs];
            store_rec(r, x.x, x.y, mass[cs], cwsq, -1, 0u, cs, ce, 0.f, 0.f, 0.f);
            if (mom3) mom3[nsys + 4ull * p + q] = make_float4(0.f, 0.f, 0.f, 0.f);
        } else if (L + 1 == kMaxLevel) {
            store_rec(r, 0.0, 0.0, 0.0, cwsq, -1, 0u, cs, ce, 0.f, 0.f, 0.f);
            const uint32_t k = atomicAdd(&meta[META_MULTI_LEAVES], 1u);
            multi_list[k] = (uint32_t)(nsys + 4ull * p + q);
        } else {
            const int prev = cs ? (int)delta[cs - 1] : -1;
            const uint32_t cid = off[cs] + (uint32_t)(L + 1 - prev - 1);
            inode_slot[cid] = (uint32_t)(nsys + 4ull * p + q);
        }
    }
    inode_mask[p] = (uint8_t)mask;
    const double wsq = ldexp(root_wsq, -2 * L);
    inode_info[p] = make_uint4(a, b, __double2loint(wsq), __double2hiint(wsq));
    if (L == 0) inode_slot[p] = sys;  // the root of system `sys` lives in slot `sys`
}

// Lists every internal node: pair i owns levels (delta_{i-1}, min(delta_i, 29)],
// internal node ids off[i] + (L - delta_{i-1} - 1).
__global__ __launch_bounds__(256) void k_tree_enum(const int8_t* __restrict__ delta,
                                                   const uint32_t* __restrict__ off, uint32_t n,
                                                   const uint32_t* __restrict__ sys_of,
                                                   const uint32_t* __restrict__ sys_off,
                                                   uint32_t* __restrict__ meta, uint64_t icap,
                                                   uint32_t* __restrict__ node_a,
                                                   uint8_t* __restrict__ inode_level,
                                                   const double2* __restrict__ pos,
                                                   const double* __restrict__ mass,
                                                   const double* __restrict__ box,
                                                   NodeRec* __restrict__ nodes,
                                                   float4* __restrict__ mom3) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    const uint32_t I = meta[META_INTERNAL];
    // A step after one that halted (any halt flag) only carries the state
    // along: its tree is not built. The node count stays a real one (the
    // keys are still sorted), so it can size the node pool.
    if (I > icap || (meta[META_FLAGS] & kHaltFlags)) {
        if (i == 0 && I > icap) atomicOr(&meta[META_FLAGS], (uint32_t)FLAG_TREE_CAPACITY);
        return;
    }
    if (i >= n) return;
    const uint32_t sys = sys_of[i];
    if (sys_off[sys] == i && sys_off[sys + 1] == i + 1) {
        // a system of one body: its root is a leaf
        NodeRec* r = &nodes[sys];
        const double2 x = pos[i];
        r->cx = x.x;
        r->cy = x.y;
        r->m = mass[i];
        r->mf = (float)mass[i];
        r->wsq = box[(uint64_t)kBoxStride * sys + 4];
        r->c0 = -1;
        r->mask = 0;
        r->bs = (int32_t)i;
        r->be = (int32_t)i;
        r->ixx = r->ixy = r->iyy = 0.f;
        if (mom3) mom3[sys] = make_float4(0.f, 0.f, 0.f, 0.f);
        return;
    }
    if (i + 1 >= n) return;
    const int prev = i ? (int)delta[i - 1] : -1;
    const int d = delta[i];
    const int hi = d < kMaxLevel - 1 ? d : kMaxLevel - 1;
    const uint32_t o = off[i];
    for (int L = prev + 1; L <= hi; L++) {
        const uint32_t p = o + (uint32_t)(L - prev - 1);
        node_a[p] = i;
        inode_level[p] = (uint8_t)L;
    }
}

__global__ __launch_bounds__(256) void k_tree_nodes(
    const uint64_t* __restrict__ keys, uint32_t nsys, const uint32_t* __restrict__ sys_of,
    const uint32_t* __restrict__ sys_off, const double2* __restrict__ pos,
    const double* __restrict__ mass, const int8_t* __restrict__ delta,
    const uint32_t* __restrict__ off, const double* __restrict__ box, uint32_t* __restrict__ meta,
    uint64_t icap, const uint32_t* __restrict__ node_a, NodeRec* __restrict__ nodes,
    uint32_t* __restrict__ inode_slot, const uint8_t* __restrict__ inode_level,
    uint8_t* __restrict__ inode_mask, uint4* __restrict__ inode_info,
    uint32_t* __restrict__ multi_list, float4* __restrict__ mom3) {
    const uint32_t I = meta[META_INTERNAL];
    if (I > icap || (meta[META_FLAGS] & kHaltFlags)) return;
    const uint32_t p = blockIdx.x * 256 + threadIdx.x;
    if (p >= I) return;
    build_node(p, keys, nsys, sys_of, sys_off, pos, mass, delta, off, box, node_a, nodes,
               inode_slot, inode_level, inode_mask, inode_info, multi_list, meta, mom3);
}

// Multi-body leaf COM, one wave per leaf, in the order of compute_leaf_com
// (BarnesHutCuda.cu:277-309): 128 strided partial sums (lane l holds slots l
// and l+64), then the halving tree reduction. Second moments about the COM.
// Workgroup `blk` of `nblk` (256 threads each).
__device__ __forceinline__ void leaf_multi(uint32_t blk, uint32_t nblk,
                                           const double2* __restrict__ pos,
                                           const double* __restrict__ mass,
                                           const uint32_t* __restrict__ multi_list,
                                           const uint32_t* __restrict__ meta,
                                           NodeRec* __restrict__ nodes,
                                           float4* __restrict__ mom3) {
    const uint32_t count = meta[META_MULTI_LEAVES];
    const uint32_t waves = nblk * 4;
    const int l = lane_id();
    for (uint32_t k = blk * 4 + (threadIdx.x >> 6); k < count; k += waves) {
        NodeRec* r = &nodes[multi_list[k]];
        const int32_t bs = r->bs, be = r->be;
        double M[2] = {0.0, 0.0}, X[2] = {0.0, 0.0}, Y[2] = {0.0, 0.0};
#pragma unroll
        for (int h = 0; h < 2; h++) {
            for (int32_t i = bs + l + 64 * h; i <= be; i += 128) {
                const double m = mass[i];
                const double2 x = pos[i];
                M[h] += m;
                X[h] = fma(m, x.x, X[h]);
                Y[h] = fma(m, x.y, Y[h]);
            }
        }
        double m = M[0] + M[1], x = X[0] + X[1], y = Y[0] + Y[1];  // stride 64
#pragma unroll
        for (int st = 32; st >= 1; st >>= 1) {
            const double om = __shfl_down(m, st, 64), ox = __shfl_down(x, st, 64),
                         oy = __shfl_down(y, st, 64);
            if (l < st) {
                m += om;
                x += ox;
                y += oy;
            }
        }
        m = __shfl(m, 0, 64);
        const double cx = __shfl(x, 0, 64) / m, cy = __shfl(y, 0, 64) / m;
        double a = 0.0, b = 0.0, c = 0.0, t0 = 0.0, t1 = 0.0, t2 = 0.0, t3 = 0.0;
        for (int32_t i = bs + l; i <= be; i += 64) {
            const double2 p = pos[i];
            const double w = mass[i], dx = p.x - cx, dy = p.y - cy;
            a += w * dx * dx;
            b += w * dx * dy;
            c += w * dy * dy;
            t0 += w * dx * dx * dx;
            t1 += w * dx * dx * dy;
            t2 += w * dx * dy * dy;
            t3 += w * dy * dy * dy;
        }
#pragma unroll
        for (int st = 32; st >= 1; st >>= 1) {
            a += __shfl_xor(a, st, 64);
            b += __shfl_xor(b, st, 64);
            c += __shfl_xor(c, st, 64);
            t0 += __shfl_xor(t0, st, 64);
            t1 += __shfl_xor(t1, st, 64);
            t2 += __shfl_xor(t2, st, 64);
            t3 += __shfl_xor(t3, st, 64);
        }
        if (l == 0 && mom3)
            mom3[multi_list[k]] = make_float4((float)t0, (float)t1, (float)t2, (float)t3);
        if (l == 0) {
            r->cx = cx;
            r->cy = cy;
            r->m = m;
            r->mf = (float)m;
            r->ixx = (float)a;
            r->ixy = (float)b;
            r->iyy = (float)c;
        }
    }
}

// Mass / COM / second moments of an internal node from its four child slots
// in slot order q = 0..3: M += m, R = fma(m, com, R) (propagate_com_up,
// BarnesHutCuda.cu:315-344, with the multiply-add nvcc contracts), com = R/M;
// moments by the parallel-axis theorem. Empty slots hold m = 0 and com = 0,
// so summing all four adds exact zeros -- the result equals the reference's
// sum over the non-empty children -- and the four slot loads are independent.
struct Com {
    double m, cx, cy;
    float ixx, ixy, iyy;
};

__device__ __forceinline__ Com load_com(const NodeRec* __restrict__ r) {
    Com c;
    c.m = r->m;
    c.cx = r->cx;
    c.cy = r->cy;
    c.ixx = r->ixx;
    c.ixy = r->ixy;
    c.iyy = r->iyy;
    return c;
}

__device__ __forceinline__ Com combine(const Com ch[4]) {
    double M = 0.0, Rx = 0.0, Ry = 0.0;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        M += ch[q].m;
        Rx = fma(ch[q].m, ch[q].cx, Rx);
        Ry = fma(ch[q].m, ch[q].cy, Ry);
    }
    Com o;
    o.m = M;
    o.cx = Rx / M;
    o.cy = Ry / M;
    double a = 0.0, b = 0.0, c = 0.0;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const double m = ch[q].m, dx = ch[q].cx - o.cx, dy = ch[q].cy - o.cy;
        a += (double)ch[q].ixx + m * dx * dx;
        b += (double)ch[q].ixy + m * dx * dy;
        c += (double)ch[q].iyy + m * dy * dy;
    }
    o.ixx = (float)a;
    o.ixy = (float)b;
    o.iyy = (float)c;
    return o;
}

// The reference CPU quadtree's order (Quadtree.cpp:121-128): mass summed
// TL + TR + BL + BR, the COM numerator TL, BR, BL, TR as products and sums
// without contraction, then divided by the mass. Child slot q = right << 1 |
// bottom, so TL, TR, BL, BR are q = 0, 2, 1, 3.
__device__ __forceinline__ Com combine_ref(const Com ch[4]) {
#pragma clang fp contract(off)
    const Com &tl = ch[0], &tr = ch[2], &bl = ch[1], &br = ch[3];
    Com o = combine(ch);  // second moments (the reference has none)
    const double M = tl.m + tr.m + bl.m + br.m;
    const double sx = tl.cx * tl.m + br.cx * br.m + bl.cx * bl.m + tr.cx * tr.m;
    const double sy = tl.cy * tl.m + br.cy * br.m + bl.cy * bl.m + tr.cy * tr.m;
    o.m = M;
    o.cx = sx / M;
    o.cy = sy / M;
    return o;
}

// Third moments about the parent's COM from the children's (parallel axes):
// with d = child COM - parent COM, m, I, T the child's mass, second and
// third moments, T_abc += T_abc + d_a I_bc + d_b I_ac + d_c I_ab + m d_a d_b d_c.
__device__ __forceinline__ float4 combine3(const Com ch[4], const float4 t[4], const Com& o) {
    double txxx = 0.0, txxy = 0.0, txyy = 0.0, tyyy = 0.0;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const double m = ch[q].m, dx = ch[q].cx - o.cx, dy = ch[q].cy - o.cy;
        if (m == 0.0) continue;  // an empty slot
        const double ixx = ch[q].ixx, ixy = ch[q].ixy, iyy = ch[q].iyy;
        txxx += (double)t[q].x + 3.0 * dx * ixx + m * dx * dx * dx;
        txxy += (double)t[q].y + 2.0 * dx * ixy + dy * ixx + m * dx * dx * dy;
        txyy += (double)t[q].z + dx * iyy + 2.0 * dy * ixy + m * dx * dy * dy;
        tyyy += (double)t[q].w + 3.0 * dy * iyy + m * dy * dy * dy;
    }
    return make_float4((float)txxx, (float)txxy, (float)txyy, (float)tyyy);
}

__device__ __forceinline__ bool tree_halted(const uint32_t* meta, uint64_t icap) {
    return meta[META_INTERNAL] > icap || (meta[META_FLAGS] & kHaltFlags);
}

// Counting sort of the internal nodes by level, so each level of the
// bottom-up reduction touches only its own nodes: per-block level histograms
// (LDS atomics), one exclusive scan over the level-major counts, then a
// scatter. lvl_cnt has kMaxLevel rows of `nblk` counts. The same launch runs
// the multi-body leaves (independent work) in kMultiBlocks more workgroups.
constexpr uint32_t kMultiBlocks = 64;
__global__ __launch_bounds__(256) void k_level_hist(const uint32_t* __restrict__ meta,
                                                    uint64_t icap, uint32_t nblk,
                                                    const uint8_t* __restrict__ inode_level,
                                                    uint32_t* __restrict__ lvl_cnt,
                                                    const double2* __restrict__ pos,
                                                    const double* __restrict__ mass,
                                                    const uint32_t* __restrict__ multi_list,
                                                    NodeRec* __restrict__ nodes,
                                                    float4* __restrict__ mom3) {
    if (blockIdx.x >= nblk) {
        leaf_multi(blockIdx.x - nblk, kMultiBlocks, pos, mass, multi_list, meta, nodes, mom3);
        return;
    }
    __shared__ uint32_t h[kMaxLevel];
    if (threadIdx.x < kMaxLevel) h[threadIdx.x] = 0;
    __syncthreads();
    const bool ok = !tree_halted(meta, icap);
    const uint32_t I = meta[META_INTERNAL];
    const uint32_t p = blockIdx.x * 256 + threadIdx.x;
    if (ok && p < I) atomicAdd(&h[inode_level[p]], 1u);
    __syncthreads();
    if (threadIdx.x < kMaxLevel) lvl_cnt[threadIdx.x * nblk + blockIdx.x] = h[threadIdx.x];
}

// Each entry carries what the level's reduction needs besides the child
// records (node id, own slot, occupancy mask), so those are one load away.
__global__ __launch_bounds__(256) void k_level_scatter(const uint32_t* __restrict__ meta,
                                                       uint64_t icap,
                                                       const uint8_t* __restrict__ inode_level,
                                                       const uint8_t* __restrict__ inode_mask,
                                                       const uint32_t* __restrict__ inode_slot,
                                                       const uint32_t* __restrict__ lvl_off,
                                                       uint4* __restrict__ by_level) {
    __shared__ uint32_t h[kMaxLevel];
    if (threadIdx.x < kMaxLevel) h[threadIdx.x] = 0;
    __syncthreads();
    if (tree_halted(meta, icap)) return;
    const uint32_t I = meta[META_INTERNAL];
    const uint32_t p = blockIdx.x * 256 + threadIdx.x;
    if (p >= I) return;
    const uint32_t L = inode_level[p];
    const uint32_t r = atomicAdd(&h[L], 1u);
    by_level[lvl_off[L * gridDim.x + blockIdx.x] + r] =
        make_uint4(p, inode_slot[p], (uint32_t)inode_mask[p], 0u);
}

// One internal node of the reduction (by_level entry: node id, own slot,
// occupancy mask). An empty slot holds exact zeros: its load is skipped and
// combine() adds the same zeros; the third moments are loaded with the rest
// (one round trip). REF: the reference CPU order.
template <bool REF, bool M3>
__device__ __forceinline__ void com_node(const uint4 en, uint32_t nsys,
                                         const uint4* __restrict__ inode_info,
                                         NodeRec* __restrict__ nodes, float4* __restrict__ mom3) {
    const uint32_t p = en.x, mask = en.z;
    const uint4 info = inode_info[p];  // body range and wsq
    Com ch[4];
    float4 t[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const bool has = (mask >> q) & 1u;
        ch[q] = has ? load_com(&nodes[nsys + 4ull * p + q]) : Com{0.0, 0.0, 0.0, 0.f, 0.f, 0.f};
        if (M3) t[q] = has ? mom3[nsys + 4ull * p + q] : make_float4(0.f, 0.f, 0.f, 0.f);
    }
    const Com o = REF ? combine_ref(ch) : combine(ch);
    store_rec(&nodes[en.y], o.cx, o.cy, o.m, __hiloint2double((int)info.w, (int)info.z),
              (int32_t)(nsys + 4ull * p), mask, info.x, info.y, o.ixx, o.ixy, o.iyy);
    if (M3) mom3[en.y] = combine3(ch, t, o);
}

// The internal nodes of level L: segment [lvl_off[L * nblk], lvl_off[(L+1) * nblk])
// of by_level (the last level ends at I).
__device__ __forceinline__ uint32_t level_start(const uint32_t* __restrict__ lvl_off,
                                                const uint32_t* __restrict__ meta, uint32_t nblk,
                                                int L) {
    return L < kMaxLevel ? lvl_off[(uint64_t)L * nblk] : meta[META_INTERNAL];
}

template <bool REF, bool M3>
__global__ __launch_bounds__(256) void k_com_level(int L, const uint32_t* __restrict__ meta,
                                                   uint64_t icap, uint32_t nblk, uint32_t nsys,
                                                   const uint32_t* __restrict__ lvl_off,
                                                   const uint4* __restrict__ by_level,
                                                   const uint4* __restrict__ inode_info,
                                                   NodeRec* __restrict__ nodes,
                                                   float4* __restrict__ mom3) {
    // the segment bounds are loaded alongside the halt check (in bounds either way)
    const uint32_t b = level_start(lvl_off, meta, nblk, L);
    const uint32_t e = level_start(lvl_off, meta, nblk, L + 1);
    if (tree_halted(meta, icap)) return;
    for (uint32_t k = b + blockIdx.x * 256 + threadIdx.x; k < e; k += gridDim.x * 256)
        com_node<REF, M3>(by_level[k], nsys, inode_info, nodes, mom3);
}

// The deepest levels, from com_deep to 29, in one launch: one workgroup
// reduces them level after level with a barrier in between (a level's
// children are the nodes of the level below, written by this workgroup).
// com_deep is set a few levels below the deepest non-empty level of the
// context's trees, so these levels are empty or hold a few nodes (pairs of
// bodies closer than the deeper cells), and one launch replaces a launch per
// level. The launch records 1 + the deepest non-empty level and the most
// nodes of a level it took: a tree that grew so deep that this workgroup took
// a level of more than kComRefit nodes (still reduced, in more rounds) has
// the host pick com_deep again.
template <bool REF, bool M3>
__global__ __launch_bounds__(1024) void k_com_deep(int com_deep, uint32_t* __restrict__ meta,
                                                   uint64_t icap, uint32_t nblk, uint32_t nsys,
                                                   const uint32_t* __restrict__ lvl_off,
                                                   const uint4* __restrict__ by_level,
                                                   const uint4* __restrict__ inode_info,
                                                   NodeRec* __restrict__ nodes,
                                                   float4* __restrict__ mom3) {
    __shared__ uint32_t s_start[kMaxLevel + 1];
    if (tree_halted(meta, icap)) return;
    const int t = threadIdx.x;
    if (t <= kMaxLevel) s_start[t] = level_start(lvl_off, meta, nblk, t);
    __syncthreads();
    const int last = max(com_deep, 0);
    if (t == 0) {
        int deepest = 0;
        uint32_t most = 0;
        for (int L = 0; L < kMaxLevel; L++) {
            const uint32_t cnt = s_start[L + 1] - s_start[L];
            if (cnt) deepest = L + 1;
            if (L >= last) most = max(most, cnt);
        }
        atomicMax(&meta[META_COM_HI], (uint32_t)deepest);
        atomicMax(&meta[META_COM_DEEP_MAX], most);
    }
    for (int L = kMaxLevel - 1; L >= last; L--) {
        const uint32_t b = s_start[L], e = s_start[L + 1];
        for (uint32_t k = b + t; k < e; k += 1024)
            com_node<REF, M3>(by_level[k], nsys, inode_info, nodes, mom3);
        __syncthreads();
    }
}

void launch_tree_build(nb_ctx* c, int src) {
    const uint32_t n = (uint32_t)c->n;
    const uint32_t pairs = n > 1 ? n - 1 : 0;
    const uint32_t bg = (n + 255) / 256;
    // the pair deltas and counts were made by the gather (k_gather)
    launch_scan_u32(c, c->b.cnt, pairs, &c->b.meta[META_INTERNAL]);
    hipLaunchKernelGGL(k_tree_enum, dim3(bg), dim3(256), 0, c->stream, c->b.delta, c->b.cnt, n,
                       c->b.sys_of, c->b.sys_off, c->b.meta, (uint64_t)c->icap, c->b.node_a,
                       c->b.inode_level, c->b.pos[src], c->b.mass[src], c->b.box, c->b.nodes,
                       c->b.mom3);
    const uint32_t ng = (uint32_t)((c->icap + 255) / 256);
    hipLaunchKernelGGL(k_tree_nodes, dim3(ng), dim3(256), 0, c->stream, c->b.keys[0], c->nsys,
                       c->b.sys_of, c->b.sys_off, c->b.pos[src], c->b.mass[src], c->b.delta,
                       c->b.cnt, c->b.box, c->b.meta, (uint64_t)c->icap, c->b.node_a,
                       c->b.nodes, c->b.inode_slot, c->b.inode_level, c->b.inode_mask,
                       c->b.inode_info, c->b.multi_list, c->b.mom3);
    hipLaunchKernelGGL(k_level_hist, dim3(ng + kMultiBlocks), dim3(256), 0, c->stream, c->b.meta,
                       (uint64_t)c->icap, ng, c->b.inode_level, c->b.lvl_cnt, c->b.pos[src],
                       c->b.mass[src], c->b.multi_list, c->b.nodes, c->b.mom3);
    launch_scan_u32(c, c->b.lvl_cnt, (uint32_t)(kMaxLevel * ng), nullptr);
    hipLaunchKernelGGL(k_level_scatter, dim3(ng), dim3(256), 0, c->stream, c->b.meta,
                       (uint64_t)c->icap, c->b.inode_level, c->b.inode_mask, c->b.inode_slot,
                       c->b.lvl_cnt, c->b.by_level);
    // Deepest first (a level's children are all done when it starts): the
    // levels >= com_deep in one launch, then one launch per level below.
    const uint32_t cg = ng < 512 ? ng : 512;
    const bool ref = c->p.cells == NB_CELLS_REF_CPU;
    const bool m3 = c->b.mom3 != nullptr;
#define NB_COM(R, M)                                                                           \
    do {                                                                                       \
        if (L == kMaxLevel)                                                                    \
            hipLaunchKernelGGL((k_com_deep<R, M>), dim3(1), dim3(1024), 0, c->stream,          \
                               c->com_deep, c->b.meta, (uint64_t)c->icap, ng, c->nsys,         \
                               c->b.lvl_cnt, c->b.by_level, c->b.inode_info, c->b.nodes,       \
                               c->b.mom3);                                                     \
        else                                                                                   \
            hipLaunchKernelGGL((k_com_level<R, M>), dim3(cg), dim3(256), 0, c->stream, L,      \
                               c->b.meta, (uint64_t)c->icap, ng, c->nsys, c->b.lvl_cnt,        \
                               c->b.by_level, c->b.inode_info, c->b.nodes, c->b.mom3);         \
    } while (0)
    for (int L = kMaxLevel; L >= 0; L--) {
        if (L < kMaxLevel && L >= c->com_deep) continue;  // the one-workgroup launch's
        if (ref) {
            if (m3) NB_COM(true, true); else NB_COM(true, false);
        } else {
            if (m3) NB_COM(false, true); else NB_COM(false, false);
        }
    }
#undef NB_COM
}

}  // namespace nb
