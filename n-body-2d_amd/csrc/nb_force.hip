// nb_force.hip — force evaluation and integration kernels.
//
//   k_traverse    Barnes-Hut force, one wavefront per 64 Morton-adjacent
//                 bodies sharing one LDS node stack (compute_force,
//                 BarnesHutCuda.cu:382-486, re-designed for wave64)
//   k_allpairs    LDS-tiled direct sum (AllPairsSim, AllPairs.cpp:38-49)
//   k_drift       x += v dt (update_positions, BarnesHutCuda.cu:489-495)
//   k_drift_box   the same, plus the next step's box partials (one system)
//   k_direct64    FP64 direct sum, the reference for force errors
//
// Precision: the cell offset dx = com - x, its square d^2 and the opening
// test are evaluated in FP64 on the FP64 state with exactly the expressions
// of the reference path being reproduced, so every body accepts the same
// cells as that path's per-body FP64 DFS: SIDE (the CUDA path)
// d^2 = fma(dx, dx, dy*dy), wsq < theta^2 d^2; DIAG (the CPU path)
// d^2 = dx*dx + dy*dy, (w^2 + h^2) / d^2 < theta^2, where each node's wsq
// field holds its size^2 for the test in use. The interaction itself runs in
// FP32 and is summed per body with Kahan compensation; the kick v += a G dt
// is FP64. Force laws: PLUMMER m d (d^2 + eps^2)^{-3/2}; REF_CPU
// m (d / |d|) / (d^2 + eps^2) (Simulation.cpp:74-80), whose zero-distance
// terms (a body's own entry) are dropped.
#include "nb_device.h"
#include "nb_internal.h"

namespace nb {

__device__ __forceinline__ void kahan_add(float& sum, float& comp, float v) {
    const float y = v - comp;
    const float t = sum + y;
    comp = (t - sum) - y;
    sum = t;
}

struct Acc {
    float ax = 0.f, ay = 0.f, cx = 0.f, cy = 0.f;
    uint32_t n = 0;
};

// Magnitude factor s of one term s * d (FP32) for the squared distance d2:
// PLUMMER m / (d2 + eps^2)^{3/2}; REF_CPU m / (|d| (d2 + eps^2)), 0 at d2 = 0.
template <bool REFF>
__device__ __forceinline__ float term_scale(float d2, float m, float eps2f) {
    if (REFF) {
        const float s = m * __builtin_amdgcn_rsqf(d2) * __builtin_amdgcn_rcpf(d2 + eps2f);
        return d2 > 0.f ? s : 0.f;
    }
    const float rs = __builtin_amdgcn_rsqf(d2 + eps2f);
    return m * (rs * rs * rs);
}

// One monopole term of the force law.
template <bool REFF>
__device__ __forceinline__ void interact(Acc& a, double dx, double dy, double d2, float m,
                                         float eps2f) {
    const float s = term_scale<REFF>((float)d2, m, eps2f);
    kahan_add(a.ax, a.cx, (float)dx * s);
    kahan_add(a.ay, a.cy, (float)dy * s);
    a.n++;
}

// d^2 = dx*dx + dy*dy rounded as written, without a multiply-add: the
// reference CPU's lengthSquared (BarnesHut.cpp:113-114).
__device__ __forceinline__ double dist2_plain(double dx, double dy) {
#pragma clang fp contract(off)
    const double a = dx * dx;
    const double b = dy * dy;
    return a + b;
}

// Squared distance and opening test of one path. SIDE: the CUDA path's
// fma(dx, dx, dy*dy) and wsq < theta^2 d^2 (BarnesHutCuda.cu:413-415, 462);
// DIAG: the CPU path's plain d^2 and size^2 / d^2 < theta^2 (BarnesHut.cpp:113-115).
template <bool DIAG>
__device__ __forceinline__ double dist2(double dx, double dy) {
    return DIAG ? dist2_plain(dx, dy) : fma(dx, dx, dy * dy);
}

template <bool DIAG>
__device__ __forceinline__ bool opening_accepts(double size_sq, double d2, double theta_sq) {
    return DIAG ? (size_sq / d2 < theta_sq) : (size_sq < theta_sq * d2);
}

constexpr int kTravThreads = 256;
constexpr int kTravWaves = kTravThreads / 64;

// Wavefront-cooperative traversal. The stack entries are (slot, lane mask):
// the mask holds the lanes that still need the subtree. A popped internal
// node is tested by every lane in its mask; lanes that accept take its
// monopole, the others are pushed with its non-empty children (c0+3 .. c0,
// so c0 pops first, BarnesHutCuda.cu:474-481). Each lane therefore sees
// exactly the nodes, in exactly the order, of its own per-body DFS, while the
// wave fetches each node once (a wave-uniform scalar load).
// Masked walk of one wavefront (see above). s_node/s_mask: kStackDepth
// entries of LDS owned by this wave. Returns false on stack overflow.
template <bool GUARD, bool DIAG, bool REFF>
__device__ __forceinline__ bool exact_walk(const NodeRec* __restrict__ nodes,
                                           const double2* __restrict__ pos,
                                           const double* __restrict__ mass, double2 p,
                                           int32_t me, bool valid, double theta_sq,
                                           float eps2f, int32_t* s_node, uint64_t* s_mask,
                                           Acc& a) {
    const int l = lane_id();
    const uint64_t act = ballot64(valid);
    int sp = 1;
    s_node[0] = 0;
    s_mask[0] = act;
    while (sp > 0) {
        --sp;
        const int32_t node = readfirst(s_node[sp]);
        const uint64_t m = readfirst64(s_mask[sp]);
        const NodeRec* r = nodes + node;
        const double rcx = r->cx, rcy = r->cy;
        const int32_t c0 = r->c0, bs = r->bs, be = r->be;
        const bool mine = (m >> l) & 1ull;
        const double dx = rcx - p.x, dy = rcy - p.y;
        const double d2 = dist2<DIAG>(dx, dy);
        if (c0 < 0) {
            // Leaf (empty slots are never pushed). BarnesHutCuda.cu:421-458; a
            // single-body leaf other than the body itself never sits at its
            // position (equal positions share one leaf), so the index test is
            // also the CPU path's position test (BarnesHut.cpp:108).
            if (mine) {
                if (bs == be) {
                    if (bs != me) interact<REFF>(a, dx, dy, d2, r->mf, eps2f);
                } else if (me < bs || me > be) {
                    interact<REFF>(a, dx, dy, d2, r->mf, eps2f);
                } else {
                    for (int32_t j = bs; j <= be; j++) {
                        if (j == me) continue;
                        const double2 q = pos[j];
                        const double ex = q.x - p.x, ey = q.y - p.y;
                        interact<REFF>(a, ex, ey, dist2<DIAG>(ex, ey), (float)mass[j], eps2f);
                    }
                }
            }
        } else {
            bool accept = mine && opening_accepts<DIAG>(r->wsq, d2, theta_sq);
            if (GUARD) accept = accept && !(me >= bs && me <= be);
            if (accept) interact<REFF>(a, dx, dy, d2, r->mf, eps2f);
            const uint64_t open = ballot64(mine && !accept);
            if (open) {
                if (sp + 4 > kStackDepth) return false;
                const int32_t cm = r->mask;
#pragma unroll
                for (int q = 3; q >= 0; q--) {
                    if ((cm >> q) & 1) {
                        s_node[sp] = c0 + q;
                        s_mask[sp] = open;
                        sp++;
                    }
                }
            }
        }
    }
    return true;
}

template <bool KICK, bool GUARD, bool DIAG, bool REFF>
__global__ __launch_bounds__(kTravThreads) void k_traverse(
    const NodeRec* __restrict__ nodes, const double2* __restrict__ pos,
    const double* __restrict__ mass, double2* __restrict__ vel, double2* __restrict__ acc_out,
    uint32_t* __restrict__ inter_out, uint32_t n, const StepConsts* __restrict__ k,
    uint32_t* __restrict__ meta, int count_step) {
    if (meta[META_FLAGS] & kHaltFlags) return;
    __shared__ int32_t s_node[kTravWaves][kStackDepth];
    __shared__ uint64_t s_mask[kTravWaves][kStackDepth];
    const int w = threadIdx.x >> 6;
    const uint32_t blk = xcd_remap(blockIdx.x, gridDim.x);
    const uint32_t i = (blk * kTravWaves + w) * 64 + lane_id();
    const bool valid = i < n;
    if (count_step && blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&meta[META_STEPS_DONE], 1u);
    if (ballot64(valid) == 0) return;
    const double2 p = valid ? pos[i] : make_double2(0.0, 0.0);
    Acc a;
    const bool ok = exact_walk<GUARD, DIAG, REFF>(nodes, pos, mass, p, (int32_t)i, valid,
                                                  k->theta_sq, (float)k->eps2, s_node[w],
                                                  s_mask[w], a);
    if (!ok && lane_id() == 0) atomicOr(&meta[META_FLAGS], (uint32_t)FLAG_STACK);
    if (!valid) return;
    const double ax = (double)a.ax - (double)a.cx;
    const double ay = (double)a.ay - (double)a.cy;
    if (KICK) {
        const double gdt = k->G_dt;
        double2 v = vel[i];
        v.x += ax * gdt;
        v.y += ay * gdt;
        vel[i] = v;
    } else {
        const double G = k->G;
        acc_out[i] = make_double2(ax * G, ay * G);
        inter_out[i] = a.n;
    }
}

// ---------------------------------------------------------------------------
// Group traversal: one wavefront per 64 Morton-adjacent bodies. Phase 1 walks
// the tree 64 cells at a time (one per lane) from a shared LDS stack; a cell
// is accepted when wsq < theta^2 dmin^2, dmin = distance from its COM to the
// group's bounding box -- then every member's own test holds too. Accepted
// cells and single-body leaves go to an LDS interaction list as FP32 offsets
// from the group centre (the FP64 subtraction is done once per entry).
// Phase 2: every lane sums the list (broadcast LDS reads), 14 flops per entry,
// FP32 tile sums added with Kahan compensation. The list is flushed whenever
// fewer than 64 free entries remain. A step takes up to 64 cells from the top
// of the stack and pushes at most 4 children for each (net +3). The take is
// throttled to keep kGReserve entries free: from any stack depth sp, taking
// one cell at a time (depth first) never needs more than sp + 3 * 30 + 1
// entries, because a cell at level L has at most 30 - L levels below it, so
// the stack cannot overflow.
// ---------------------------------------------------------------------------
constexpr int kGReserve = 3 * kMaxLevel + 4;
constexpr int kGList = 256;   // the list is flushed past `flush` (<= kGList - 64) entries
// Before its walk a group is split at its largest Morton jump when the two
// parts are much more compact than the whole. As a safety net, a group whose
// interaction list still passes `budget` entries (default kGroupBudget), or
// whose walk takes more than kGMaxSteps steps, is split into pieces of 8
// bodies (a piece of at most 8 into single bodies), shared with other
// wavefronts through the work queue, and walked again. A typical group walk
// takes 15-25 steps; the long ones (some 440 steps, thousands of entries)
// are groups whose wide box reaches into dense regions, e.g. bodies thrown
// across a disk. Such a walk only matters when it outlasts the launch, so the
// step limit is what the launch's resident wavefronts get through anyway:
// kGWalkSteps per round of walks, at least kGMinSteps. On a 1M-body disk
// (5 rounds) that is 96 steps and the split pieces start while most
// wavefronts are still busy; on a 5.6M-body ensemble (26 rounds) the long
// walks hide behind the others, and splitting them, each piece as long
// again, made the launch 18x slower. (A flat limit of 48 split ordinary walks
// too, in cascades.) Both this limit and the list budget hold for theta >=
// 0.5; below it walks and lists grow (the accepted cells of a 2-D walk go as
// 1/theta^2), so both limits grow by 0.5/theta -- without that nearly every
// group splits, and the pieces, each walked again from the root, split again
// into single bodies that use one lane of their wavefront (a 1M disk at
// theta 0.15: 949 ms per step instead of 4.2). Growing them by (0.5/theta)^2
// instead left more stragglers unsplit (theta 0.3: 1.91 vs 1.76 ms).
constexpr uint32_t kGMinSteps = 96;
constexpr uint32_t kGWalkSteps = 20;

// ES = 2 when two lanes share a body (32 bodies per wavefront): each sums
// every other entry, from e0.
template <bool ZERO_EPS, bool REFF, int ES>
__device__ __forceinline__ void eval_list(const float4* __restrict__ list, int e0, int cnt,
                                          float bx, float by, float eps2f, Acc& a) {
    float px = 0.f, py = 0.f;
#pragma unroll 8
    for (int e = e0; e < cnt; e += ES) {
        const float4 d = list[e];
        const float dx = d.x - bx, dy = d.y - by;
        float w;
        if (REFF) {
            const float d2 = fmaf(dx, dx, dy * dy);
            const float s = d.z * __builtin_amdgcn_rsqf(d2) * __builtin_amdgcn_rcpf(d2 + eps2f);
            w = d2 > 0.f ? s : 0.f;  // the body's own entry
        } else {
            const float r2 = fmaf(dx, dx, fmaf(dy, dy, eps2f));
            float rs = __builtin_amdgcn_rsqf(r2);
            if (ZERO_EPS) rs = r2 > 0.f ? rs : 0.f;
            w = d.z * (rs * rs * rs);
        }
        px = fmaf(dx, w, px);
        py = fmaf(dy, w, py);
    }
    kahan_add(a.ax, a.cx, px);
    kahan_add(a.ay, a.cy, py);
}

// The learned far field's near cells: every entry also carries the cell's
// second moments about the COM (ixx in the entry's w, ixy and iyy in `quad`),
// and each body adds the cell's softened quadrupole term at its own position
// to the monopole,
//   a = m d / R^3 + (3 / (2 R^5)) [(5 (d.I d) / R^2 - tr I) d - 2 I d],
// d = COM - body, R^2 = |d|^2 + eps^2: 36 flops per entry.
template <bool ZERO_EPS, int ES>
__device__ __forceinline__ void eval_list_quad(const float4* __restrict__ list,
                                               const float2* __restrict__ quad, int e0, int cnt,
                                               float bx, float by, float eps2f, Acc& a) {
    float px = 0.f, py = 0.f;
#pragma unroll 4
    for (int e = e0; e < cnt; e += ES) {
        const float4 d = list[e];
        const float2 q = quad[e];
        const float dx = d.x - bx, dy = d.y - by;
        const float r2 = fmaf(dx, dx, fmaf(dy, dy, eps2f));
        float rs = __builtin_amdgcn_rsqf(r2);
        if (ZERO_EPS) rs = r2 > 0.f ? rs : 0.f;
        const float rs2 = rs * rs, rs3 = rs2 * rs, rs5 = rs3 * rs2;
        const float idx = fmaf(d.w, dx, q.x * dy), idy = fmaf(q.x, dx, q.y * dy);
        const float S = fmaf(dx, idx, dy * idy);
        const float cq = 1.5f * rs5;
        const float cd = fmaf(cq, fmaf(5.f * S, rs2, -(d.w + q.y)), d.z * rs3);
        const float ci = -2.f * cq;
        px = fmaf(dx, cd, fmaf(ci, idx, px));
        py = fmaf(dy, cd, fmaf(ci, idy, py));
    }
    kahan_add(a.ax, a.cx, px);
    kahan_add(a.ay, a.cy, py);
}

// Work sharing: a subgroup whose list passes the budget is split into pieces
// of 8 bodies (single bodies below that); all pieces but one go to a queue in device memory (meta[META_QHEAD], meta[META_QTAIL]
// count pushes and pops; an entry is {group, first lane << 8 | lanes} + 1 in
// the high word so that 0 means "not yet written"), and a wavefront that has
// finished its own group takes queued subgroups before it exits. Every
// wavefront that pushes also drains the queue, so each entry is taken once
// and the queue is empty (all zero) when the kernel ends. Each subgroup's
// result depends only on its own walk, so the outcome does not depend on
// which wavefront takes it.
__device__ __forceinline__ bool queue_push(uint32_t* meta, uint64_t* queue, uint32_t cap,
                                           uint32_t g, uint32_t sub) {
    bool ok = false;
    if (lane_id() == 0) {
        const uint32_t at = atomicAdd(&meta[META_QHEAD], 1u);
        if (at < cap) {
            __threadfence();
            __hip_atomic_store(&queue[at], ((uint64_t)(sub + 1u) << 32) | g, __ATOMIC_RELEASE,
                               __HIP_MEMORY_SCOPE_AGENT);
            ok = true;
        }
    }
    return __shfl((int)ok, 0, 64) != 0;
}

__device__ __forceinline__ bool queue_pop(uint32_t* meta, uint64_t* queue, uint32_t cap,
                                          uint32_t& g, uint32_t& sub) {
    uint64_t v = 0;
    if (lane_id() == 0) {
        for (;;) {
            const uint32_t h = __hip_atomic_load(&meta[META_QHEAD], __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT);
            if (h == 0) break;  // nothing was queued in this launch (the common case)
            const uint32_t t = __hip_atomic_load(&meta[META_QTAIL], __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT);
            if (t >= min(h, cap)) break;
            if (atomicCAS(&meta[META_QTAIL], t, t + 1u) == t) {
                // the pusher of entry t has counted it and is about to store it
                do {
                    v = __hip_atomic_load(&queue[t], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
                } while ((v >> 32) == 0);
                __hip_atomic_store(&queue[t], (uint64_t)0, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
        }
    }
    const uint32_t lo = __shfl((uint32_t)v, 0, 64), hi = __shfl((uint32_t)(v >> 32), 0, 64);
    g = lo;
    sub = hi - 1u;
    return hi != 0;
}

// LEARN: each body adds the softened quadrupole terms of the accepted cells
// near its subgroup (within near_sq squared box half-diagonals of the box
// centre) at its own position (eval_list_quad); the farther cells' terms, and
// the far-field network, are left to k_mlp_apply, at the subgroup centre. The
// two kinds of entry fill the LDS list from its two ends. The tree slot
// of every accepted internal cell is also appended to `cell_slot`, in the
// region of the group the wavefront was launched for; each finished subgroup
// k of group g gets subs[g][k] (its box centre and its cell range) from a
// per-group counter, each member its subgroup and its offset from that
// centre, and the raw acceleration (without G) goes to acc_out instead of a
// kick.
// STACK: LDS stack entries (fewer: more resident wavefronts, smaller takes
// near the top of the stack). BPW: bodies per wavefront, 64 (one per lane) or
// 32 (lanes l and l + 32 hold the same body and sum alternate list entries).
template <bool KICK, bool ZERO_EPS, bool GUARD, bool LEARN, bool REFF, int STACK, int BPW>
__global__ __launch_bounds__(64) void k_traverse_group(
    const NodeRec* __restrict__ nodes, const double2* __restrict__ pos,
    const double* __restrict__ mass, double2* __restrict__ vel, double2* __restrict__ acc_out,
    uint32_t* __restrict__ inter_out, uint32_t n, const StepConsts* __restrict__ k,
    uint32_t* __restrict__ meta, int count_step, int budget, uint32_t max_steps,
    const uint64_t* __restrict__ keys, uint32_t* __restrict__ cell_slot,
    uint32_t* __restrict__ cell_cnt, uint32_t cell_cap, SubRec* __restrict__ subs,
    float4* __restrict__ body_w, uint64_t* __restrict__ queue, uint32_t queue_cap,
    const uint32_t* __restrict__ grp_first, const uint32_t* __restrict__ grp_info, int flush,
    double near_sq) {
    static_assert(BPW == 64 || BPW == 32, "64 or 32 bodies per wavefront");
    static_assert(STACK >= kGReserve + 3 * 64, "stack too small for a full take");
    constexpr int ES = 64 / BPW;  // lanes per body
    if (meta[META_FLAGS] & kHaltFlags) return;
    __shared__ __attribute__((aligned(16))) int32_t s_stack[STACK];
    __shared__ float4 s_list[kGList];
    __shared__ float2 s_quad[LEARN ? kGList : 1];  // LEARN: ixy, iyy of each entry
    __shared__ int32_t s_todo[16];
    __shared__ unsigned long long s_own;  // bodies whose own single-body leaf is listed
    const int l = lane_id();
    const int bl = l & (BPW - 1);  // the lane's body in the group
    const int part = l / BPW;      // which of the body's lanes (BPW = 32)
    const uint32_t home = xcd_remap(blockIdx.x, gridDim.x);
    if (count_step && blockIdx.x == 0 && l == 0) atomicAdd(&meta[META_STEPS_DONE], 1u);
    const double theta_sq = k->theta_sq;
    const float eps2f = (float)k->eps2;
    bool overflow = false;
    uint32_t cbase = 0;  // cells written to the home region (LEARN)
    uint64_t qterms = 0; // near-cell quadrupole terms applied, over all members (LEARN)
    const uint64_t region = (uint64_t)home * cell_cap;
    if (LEARN && l == 0)
        __hip_atomic_store(&cell_cnt[home], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint32_t g = home;
    // Pending subgroups of group g as (first lane << 8 | lane count); a
    // single body is never split (its group test is its own opening test).
    int nt = 1;
    if (l == 0) s_todo[0] = BPW;
    for (;;) {
        // group g: bodies first .. first + gn - 1 of system gsys
        const uint32_t first = grp_first[g], ginfo = grp_info[g];
        const int gn = (int)(ginfo & 255u);
        const int32_t root = (int32_t)(ginfo >> 8);
        const uint32_t i = first + bl;
        const bool valid = bl < gn;
        const double2 p = valid ? pos[i] : pos[first];
        const int32_t me = (int32_t)i;
        // Morton jump to the next body: digits shared with it (64 = none
        // follows); the second lane of a body (BPW = 32) reports none.
        const uint64_t key = valid ? keys[i] : 0ull;
        const uint64_t key_next = __shfl_down(key, 1, 64);
        const int jump = (valid && bl + 1 < gn && part == 0) ? __clzll(key ^ key_next) : 64;
        while (nt > 0 && !overflow) {
            const int32_t e = readfirst(s_todo[--nt]);
            const int f = e >> 8, cn = e & 255;
            const bool member = valid && bl >= f && bl < f + cn;
            if (ballot64(member) == 0) continue;
            double xmin = member ? p.x : INFINITY, xmax = member ? p.x : -INFINITY;
            double ymin = member ? p.y : INFINITY, ymax = member ? p.y : -INFINITY;
#pragma unroll
            for (int d = 32; d >= 1; d >>= 1) {
                xmin = fmin(xmin, __shfl_xor(xmin, d, 64));
                xmax = fmax(xmax, __shfl_xor(xmax, d, 64));
                ymin = fmin(ymin, __shfl_xor(ymin, d, 64));
                ymax = fmax(ymax, __shfl_xor(ymax, d, 64));
            }
            // Split at the largest Morton jump inside the group when both
            // parts are far more compact than the whole (the group straddles
            // a jump of the curve); no walk is wasted on the wide box.
            if (cn > 1) {
                const int my = (member && bl < f + cn - 1) ? jump : 65;
                int best = my;
#pragma unroll
                for (int d = 32; d >= 1; d >>= 1) best = min(best, __shfl_xor(best, d, 64));
                // (jumps are reported by lanes 0..BPW-1 only: lane = body)
                const uint64_t at = ballot64(member && my == best);
                const int cut = __ffsll((unsigned long long)at);  // first body after the jump
                if (best < 64 && cut > f && cut < f + cn) {
                    const bool left = bl < cut;
                    double ax0 = member && left ? p.x : INFINITY, ax1 = member && left ? p.x : -INFINITY;
                    double ay0 = member && left ? p.y : INFINITY, ay1 = member && left ? p.y : -INFINITY;
                    double bx0 = member && !left ? p.x : INFINITY, bx1 = member && !left ? p.x : -INFINITY;
                    double by0 = member && !left ? p.y : INFINITY, by1 = member && !left ? p.y : -INFINITY;
#pragma unroll
                    for (int d = 32; d >= 1; d >>= 1) {
                        ax0 = fmin(ax0, __shfl_xor(ax0, d, 64));
                        ax1 = fmax(ax1, __shfl_xor(ax1, d, 64));
                        ay0 = fmin(ay0, __shfl_xor(ay0, d, 64));
                        ay1 = fmax(ay1, __shfl_xor(ay1, d, 64));
                        bx0 = fmin(bx0, __shfl_xor(bx0, d, 64));
                        bx1 = fmax(bx1, __shfl_xor(bx1, d, 64));
                        by0 = fmin(by0, __shfl_xor(by0, d, 64));
                        by1 = fmax(by1, __shfl_xor(by1, d, 64));
                    }
                    const double ew = fmax(xmax - xmin, ymax - ymin);
                    const double ea = fmax(ax1 - ax0, ay1 - ay0);
                    const double eb = fmax(bx1 - bx0, by1 - by0);
                    if (ew * ew > 4.0 * (ea * ea + eb * eb) && nt + 2 <= 16) {
                        if (l == 0) {
                            s_todo[nt] = (cut << 8) | (f + cn - cut);
                            s_todo[nt + 1] = (f << 8) | (cut - f);
                        }
                        nt += 2;
                        continue;
                    }
                }
            }
            const double ox = 0.5 * (xmin + xmax), oy = 0.5 * (ymin + ymax);
            const double hx = 0.5 * (xmax - xmin), hy = 0.5 * (ymax - ymin);
            const float bx = (float)(p.x - ox), by = (float)(p.y - oy);
            // LEARN: a cell whose COM is farther than this (squared) from the
            // centre takes its quadrupole term at the centre in k_mlp_apply
            // (every cell, for a single body: its offset from the centre is 0)
            const double thr = near_sq < 0.0 ? 1e300 : near_sq * fma(hx, hx, hy * hy);
            const int32_t g_first = (int32_t)(first + f);
            const int32_t g_last = (int32_t)(first + min(f + cn, gn) - 1);
            const bool capped = cn > 1;
            Acc t;
            int sp = 1, cnt = 0;
            int cnq = 0;        // LEARN: near cells, at the top end of the list
            uint32_t ccnt = 0;  // cells of this subgroup written after cbase
            uint32_t listed = 0, steps = 0;
            uint32_t listedq = 0;  // LEARN: near-cell entries summed
            bool split = false;
            if (l == 0) s_own = 0ull;
            if (l == 0) s_stack[0] = root;  // the root of this group's system
            // Each step pops up to 64 stack entries and loads their records; the
            // next step's records are loaded before the list is flushed, so
            // their latency overlaps the list's evaluation.
            bool have = false;
            int32_t node = 0;
            NodeRec r;
            auto pop = [&]() -> bool {  // false: the stack would overflow
                const int take = max(1, min(min(64, sp), (STACK - kGReserve - sp) / 3));
                if (sp + 3 > STACK) return false;  // unreachable: see kGReserve
                const int base = sp - take;
                have = l < take;
                node = have ? s_stack[base + l] : 0;
                sp = base;
                if (have) r = nodes[node];  // one 64-byte record: four independent loads
                return true;
            };
            overflow = !pop();
            while (!overflow) {
                steps++;
                const int32_t c0 = have ? r.c0 : -2;
                const double dxc = have ? r.cx - ox : 0.0, dyc = have ? r.cy - oy : 0.0;
                bool add = false, open = false;
                int32_t bs = 0, be = 0;
                if (have) {
                    if (c0 < 0) {
                        bs = r.bs;
                        be = r.be;
                        add = bs == be;
                    } else {
                        const double ex = fmax(fabs(dxc) - hx, 0.0), ey = fmax(fabs(dyc) - hy, 0.0);
                        const double dmin2 = fma(ex, ex, ey * ey);
                        add = r.wsq < theta_sq * dmin2;
                        if (GUARD) add = add && !(r.bs <= g_last && r.be >= g_first);
                        open = !add;
                    }
                }
                // append accepted cells / single-body leaves to the list: the
                // learned walk's near cells, whose quadrupole term every body
                // adds at its own position, fill it from the top end
                const bool nearc = LEARN && add && c0 >= 0 && !(fma(dxc, dxc, dyc * dyc) > thr);
                const uint64_t am = ballot64(add && !nearc);
                const uint64_t an = LEARN ? ballot64(nearc) : 0ull;
                if (add) {
                    if (LEARN && nearc) {
                        const int at = kGList - 1 - cnq - (int)__popcll(an & lanemask_lt());
                        s_list[at] = make_float4((float)dxc, (float)dyc, r.mf, r.ixx);
                        s_quad[at] = make_float2(r.ixy, r.iyy);
                    } else {
                        const int at = cnt + (int)__popcll(am & lanemask_lt());
                        s_list[at] = make_float4((float)dxc, (float)dyc, r.mf, 0.f);
                    }
                    // a member's own single-body leaf: a zero term, not counted
                    if (c0 < 0 && bs >= g_first && bs <= g_last)
                        atomicOr(&s_own, 1ull << (bs - (int32_t)first));
                }
                cnt += (int)__popcll(am);
                cnq += (int)__popcll(an);
                if (LEARN) {
                    const bool cell = add && c0 >= 0;
                    const uint64_t cmk = ballot64(cell);
                    if (cell) {
                        const uint32_t at = cbase + ccnt + (uint32_t)__popcll(cmk & lanemask_lt());
                        if (at < cell_cap) cell_slot[region + at] = (uint32_t)node;
                    }
                    ccnt += (uint32_t)__popcll(cmk);
                }
                // push the non-empty children of opened cells: a lane's slot
                // is the children of the lanes below it, counted from one
                // ballot per child position (no cross-lane scan)
                const int32_t cm = open ? r.mask : 0;
                const uint64_t q0 = ballot64(cm & 1), q1 = ballot64(cm & 2);
                const uint64_t q2 = ballot64(cm & 4), q3 = ballot64(cm & 8);
                const int total = __popcll(q0) + __popcll(q1) + __popcll(q2) + __popcll(q3);
                if (open) {
                    int at = sp + (int)(mask_below(q0) + mask_below(q1) + mask_below(q2) +
                                        mask_below(q3));
#pragma unroll
                    for (int q = 3; q >= 0; q--)
                        if ((cm >> q) & 1) s_stack[at++] = c0 + q;
                }
                sp += total;
                // multi-body leaves (level 30): every lane sums their bodies directly
                uint64_t mm = ballot64(have && c0 < 0 && be > bs);
                while (mm) {
                    const int src = __ffsll((unsigned long long)mm) - 1;
                    mm &= mm - 1;
                    const int32_t js = __shfl(bs, src, 64), je = __shfl(be, src, 64);
                    for (int32_t j = js + part; j <= je; j += ES) {
                        if (j == me) continue;
                        const double2 q = pos[j];
                        const double ex = q.x - p.x, ey = q.y - p.y;
                        interact<REFF>(t, ex, ey, fma(ex, ex, ey * ey), (float)mass[j], eps2f);
                    }
                }
                if (capped && (listed + (uint32_t)(cnt + cnq) > (uint32_t)budget ||
                               steps > max_steps) &&
                    nt + 8 <= 16) {
                    split = true;
                    break;
                }
                const bool more = sp > 0;
                if (more && !pop()) {
                    overflow = true;
                    break;
                }
                if (cnt + cnq > flush) {
                    eval_list<ZERO_EPS, REFF, ES>(s_list, part, cnt, bx, by, eps2f, t);
                    if (LEARN && cnq > 0)
                        eval_list_quad<ZERO_EPS, ES>(s_list + (kGList - cnq), s_quad + (kGList - cnq),
                                                     part, cnq, bx, by, eps2f, t);
                    listed += cnt + cnq;
                    listedq += cnq;
                    cnt = cnq = 0;
                }
                if (!more) break;
            }
            if (overflow) break;
            if (split) {
                // Pieces of 8 bodies (single bodies below that): all but the first go
                // to the shared queue (or stay here when it is full), so the
                // pieces of a heavy group are walked by several wavefronts.
                const int pieces = cn > 8 ? (cn + 7) / 8 : cn;
                const int len = (cn + pieces - 1) / pieces;
                for (int q = pieces - 1; q >= 1; q--) {
                    const int pf = f + q * len, pc = min(len, f + cn - pf);
                    if (pc <= 0) continue;
                    const uint32_t e2 = (uint32_t)((pf << 8) | pc);
                    if (!queue_push(meta, queue, queue_cap, g, e2)) {
                        if (l == 0) s_todo[nt] = (int32_t)e2;
                        nt++;
                    }
                }
                if (l == 0) s_todo[nt] = (f << 8) | min(len, cn);
                nt++;
                continue;
            }
            if (cnt + cnq > 0) {
                eval_list<ZERO_EPS, REFF, ES>(s_list, part, cnt, bx, by, eps2f, t);
                if (LEARN && cnq > 0)
                    eval_list_quad<ZERO_EPS, ES>(s_list + (kGList - cnq), s_quad + (kGList - cnq),
                                                 part, cnq, bx, by, eps2f, t);
                listed += cnt + cnq;
                listedq += cnq;
            }
            // the subgroup is finished: its members write their results
            if (LEARN) {
                uint32_t ks = 0;
                if (l == 0) ks = atomicAdd(&cell_cnt[g], 1u);
                ks = __shfl(ks, 0, 64) & 63u;
                if (member && part == 0)
                    body_w[i] = make_float4(bx, by, __int_as_float((int)ks), 0.f);
                if (l == 0) {
                    SubRec sr;
                    sr.ox = ox;
                    sr.oy = oy;
                    sr.start = (uint32_t)(region + cbase);
                    sr.count = ccnt;
                    sr.thr = thr;
                    subs[(uint64_t)g * 64 + ks] = sr;
                }
                cbase += ccnt;
                qterms += (uint64_t)listedq * (uint64_t)__popcll(ballot64(member && part == 0));
            }
            double ax = (double)t.ax - (double)t.cx;
            double ay = (double)t.ay - (double)t.cy;
            uint32_t tn = t.n;
            if (BPW == 32) {  // the body's two lanes summed alternate entries
                ax += __shfl_xor(ax, 32, 64);
                ay += __shfl_xor(ay, 32, 64);
                tn += (uint32_t)__shfl_xor((int)tn, 32, 64);
            }
            if (member && part == 0) {
                // the body's own single-body leaf, when listed, is a zero term
                const uint32_t terms = listed - (uint32_t)((s_own >> bl) & 1ull) + tn;
                if (LEARN) {
                    acc_out[i] = make_double2(ax, ay);
                    if (!KICK) inter_out[i] = terms;
                } else if (KICK) {
                    const double gdt = k->G_dt;
                    double2 v = vel[i];
                    v.x += ax * gdt;
                    v.y += ay * gdt;
                    vel[i] = v;
                } else {
                    const double G = k->G;
                    acc_out[i] = make_double2(ax * G, ay * G);
                    inter_out[i] = terms;
                }
            }
        }
        // this group is done (or the stack overflowed): take a queued subgroup
        uint32_t qg, qsub;
        if (!queue_pop(meta, queue, queue_cap, qg, qsub)) break;
        if (overflow) continue;  // drain the queue without walking
        g = qg;
        if (l == 0) s_todo[0] = (int32_t)qsub;
        nt = 1;
    }
    if (overflow && l == 0) atomicOr(&meta[META_FLAGS], (uint32_t)FLAG_STACK);
    if (LEARN && l == 0) {
        atomicAdd((unsigned long long*)&meta[META_CELLS_TOTAL], (unsigned long long)cbase);
        atomicAdd((unsigned long long*)&meta[META_QUAD_TERMS], (unsigned long long)qterms);
        atomicMax(&meta[META_CELLS_NEEDED], cbase);  // sizes the regions at creation
        if (cbase > cell_cap) atomicOr(&meta[META_FLAGS], (uint32_t)FLAG_CELL_CAPACITY);
    }
}

// The group walk with the context's walk settings (stack depth, bodies per
// wavefront; 0 = the defaults).
template <bool K, bool Z, bool GD, bool LN, bool RF>
void launch_walk(nb_ctx* c, int src, int count, int budget, uint32_t max_steps, int flush) {
    const uint32_t n = (uint32_t)c->n;
    const int stack = walk_stack(c), bpw = walk_bodies(c);
#define NB_WALK(S, B)                                                                          \
    hipLaunchKernelGGL((k_traverse_group<K, Z, GD, LN, RF, S, B>), dim3(c->ngroups), dim3(64), \
                       0, c->stream, c->b.nodes, c->b.pos[src], c->b.mass[src], c->b.vel[src], \
                       c->b.acc, c->b.inter, n, c->b.consts, c->b.meta, count, budget,         \
                       max_steps, c->b.keys[0], c->b.cell_slot, c->b.cell_cnt, c->cell_cap,    \
                       c->b.subs, c->b.body_w, c->b.queue, c->queue_cap, c->b.grp_first,       \
                       c->b.grp_info, flush, quad_near_sq(c))
    if (stack == 1024) {
        if (bpw == 32) NB_WALK(1024, 32); else NB_WALK(1024, 64);
    } else {
        if (bpw == 32) NB_WALK(2048, 32); else NB_WALK(2048, 64);
    }
#undef NB_WALK
}

void launch_traverse(nb_ctx* c, int src, bool kick, bool count_step) {
    const uint32_t n = (uint32_t)c->n;
    if (n == 0) return;
    if (c->p.traversal == NB_TRAV_GROUP) {
        const uint32_t groups = c->ngroups;
        const bool z = !(c->p.softening > 0.0);
        const bool gd = c->p.self_guard != 0;
        const double th = c->p.theta;
        const double grow = th >= 0.5 ? 1.0 : th > 1e-3 ? 0.5 / th : 1e6;
        const int budget = (int)std::min(
            (double)(c->p.group_budget > 0 ? c->p.group_budget : kGroupBudget) * grow, 1e9);
        // resident wavefronts per CU: the LDS stack and list bound them
        const bool learn = c->p.far_field == NB_FAR_LEARNED;
        const uint32_t lds = (uint32_t)walk_stack(c) * 4 + kGList * (learn ? 24 : 16) + 160;
        const uint64_t per_cu = std::max<uint32_t>(1, (160u << 10) / lds);
        const uint64_t rounds = ((uint64_t)groups + (uint64_t)c->cus * per_cu - 1) /
                                ((uint64_t)c->cus * per_cu);
        const uint32_t max_steps = (uint32_t)std::min(
            std::max((double)kGMinSteps * grow, (double)(rounds * kGWalkSteps)), 1e9);
        const int flush = list_flush(c);
        const bool reff = c->p.force_kernel == NB_FORCE_REF_CPU;
        const int count = (int)(count_step && !learn);
#define NB_TRAVG(K, Z, GD, LN, RF) launch_walk<K, Z, GD, LN, RF>(c, src, count, budget, max_steps, flush)
#define NB_TRAVG_GD(K, Z, LN, RF)                                                              \
    {                                                                                          \
        if (gd)                                                                                \
            NB_TRAVG(K, Z, true, LN, RF);                                                      \
        else                                                                                   \
            NB_TRAVG(K, Z, false, LN, RF);                                                     \
    }
        // the learned far field runs with the Plummer law (checked at creation);
        // REF_CPU drops zero-distance terms itself, whatever eps is
        if (learn) {
            if (z) NB_TRAVG_GD(false, true, true, false) else NB_TRAVG_GD(false, false, true, false)
        } else if (reff) {
            if (kick) NB_TRAVG_GD(true, false, false, true) else NB_TRAVG_GD(false, false, false, true)
        } else if (kick) {
            if (z) NB_TRAVG_GD(true, true, false, false) else NB_TRAVG_GD(true, false, false, false)
        } else {
            if (z) NB_TRAVG_GD(false, true, false, false) else NB_TRAVG_GD(false, false, false, false)
        }
#undef NB_TRAVG_GD
#undef NB_TRAVG
        return;
    }
    const uint32_t blocks = (n + kTravThreads - 1) / kTravThreads;
    const bool guard = c->p.self_guard != 0;
    const bool diag = c->p.mac == NB_MAC_DIAG;
    const bool reff = c->p.force_kernel == NB_FORCE_REF_CPU;
#define NB_TRAV(K, G, D, R)                                                                    \
    hipLaunchKernelGGL((k_traverse<K, G, D, R>), dim3(blocks), dim3(kTravThreads), 0,          \
                       c->stream, c->b.nodes, c->b.pos[src], c->b.mass[src], c->b.vel[src],    \
                       c->b.acc, c->b.inter, n, c->b.consts, c->b.meta, (int)count_step)
#define NB_TRAV_R(K, G, D)                                                                     \
    {                                                                                          \
        if (reff)                                                                              \
            NB_TRAV(K, G, D, true);                                                            \
        else                                                                                   \
            NB_TRAV(K, G, D, false);                                                           \
    }
#define NB_TRAV_D(K, G)                                                                        \
    {                                                                                          \
        if (diag)                                                                              \
            NB_TRAV_R(K, G, true) else NB_TRAV_R(K, G, false)                                  \
    }
#define NB_TRAV_G(K)                                                                           \
    {                                                                                          \
        if (guard)                                                                             \
            NB_TRAV_D(K, true) else NB_TRAV_D(K, false)                                        \
    }
    if (kick) NB_TRAV_G(true) else NB_TRAV_G(false)
#undef NB_TRAV_G
#undef NB_TRAV_D
#undef NB_TRAV_R
#undef NB_TRAV
}

// ---------------------------------------------------------------------------
// All-pairs: each thread owns R bodies, the workgroup streams 256-body source
// tiles through LDS; the j range is split across blockIdx.y so a
// 262,144-body launch still fills 256 CUs. Positions stay FP64 (in LDS and
// registers) and each offset dx = x_j - x_i is one FP64 subtraction rounded
// once to FP32: accurate to half an FP32 ulp of dx itself however close two
// bodies are (rounding the absolute positions to FP32 first leaves an error
// of ~6e-8 |x| / |dx|, which at eps = 0 dominates close pairs). Without
// packed FP32 an FP64 add issues at the FP32 rate, so this costs two VALU
// operations per pair (the two conversions) over the FP32 subtraction. Each
// 256-term tile is summed plainly, and tile sums are added to the body's
// running total with Kahan compensation.
// ---------------------------------------------------------------------------
constexpr int kApThreads = 256;

template <bool ZERO_EPS, bool REFF, int kApR>
__global__ __launch_bounds__(kApThreads) void k_allpairs(const double2* __restrict__ pos,
                                                         const double* __restrict__ mass,
                                                         uint32_t n, uint32_t chunk,
                                                         const StepConsts* __restrict__ k,
                                                         double2* __restrict__ part) {
    __shared__ double2 tile[kApThreads];
    __shared__ float tmass[kApThreads];
    const uint32_t ib = blockIdx.x * (kApThreads * kApR) + threadIdx.x;
    double2 pi[kApR];
    float ax[kApR], ay[kApR], cx[kApR], cy[kApR];
#pragma unroll
    for (int r = 0; r < kApR; r++) {
        const uint32_t i = ib + r * kApThreads;
        pi[r] = i < n ? pos[i] : make_double2(0.0, 0.0);
        ax[r] = ay[r] = cx[r] = cy[r] = 0.f;
    }
    const float eps2f = (float)k->eps2;
    const uint32_t j0 = blockIdx.y * chunk;
    const uint32_t j1 = min(j0 + chunk, n);
    for (uint32_t jt = j0; jt < j1; jt += kApThreads) {
        const uint32_t j = jt + threadIdx.x;
        if (j < j1) {
            tile[threadIdx.x] = pos[j];
            tmass[threadIdx.x] = (float)mass[j];
        } else {
            tile[threadIdx.x] = make_double2(0.0, 0.0);
            tmass[threadIdx.x] = 0.f;
        }
        __syncthreads();
        float px[kApR], py[kApR];
#pragma unroll
        for (int r = 0; r < kApR; r++) px[r] = py[r] = 0.f;
#pragma unroll 8
        for (int jj = 0; jj < kApThreads; jj++) {
            const double2 s = tile[jj];
            const float mj = tmass[jj];
#pragma unroll
            for (int r = 0; r < kApR; r++) {
                const float dx = (float)(s.x - pi[r].x);
                const float dy = (float)(s.y - pi[r].y);
                float w;
                if (REFF) {  // Simulation::force (Simulation.cpp:74-80)
                    const float d2 = fmaf(dx, dx, dy * dy);
                    const float t =
                        mj * __builtin_amdgcn_rsqf(d2) * __builtin_amdgcn_rcpf(d2 + eps2f);
                    w = d2 > 0.f ? t : 0.f;  // the body itself
                } else {
                    const float r2 = fmaf(dx, dx, fmaf(dy, dy, eps2f));
                    float rs = __builtin_amdgcn_rsqf(r2);
                    if (ZERO_EPS) rs = r2 > 0.f ? rs : 0.f;
                    w = mj * (rs * rs * rs);
                }
                px[r] = fmaf(dx, w, px[r]);
                py[r] = fmaf(dy, w, py[r]);
            }
        }
#pragma unroll
        for (int r = 0; r < kApR; r++) {
            kahan_add(ax[r], cx[r], px[r]);
            kahan_add(ay[r], cy[r], py[r]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < kApR; r++) {
        const uint32_t i = ib + r * kApThreads;
        if (i < n)
            part[(uint64_t)blockIdx.y * n + i] =
                make_double2((double)ax[r] - (double)cx[r], (double)ay[r] - (double)cy[r]);
    }
}

// Sums the split partials in a fixed order, then kicks or stores the result.
__global__ __launch_bounds__(256) void k_allpairs_finish(
    const double2* __restrict__ part, uint32_t split, uint32_t n, const StepConsts* __restrict__ k,
    double2* __restrict__ vel, double2* __restrict__ acc_out, uint32_t* __restrict__ inter_out,
    int kick, uint32_t* __restrict__ meta, int count_step) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (count_step && i == 0) atomicAdd(&meta[META_STEPS_DONE], 1u);
    if (i >= n) return;
    double ax = 0.0, ay = 0.0;
    for (uint32_t s = 0; s < split; s++) {
        const double2 v = part[(uint64_t)s * n + i];
        ax += v.x;
        ay += v.y;
    }
    if (kick) {
        double2 v = vel[i];
        v.x += ax * k->G_dt;
        v.y += ay * k->G_dt;
        vel[i] = v;
    } else {
        acc_out[i] = make_double2(ax * k->G, ay * k->G);
        inter_out[i] = n - 1;
    }
}

// Bodies per thread (c->ap_rows: 2, 4 or 8) and the j split (c->ap_split)
// are the all-pairs launch's tunables (nb_autotune).
void launch_allpairs(nb_ctx* c, int src, bool kick, bool count_step) {
    const uint32_t n = (uint32_t)c->n;
    if (n == 0) return;
    const uint32_t R = c->ap_rows;
    const uint32_t ib = (n + kApThreads * R - 1) / (kApThreads * R);
    const uint32_t split = c->ap_split;
    uint32_t chunk = (n + split - 1) / split;
    chunk = (chunk + kApThreads - 1) / kApThreads * kApThreads;
    const bool z = !(c->p.softening > 0.0);
    const bool reff = c->p.force_kernel == NB_FORCE_REF_CPU;
#define NB_AP(Z, RF, RR)                                                                       \
    hipLaunchKernelGGL((k_allpairs<Z, RF, RR>), dim3(ib, split), dim3(kApThreads), 0,          \
                       c->stream, c->b.pos[src], c->b.mass[src], n, chunk, c->b.consts,        \
                       c->b.ap_part)
#define NB_AP2(RR)                                                                             \
    {                                                                                          \
        if (reff)                                                                              \
            NB_AP(false, true, RR);                                                            \
        else if (z)                                                                            \
            NB_AP(true, false, RR);                                                            \
        else                                                                                   \
            NB_AP(false, false, RR);                                                           \
    }
    if (R == 2) {
        NB_AP2(2);
    } else if (R == 8) {
        NB_AP2(8);
    } else {
        NB_AP2(4);
    }
#undef NB_AP2
#undef NB_AP
    hipLaunchKernelGGL(k_allpairs_finish, dim3((n + 255) / 256), dim3(256), 0, c->stream,
                       c->b.ap_part, split, n, c->b.consts, c->b.vel[src], c->b.acc, c->b.inter,
                       (int)kick, c->b.meta, (int)count_step);
}

// ---------------------------------------------------------------------------
// Drift x += v dt. HBM traffic: 16 B vel + 16 B pos read, 16 B pos written.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_drift(double2* __restrict__ pos,
                                               const double2* __restrict__ vel, uint32_t n,
                                               const StepConsts* __restrict__ k,
                                               uint32_t* __restrict__ meta, int check,
                                               int count_step) {
    if (check && (meta[META_FLAGS] & kHaltFlags)) return;
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (count_step && i == 0) atomicAdd(&meta[META_STEPS_DONE], 1u);
    if (i >= n) return;
    const double dt = k->dt;
    double2 x = pos[i];
    const double2 v = vel[i];
    x.x += v.x * dt;
    x.y += v.y * dt;
    pos[i] = x;
}

// The drift of a one-system Barnes-Hut context, which also leaves the box
// partials of the new positions for the next step's box (one per workgroup,
// bbox_blocks workgroups striding over the bodies, as k_bbox_partial does).
// A halted step leaves the positions, and so their partials, as they were.
// HBM traffic: 32 B read + 16 B written per body.
constexpr int kDriftRounds = 8;  // bodies per thread loaded before the first update

__global__ __launch_bounds__(kBboxThreads) void k_drift_box(double2* __restrict__ pos,
                                                            const double2* __restrict__ vel,
                                                            uint32_t n,
                                                            const StepConsts* __restrict__ k,
                                                            uint32_t* __restrict__ meta,
                                                            int check, int count_step,
                                                            double* __restrict__ part) {
    if (check && (meta[META_FLAGS] & kHaltFlags)) return;
    if (count_step && blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&meta[META_STEPS_DONE], 1u);
    const double dt = k->dt;
    const uint64_t stride = (uint64_t)gridDim.x * kBboxThreads;
    double xmin = INFINITY, xmax = -INFINITY, ymin = INFINITY, ymax = -INFINITY;
    for (uint64_t base = (uint64_t)blockIdx.x * kBboxThreads + threadIdx.x; base < n;
         base += kDriftRounds * stride) {
        double2 x[kDriftRounds], v[kDriftRounds];
#pragma unroll
        for (int r = 0; r < kDriftRounds; r++) {
            const uint64_t i = base + r * stride;
            if (i < n) {
                x[r] = pos[i];
                v[r] = vel[i];
            }
        }
#pragma unroll
        for (int r = 0; r < kDriftRounds; r++) {
            const uint64_t i = base + r * stride;
            if (i < n) {
                x[r].x += v[r].x * dt;
                x[r].y += v[r].y * dt;
                pos[i] = x[r];
                xmin = fmin(xmin, x[r].x);
                xmax = fmax(xmax, x[r].x);
                ymin = fmin(ymin, x[r].y);
                ymax = fmax(ymax, x[r].y);
            }
        }
    }
    box_block_store<kBboxThreads>(xmin, xmax, ymin, ymax, part);
}

void launch_drift(nb_ctx* c, int src, bool check, bool count_step) {
    const uint32_t n = (uint32_t)c->n;
    if (box_from_drift(c)) {
        hipLaunchKernelGGL(k_drift_box, dim3(c->bbox_blocks), dim3(kBboxThreads), 0, c->stream,
                           c->b.pos[src], c->b.vel[src], n, c->b.consts, c->b.meta, (int)check,
                           (int)count_step, c->b.bbox_part);
        // a halted drift skips both the update and the partials (which then
        // still describe the unchanged positions); a captured drift runs at
        // the replays, which ensure_box_parts precedes
        if (!c->capturing) c->box_parts_ok = true;
        return;
    }
    hipLaunchKernelGGL(k_drift, dim3(n ? (n + 255) / 256 : 1), dim3(256), 0, c->stream,
                       c->b.pos[src], c->b.vel[src], n, c->b.consts, c->b.meta, (int)check,
                       (int)count_step);
}

// ---------------------------------------------------------------------------
// FP64 direct sum (the force-error reference): thread per target, sources through LDS in FP64,
// j range split across blockIdx.y, Kahan-compensated FP64 sums, IEEE sqrt and
// division.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_direct64(const double2* __restrict__ pos,
                                                  const double* __restrict__ mass, uint32_t n,
                                                  double eps2, const int64_t* __restrict__ tgt,
                                                  uint32_t m, uint32_t chunk,
                                                  double2* __restrict__ part) {
    __shared__ double3 tile[256];
    const uint32_t t = blockIdx.x * 256 + threadIdx.x;
    int64_t ti = t < m ? tgt[t] : -1;
    if (ti >= (int64_t)n) ti = -1;  // out of range: never read (finished as NaN)
    const double2 p = ti >= 0 ? pos[ti] : make_double2(0.0, 0.0);
    double sx = 0.0, sy = 0.0, cx = 0.0, cy = 0.0;
    const uint32_t j0 = blockIdx.y * chunk, j1 = min(j0 + chunk, n);
    for (uint32_t jt = j0; jt < j1; jt += 256) {
        const uint32_t j = jt + threadIdx.x;
        if (j < j1) {
            const double2 q = pos[j];
            tile[threadIdx.x] = make_double3(q.x, q.y, mass[j]);
        }
        __syncthreads();
        const uint32_t cnt = min(256u, j1 - jt);
        for (uint32_t jj = 0; jj < cnt; jj++) {
            if ((int64_t)(jt + jj) == ti) continue;
            const double3 s = tile[jj];
            const double dx = s.x - p.x, dy = s.y - p.y;
            const double r2 = dx * dx + dy * dy + eps2;
            if (r2 == 0.0) continue;
            const double w = s.z / (r2 * sqrt(r2));
            const double vx = dx * w - cx, tx = sx + vx;
            cx = (tx - sx) - vx;
            sx = tx;
            const double vy = dy * w - cy, ty = sy + vy;
            cy = (ty - sy) - vy;
            sy = ty;
        }
        __syncthreads();
    }
    if (t < m) part[(uint64_t)blockIdx.y * m + t] = make_double2(sx - cx, sy - cy);
}

// Sums the split partials of each target in split order, then scales by G:
// the same additions, in the same order, as the host-side sum of
// nb_direct_sum_f64. A target outside [0, n) gets NaN.
__global__ __launch_bounds__(256) void k_direct64_finish(const double2* __restrict__ part,
                                                         uint32_t split, uint32_t m, uint32_t n,
                                                         const int64_t* __restrict__ tgt, double G,
                                                         double2* __restrict__ acc) {
    const uint32_t t = blockIdx.x * 256 + threadIdx.x;
    if (t >= m) return;
    double ax = 0.0, ay = 0.0;
    for (uint32_t k = 0; k < split; k++) {
        const double2 v = part[(uint64_t)k * m + t];
        ax += v.x;
        ay += v.y;
    }
    const int64_t ti = tgt[t];
    acc[t] = (ti >= 0 && ti < (int64_t)n) ? make_double2(ax * G, ay * G)
                                          : make_double2(__longlong_as_double(0x7ff8000000000000ll),
                                                         __longlong_as_double(0x7ff8000000000000ll));
}

hipError_t direct_sum_f64_finish(const double2* part, uint32_t split, uint64_t m, uint64_t n,
                                 const int64_t* targets, double G, double2* acc, hipStream_t s) {
    hipLaunchKernelGGL(k_direct64_finish, dim3((uint32_t)((m + 255) / 256)), dim3(256), 0, s,
                       part, split, (uint32_t)m, (uint32_t)n, targets, G, acc);
    return hipGetLastError();
}

hipError_t direct_sum_f64(uint64_t n, const double2* pos, const double* mass, double eps2,
                          uint64_t m, const int64_t* targets, double2* part, uint32_t split,
                          hipStream_t s) {
    uint32_t chunk = (uint32_t)((n + split - 1) / split);
    chunk = (chunk + 255) / 256 * 256;
    hipLaunchKernelGGL(k_direct64, dim3((uint32_t)((m + 255) / 256), split), dim3(256), 0, s, pos,
                       mass, (uint32_t)n, eps2, targets, (uint32_t)m, chunk, part);
    return hipGetLastError();
}

}  // namespace nb
