// nb_mlp.hip — the learned far-field correction: training pairs generated by
// exact direct sums on the GPU, hand-written training kernels (forward +
// back-propagation in FP32 with the network's BF16 rounding, Adam), MFMA
// inference (nb_mlp.h), and the C ABI around them.
//
// Pair generation: sample bodies walk the tree with the per-body opening test
// at theta (cells holding the body itself are never accepted); every accepted
// internal cell with at most kPairMaxBodies bodies becomes a pair. Its inputs
// include the cell's third moments (the tree keeps them for the learned far
// field); its target is the exact FP64 Plummer acceleration of the cell's
// bodies at the sample body minus the cell's monopole, along (u, w) in units
// of M s^2 r / R^5 (the network's output frame, nb_mlp.h), less the softened
// quadrupole term that the apply adds exactly, over s/R.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "nb_device.h"
#include "nb_internal.h"
#include "nb_mlp.h"

namespace nb {

constexpr uint32_t kPairMaxBodies = 8192;
constexpr int kTrainPerBlock = 64;  // samples per training workgroup

__device__ __forceinline__ uint32_t hash_u32(uint64_t a) {
    a ^= a >> 33;
    a *= 0xff51afd7ed558ccdull;
    a ^= a >> 33;
    a *= 0xc4ceb9fe1a85ec53ull;
    a ^= a >> 33;
    return (uint32_t)a;
}

// ---------------------------------------------------------------------------
// Pair collection: wave-cooperative masked walk (as the EXACT traversal), in
// two passes so the pair order is deterministic: COUNT stores each sample's
// number of pairs, an exclusive scan turns them into offsets, WRITE stores
// the pairs of sample k from offset[k] on (up to cap pairs in all).
// ---------------------------------------------------------------------------
template <bool WRITE>
__global__ __launch_bounds__(64) void k_mlp_collect(const NodeRec* __restrict__ nodes,
                                                    const double2* __restrict__ pos, uint32_t n,
                                                    uint32_t samples, uint64_t seed,
                                                    double theta_sq, uint32_t* __restrict__ counts,
                                                    uint32_t cap,
                                                    uint32_t* __restrict__ pair_body,
                                                    uint32_t* __restrict__ pair_slot) {
    __shared__ int32_t s_node[kStackDepth];
    __shared__ uint64_t s_mask[kStackDepth];
    const int l = lane_id();
    const uint32_t k = blockIdx.x * 64 + l;
    const bool valid = k < samples && n > 1;
    const int32_t me = valid ? (int32_t)(hash_u32(seed * 0x9e3779b97f4a7c15ull + k) % n) : -1;
    const double2 p = valid ? pos[me] : make_double2(0.0, 0.0);
    uint32_t at = (WRITE && valid) ? counts[k] : 0u;
    const uint64_t act = ballot64(valid);
    if (act == 0) return;
    int sp = 1;
    s_node[0] = 0;
    s_mask[0] = act;
    while (sp > 0) {
        --sp;
        const int32_t node = readfirst(s_node[sp]);
        const uint64_t m = readfirst64(s_mask[sp]);
        const NodeRec* r = nodes + node;
        const int32_t c0 = r->c0, bs = r->bs, be = r->be;
        if (c0 < 0) continue;  // leaves are exact already
        const bool mine = (m >> l) & 1ull;
        const double dx = r->cx - p.x, dy = r->cy - p.y;
        const double d2 = fma(dx, dx, dy * dy);
        const bool inside = me >= bs && me <= be;
        const bool accept = mine && !inside && (r->wsq < theta_sq * d2);
        if (accept && (uint32_t)(be - bs + 1) <= kPairMaxBodies) {
            if (WRITE && at < cap) {
                pair_body[at] = (uint32_t)me;
                pair_slot[at] = (uint32_t)node;
            }
            at++;
        }
        const uint64_t open = ballot64(mine && !accept);
        if (open && sp + 4 <= kStackDepth) {
            const int32_t cm = r->mask;
#pragma unroll
            for (int q = 3; q >= 0; q--)
                if ((cm >> q) & 1) {
                    s_node[sp] = c0 + q;
                    s_mask[sp] = open;
                    sp++;
                }
        }
    }
    if (!WRITE && valid) counts[k] = at;
}

// Features and exact-minus-monopole targets, one wave per pair.
__global__ __launch_bounds__(256) void k_mlp_targets(
    const NodeRec* __restrict__ nodes, const float4* __restrict__ mom3,
    const double2* __restrict__ pos,
    const double* __restrict__ mass, const double* __restrict__ box, double eps2,
    const uint32_t* __restrict__ pair_body, const uint32_t* __restrict__ pair_slot,
    const uint32_t* __restrict__ total, uint64_t cap, float* __restrict__ X,
    float* __restrict__ T) {
    const uint64_t npairs = *total < cap ? *total : cap;
    const int l = lane_id();
    const float inv_mroot = (float)(1.0 / nodes[0].m);
    const float eps2f = (float)eps2;
    const float inv_ell = mlp_inv_ell(eps2f, box[4]);
    for (uint64_t q = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6); q < npairs;
         q += (uint64_t)gridDim.x * 4) {
        const double2 x = pos[pair_body[q]];
        const NodeRec* r = &nodes[pair_slot[q]];
        double ax = 0.0, ay = 0.0;
        for (int32_t j = r->bs + l; j <= r->be; j += 64) {
            const double2 pj = pos[j];
            const double dx = pj.x - x.x, dy = pj.y - x.y;
            const double r2 = dx * dx + dy * dy + eps2;
            const double w = mass[j] / (r2 * sqrt(r2));
            ax += dx * w;
            ay += dy * w;
        }
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) {
            ax += __shfl_xor(ax, d, 64);
            ay += __shfl_xor(ay, d, 64);
        }
        if (l == 0) {
            const double dx = r->cx - x.x, dy = r->cy - x.y;
            const double r2 = dx * dx + dy * dy + eps2;
            const double w = r->m / (r2 * sqrt(r2));
            float f[kMlpIn], ux, uy, qd[2];
            const float scale = mlp_features((float)dx, (float)dy, (float)sqrt(r->wsq), r->mf,
                                             r->ixx, r->ixy, r->iyy, mom3[pair_slot[q]], eps2f,
                                             inv_mroot, inv_ell, f, ux, uy, qd);
#pragma unroll
            for (int i = 0; i < kMlpIn; i++) X[q * kMlpIn + i] = f[i];
            // exact - monopole along (u, w), w = (-uy, ux), less the quadrupole
            // term, over s/R (the octupole's scale in these units)
            const double cx = ax - dx * w, cy = ay - dy * w;
            T[q * 2 + 0] = (float)(((cx * ux + cy * uy) / (double)scale - (double)qd[0]) / (double)f[2]);
            T[q * 2 + 1] = (float)(((cy * ux - cx * uy) / (double)scale - (double)qd[1]) / (double)f[2]);
        }
    }
}

// ---------------------------------------------------------------------------
// Training: gradient of the mean squared error over a minibatch, FP32 with
// the inference path's BF16 rounding of inputs, weights and hidden
// activations (straight-through for the rounding). One workgroup per 64
// samples writes its partial gradient; k_mlp_adam sums the partials.
// ---------------------------------------------------------------------------
struct MlpOffsets {
    static constexpr int W1 = 0, b1 = W1 + kMlpHidden * kMlpIn, W2 = b1 + kMlpHidden,
                         b2 = W2 + kMlpHidden * kMlpHidden, W3 = b2 + kMlpHidden,
                         b3 = W3 + kMlpOut * kMlpHidden;
};
static_assert(MlpOffsets::b3 + kMlpOut == kMlpParams, "parameter layout");

__global__ __launch_bounds__(256) void k_mlp_grad(const float* __restrict__ w,
                                                  const float* __restrict__ X,
                                                  const float* __restrict__ T, uint64_t npairs,
                                                  const uint32_t* __restrict__ explicit_idx,
                                                  uint32_t batch, uint64_t seed, uint32_t step,
                                                  float* __restrict__ gpart,
                                                  float* __restrict__ loss_part) {
    using O = MlpOffsets;
    __shared__ float sw[kMlpParams];
    __shared__ float sx[kTrainPerBlock][kMlpIn], st[kTrainPerBlock][2];
    __shared__ float sh1[kTrainPerBlock][65], sh2[kTrainPerBlock][65], sdh2[kTrainPerBlock][65];
    __shared__ float sdy[kTrainPerBlock][2];
    __shared__ float sloss[256 / 64];
    const int t = threadIdx.x;
    for (int i = t; i < kMlpParams; i += 256) {
        const bool bias = (i >= O::b1 && i < O::W2) || (i >= O::b2 && i < O::W3) || i >= O::b3;
        sw[i] = bias ? w[i] : bf16_round(w[i]);
    }
    const uint32_t base = blockIdx.x * kTrainPerBlock;
    for (int i = t; i < kTrainPerBlock; i += 256) {
        const uint32_t b = base + i;
        bool ok = b < batch;
        uint64_t src = 0;
        if (ok)
            src = explicit_idx ? explicit_idx[b]
                               : hash_u32(seed * 0x9e3779b97f4a7c15ull + (uint64_t)step * 0x100000001b3ull + b) % npairs;
#pragma unroll
        for (int k = 0; k < kMlpIn; k++) sx[i][k] = ok ? bf16_round(X[src * kMlpIn + k]) : 0.f;
        st[i][0] = ok ? T[src * 2 + 0] : 0.f;
        st[i][1] = ok ? T[src * 2 + 1] : 0.f;
    }
    __syncthreads();
    const int s = t >> 2, j0 = (t & 3) * 16;
    // layer 1
    for (int j = j0; j < j0 + 16; j++) {
        float a = sw[O::b1 + j];
#pragma unroll
        for (int k = 0; k < kMlpIn; k++) a = fmaf(sw[O::W1 + j * kMlpIn + k], sx[s][k], a);
        sh1[s][j] = bf16_round(fmaxf(a, 0.f));
    }
    __syncthreads();
    // layer 2
    for (int j = j0; j < j0 + 16; j++) {
        float a = sw[O::b2 + j];
        for (int k = 0; k < kMlpHidden; k++) a = fmaf(sw[O::W2 + j * 64 + k], sh1[s][k], a);
        sh2[s][j] = bf16_round(fmaxf(a, 0.f));
    }
    __syncthreads();
    // layer 3, loss, dy
    float lsum = 0.f;
    if (t < 2 * kTrainPerBlock) {
        const int ss = t >> 1, o = t & 1;
        float a = sw[O::b3 + o];
        for (int k = 0; k < kMlpHidden; k++) a = fmaf(sw[O::W3 + o * 64 + k], sh2[ss][k], a);
        const bool ok = base + ss < batch;
        const float e = ok ? a - st[ss][o] : 0.f;
        lsum = e * e;
        sdy[ss][o] = 2.f * e / (float)batch;
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) lsum += __shfl_xor(lsum, d, 64);
    if (lane_id() == 0) sloss[t >> 6] = lsum;
    __syncthreads();
    // dh2 = (W3^T dy) * [h2 > 0]
    for (int j = j0; j < j0 + 16; j++) {
        const float g = sw[O::W3 + j] * sdy[s][0] + sw[O::W3 + 64 + j] * sdy[s][1];
        sdh2[s][j] = sh2[s][j] > 0.f ? g : 0.f;
    }
    __syncthreads();
    float* gp = gpart + (uint64_t)blockIdx.x * kMlpParams;
    // dW3, db3 (threads 0..127: o = t/64, j = t%64)
    if (t < 128) {
        const int o = t >> 6, j = t & 63;
        float a = 0.f;
        for (int ss = 0; ss < kTrainPerBlock; ss++) a = fmaf(sdy[ss][o], sh2[ss][j], a);
        gp[O::W3 + o * 64 + j] = a;
    } else if (t < 130) {
        const int o = t - 128;
        float a = 0.f;
        for (int ss = 0; ss < kTrainPerBlock; ss++) a += sdy[ss][o];
        gp[O::b3 + o] = a;
    }
    // dW2 (16 outputs per thread), db2
    for (int q = 0; q < 16; q++) {
        const int idx = t * 16 + q, j = idx >> 6, k = idx & 63;
        float a = 0.f;
        for (int ss = 0; ss < kTrainPerBlock; ss++) a = fmaf(sdh2[ss][j], sh1[ss][k], a);
        gp[O::W2 + idx] = a;
    }
    if (t < 64) {
        float a = 0.f;
        for (int ss = 0; ss < kTrainPerBlock; ss++) a += sdh2[ss][t];
        gp[O::b2 + t] = a;
    }
    __syncthreads();
    // dh1 = (W2^T dh2) * [h1 > 0], stored over sh2
    for (int k = j0; k < j0 + 16; k++) {
        float g = 0.f;
        for (int j = 0; j < kMlpHidden; j++) g = fmaf(sw[O::W2 + j * 64 + k], sdh2[s][j], g);
        sh2[s][k] = sh1[s][k] > 0.f ? g : 0.f;
    }
    __syncthreads();
    // dW1 (kMlpIn * 64 / 256 outputs per thread), db1
    static_assert(kMlpHidden * kMlpIn % 256 == 0, "dW1 split");
    for (int q = 0; q < kMlpHidden * kMlpIn / 256; q++) {
        const int idx = t * (kMlpHidden * kMlpIn / 256) + q, j = idx / kMlpIn, k = idx % kMlpIn;
        float a = 0.f;
        for (int ss = 0; ss < kTrainPerBlock; ss++) a = fmaf(sh2[ss][j], sx[ss][k], a);
        gp[O::W1 + idx] = a;
    }
    if (t < 64) {
        float a = 0.f;
        for (int ss = 0; ss < kTrainPerBlock; ss++) a += sh2[ss][t];
        gp[O::b1 + t] = a;
    }
    if (t == 0) loss_part[blockIdx.x] = sloss[0] + sloss[1] + sloss[2] + sloss[3];
}

// Sums the partial gradients (fixed order); Adam update of the FP32 master
// weights when lr > 0, else writes the summed gradient to `grad_out`.
__global__ __launch_bounds__(256) void k_mlp_adam(float* __restrict__ w, float* __restrict__ m,
                                                  float* __restrict__ v,
                                                  const float* __restrict__ gpart,
                                                  uint32_t parts, float lr, float beta1,
                                                  float beta2, float eps, uint32_t step,
                                                  float* __restrict__ grad_out) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= (uint32_t)kMlpParams) return;
    float g = 0.f;
    for (uint32_t b = 0; b < parts; b++) g += gpart[(uint64_t)b * kMlpParams + i];
    if (grad_out) grad_out[i] = g;
    if (lr <= 0.f) return;
    const float mi = beta1 * m[i] + (1.f - beta1) * g;
    const float vi = beta2 * v[i] + (1.f - beta2) * g * g;
    m[i] = mi;
    v[i] = vi;
    const float mh = mi / (1.f - powf(beta1, (float)step));
    const float vh = vi / (1.f - powf(beta2, (float)step));
    w[i] -= lr * mh / (sqrtf(vh) + eps);
}

// Batch inference on MFMA: each wave loads the weights once and runs 64-row
// batches of X as four interleaved tiles.
__global__ __launch_bounds__(256) void k_mlp_forward(const float* __restrict__ X, uint64_t count,
                                                     const void* __restrict__ wt,
                                                     const float* __restrict__ bias,
                                                     float* __restrict__ Y) {
    __shared__ __attribute__((aligned(16))) float sb[kMlpBiasLen];
    for (int i = threadIdx.x; i < kMlpBiasLen; i += 256) sb[i] = bias[i];
    __syncthreads();
    const int l = lane_id();
    MlpWeights w;
    mlp_load_weights(w, wt, l);
    const uint64_t batches = (count + 63) / 64;
    const uint64_t wave = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const uint64_t waves = (uint64_t)gridDim.x * 4;
    const int part = l >> 4;  // 0: f0..f3, 1: f4..f7, 2: f8..f11, 3: zero padding
    for (uint64_t bt = wave; bt < batches; bt += waves) {
        f32x4 x[4], y[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const uint64_t row = bt * 64 + 16 * q + (l & 15);
            x[q] = f32x4{0.f, 0.f, 0.f, 0.f};
            if (row < count && part < kMlpIn / 4) {
                x[q].x = X[row * kMlpIn + part * 4 + 0];
                x[q].y = X[row * kMlpIn + part * 4 + 1];
                x[q].z = X[row * kMlpIn + part * 4 + 2];
                x[q].w = X[row * kMlpIn + part * 4 + 3];
            }
        }
        mlp_forward_tiles<2>(x, w, sb, l, y);
        mlp_forward_tiles<2>(x + 2, w, sb, l, y + 2);
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const uint64_t row = bt * 64 + 16 * q + l;
            if (l < 16 && row < count) {
                Y[row * 2 + 0] = y[q].x;
                Y[row * 2 + 1] = y[q].y;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Far-field network in the step: one wavefront per group of 64 bodies. For
// each finished subgroup the accepted cells (tree slots written by the walk,
// possibly in another group's region when another wavefront walked it)
// go through the network 64 at a time at the subgroup's box centre (lane l
// prepares cell l; tile q of the MFMA pass is cells 16q..16q+15). A subgroup
// ends with C = the sum of the network's terms (s/R times its output, in
// acceleration units) and of the far cells' softened quadrupole terms at the
// centre, and J = the sum of those terms' Jacobians; every member at offset b
// from the centre adds C - J b to the walk's acceleration (which already
// holds the near cells' quadrupole terms at the member's own position), and
// kicks (or stores G a). A cell is far when its COM lies beyond the
// subgroup's threshold from the centre, decided in FP64 exactly as the walk
// decided it. The sums run in a fixed order.
// ---------------------------------------------------------------------------
constexpr int kApplyWaves = 4;

template <bool KICK>
__global__ __launch_bounds__(64 * kApplyWaves) void k_mlp_apply(
    const uint32_t* __restrict__ cell_slot, const uint32_t* __restrict__ cell_cnt,
    uint32_t cell_cap, const SubRec* __restrict__ subs, const float4* __restrict__ body_w,
    double2* __restrict__ acc, double2* __restrict__ vel, uint32_t n,
    const StepConsts* __restrict__ k, uint32_t* __restrict__ meta,
    const NodeRec* __restrict__ nodes, const float4* __restrict__ mom3,
    const double* __restrict__ box,
    const void* __restrict__ tiles, const float* __restrict__ bias, int count_step,
    uint32_t groups, const uint32_t* __restrict__ grp_first,
    const uint32_t* __restrict__ grp_info) {
    if (meta[META_FLAGS] & kHaltFlags) return;
    __shared__ __attribute__((aligned(16))) float sb[kMlpBiasLen];
    __shared__ float su[kApplyWaves][6 * 64];        // [C_x, C_y, J00, J01, J10, J11][subgroup]
    __shared__ __attribute__((aligned(16))) float sf[kApplyWaves][64 * kMlpIn];  // cell inputs
    __shared__ float2 sr[kApplyWaves][64];           // cell scale * u
    __shared__ float4 sq[kApplyWaves][64];           // far cell's quadrupole term (units), s/R
    for (int i = threadIdx.x; i < kMlpBiasLen; i += 64 * kApplyWaves) sb[i] = bias[i];
    if (count_step && blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&meta[META_STEPS_DONE], 1u);
    __syncthreads();
    const int l = lane_id(), wv = threadIdx.x >> 6;
    const uint32_t g = blockIdx.x * kApplyWaves + wv;
    if (g >= groups) return;
    const uint32_t first = grp_first[g], ginfo = grp_info[g];
    const uint32_t gn = ginfo & 255u, sys = ginfo >> 8;
    float* u = su[wv];
    float* f_all = sf[wv];
    const float eps2f = (float)k->eps2;
    const float inv_mroot = (float)(1.0 / nodes[sys].m);
    const float inv_ell = mlp_inv_ell(eps2f, box[(uint64_t)kBoxStride * sys + 4]);
    MlpWeights w;
    mlp_load_weights(w, tiles, l);
    const int nsub = (int)min(cell_cnt[g], 64u);
    const uint64_t slots = (uint64_t)groups * cell_cap;
    const int col = l & 15;
    // Lane k holds subgroup k's record; the members' own data, needed only at
    // the end, load now. Every subgroup's sums start at zero (one without
    // cells keeps C = J = 0).
    SubRec mys{0.0, 0.0, 0u, 0u, 0.0};
    if (l < nsub) mys = subs[(uint64_t)g * 64 + l];
    const bool mem = (uint32_t)l < gn;
    const uint32_t i = first + (mem ? (uint32_t)l : 0u);
    const float4 bw = mem ? body_w[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    const double2 a0 = mem ? acc[i] : make_double2(0.0, 0.0);
    for (int e = l; e < 6 * 64; e += 64) u[e] = 0.f;
    auto sub_begin = [&](int kk) -> uint64_t {
        return (uint64_t)(uint32_t)__shfl((int)mys.start, kk, 64);
    };
    auto sub_end = [&](int kk) -> uint64_t {
        return min(sub_begin(kk) + (uint32_t)__shfl((int)mys.count, kk, 64), slots);
    };
    // The batches of 64 cells of all subgroups, in order, as one sequence:
    // the cells' slot indices are loaded two batches ahead and their records
    // one batch ahead, across subgroup boundaries, so the subgroup records,
    // slot indices and tree records are never waited for in turn.
    struct It {
        int ks;
        uint64_t b0, end;
    };
    auto seek = [&](int kk) -> It {  // the first batch of the first non-empty subgroup >= kk
        for (; kk < nsub; kk++) {
            const uint64_t b = sub_begin(kk), e = sub_end(kk);
            if (b < e) return It{kk, b, e};
        }
        return It{nsub, 0, 0};
    };
    auto advance = [&](It t) -> It {
        if (t.ks >= nsub) return t;
        t.b0 += 64;
        return t.b0 < t.end ? t : seek(t.ks + 1);
    };
    auto load_slot = [&](const It& t) -> uint32_t {
        const uint64_t cell = t.b0 + l;
        return (t.ks < nsub && cell < t.end) ? cell_slot[cell] : 0xffffffffu;
    };
    // raw fields of a cell's records, converted when used
    struct CellIn {
        double cx, cy, wsq;
        float mf, ixx, ixy, iyy;
        float4 t3;
        bool ok;
    };
    auto load_cell = [&](uint32_t slot) {
        CellIn c{0.0, 0.0, 1.0, 1.f, 0.f, 0.f, 0.f, make_float4(0.f, 0.f, 0.f, 0.f), false};
        if (slot != 0xffffffffu) {
            const NodeRec* r = nodes + slot;
            c.t3 = mom3[slot];
            c.cx = r->cx;
            c.cy = r->cy;
            c.wsq = r->wsq;
            c.mf = r->mf;
            c.ixx = r->ixx;
            c.ixy = r->ixy;
            c.iyy = r->iyy;
            c.ok = true;
        }
        return c;
    };
    It cur = seek(0);
    It nxt = advance(cur);
    CellIn rec = load_cell(load_slot(cur));
    uint32_t slot_n = load_slot(nxt);
    float cx = 0.f, cy = 0.f, j00 = 0.f, j01 = 0.f, j10 = 0.f, j11 = 0.f;
    while (cur.ks < nsub) {
        const CellIn c = rec;
        const It nn = advance(nxt);
        rec = load_cell(slot_n);      // the next batch's records
        slot_n = load_slot(nn);       // and the one after's slot indices
        const int ks = cur.ks;
        const double ox = __shfl(mys.ox, ks, 64), oy = __shfl(mys.oy, ks, 64);
        const double thr = __shfl(mys.thr, ks, 64);
        const bool ok = c.ok;
        float dx = 1.f, dy = 0.f, s = 1.f, M = 1.f, ixx = 0.f, ixy = 0.f, iyy = 0.f;
        float4 t3 = make_float4(0.f, 0.f, 0.f, 0.f);
        if (ok) {
            t3 = c.t3;
            dx = (float)(c.cx - ox);
            dy = (float)(c.cy - oy);
            s = sqrtf((float)c.wsq);
            M = c.mf;
            ixx = c.ixx;
            ixy = c.ixy;
            iyy = c.iyy;
        }
        float f[kMlpIn], ux, uy, qd[2];
        float scale = mlp_features(dx, dy, s, M, ixx, ixy, iyy, t3, eps2f, inv_mroot, inv_ell, f,
                                   ux, uy, qd);
        scale = ok ? scale : 0.f;
        // A far cell (the walk's rule, in FP64 on the same values) adds its
        // quadrupole term here, at the centre, and its Jacobian moves the
        // term to each member; a near cell's term every member added itself.
        bool far = false;
        if (ok) {
            const double ex = c.cx - ox, ey = c.cy - oy;
            far = fma(ex, ex, ey * ey) > thr;
        }
        if (far) {
            float J[4];
            quad_jacobian(dx, dy, ixx, ixy, iyy, eps2f, J);
            j00 += J[0];
            j01 += J[1];
            j10 += J[2];
            j11 += J[3];
        }
        sq[wv][l] = make_float4(far ? qd[0] : 0.f, far ? qd[1] : 0.f, f[2], 0.f);
#pragma unroll
        for (int h = 0; h < kMlpIn / 4; h++)
            *reinterpret_cast<float4*>(f_all + kMlpIn * l + 4 * h) =
                make_float4(f[4 * h], f[4 * h + 1], f[4 * h + 2], f[4 * h + 3]);
        sr[wv][l] = make_float2(scale * ux, scale * uy);
        __builtin_amdgcn_wave_barrier();
        f32x4 x[4], y[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            x[q] = f32x4{0.f, 0.f, 0.f, 0.f};
            if (l < 16 * (kMlpIn / 4)) {
                const float4 v = *reinterpret_cast<const float4*>(
                    f_all + kMlpIn * (16 * q + col) + 4 * (l >> 4));
                x[q] = f32x4{v.x, v.y, v.z, v.w};
            }
        }
        mlp_forward_tiles<2>(x, w, sb, l, y);
        mlp_forward_tiles<2>(x + 2, w, sb, l, y + 2);
        // lanes 0-15: cells 16q + l, rotated back to (x, y)
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const float2 su_ = sr[wv][(16 * q + l) & 63];
            const float4 qq = sq[wv][(16 * q + l) & 63];
            if (l < 16) {
                // (s/R) * output along (u, w), plus a far cell's quadrupole term
                const float yu = fmaf(y[q].x, qq.z, qq.x), yw = fmaf(y[q].y, qq.z, qq.y);
                cx += yu * su_.x - yw * su_.y;
                cy += yu * su_.y + yw * su_.x;
            }
        }
        __builtin_amdgcn_wave_barrier();
        if (nxt.ks != ks) {  // the subgroup's last batch: its sums, in a fixed order
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                cx += __shfl_xor(cx, d, 64);
                cy += __shfl_xor(cy, d, 64);
                j00 += __shfl_xor(j00, d, 64);
                j01 += __shfl_xor(j01, d, 64);
                j10 += __shfl_xor(j10, d, 64);
                j11 += __shfl_xor(j11, d, 64);
            }
            if (l == 0) {
                u[0 * 64 + ks] = cx;
                u[1 * 64 + ks] = cy;
                u[2 * 64 + ks] = j00;
                u[3 * 64 + ks] = j01;
                u[4 * 64 + ks] = j10;
                u[5 * 64 + ks] = j11;
            }
            cx = cy = j00 = j01 = j10 = j11 = 0.f;
        }
        cur = nxt;
        nxt = nn;
    }
    __builtin_amdgcn_wave_barrier();
    if (!mem) return;
    const int ks = __float_as_int(bw.z) & 63;
    // the body sits at b = (bw.x, bw.y) from the centre: d = COM - point moves
    // by -b, so a far cell's term there is Q - J b
    const float ax = u[0 * 64 + ks] - fmaf(u[2 * 64 + ks], bw.x, u[3 * 64 + ks] * bw.y);
    const float ay = u[1 * 64 + ks] - fmaf(u[4 * 64 + ks], bw.x, u[5 * 64 + ks] * bw.y);
    const double2 a = make_double2(a0.x + (double)ax, a0.y + (double)ay);
    if (KICK) {
        const double gdt = k->G_dt;
        double2 v = vel[i];
        v.x += a.x * gdt;
        v.y += a.y * gdt;
        vel[i] = v;
    } else {
        acc[i] = make_double2(a.x * k->G, a.y * k->G);
    }
}

void launch_mlp_apply(nb_ctx* c, int src, bool kick, bool count_step) {
    const uint32_t n = (uint32_t)c->n;
    const uint32_t groups = c->ngroups;
    const uint32_t blocks = (groups + kApplyWaves - 1) / kApplyWaves;
    const int count = count_step ? 1 : 0;
#define NB_APPLY(K)                                                                            \
    hipLaunchKernelGGL((k_mlp_apply<K>), dim3(blocks), dim3(64 * kApplyWaves), 0, c->stream,   \
                       c->b.cell_slot, c->b.cell_cnt, c->cell_cap, c->b.subs, c->b.body_w,     \
                       c->b.acc, c->b.vel[src], n, c->b.consts, c->b.meta, c->b.nodes,         \
                       c->b.mom3, c->b.box, c->b.mlp_tiles, c->b.mlp_bias, count, groups,      \
                       c->b.grp_first, c->b.grp_info)
    if (kick) NB_APPLY(true); else NB_APPLY(false);
#undef NB_APPLY
}

}  // namespace nb

using namespace nb;

namespace {

// Copy on the context's stream, which does not synchronise with the null
// stream, then wait for it.
hipError_t scopy(nb_ctx* c, void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
    const hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, c->stream);
    return e == hipSuccess ? hipStreamSynchronize(c->stream) : e;
}

int mlp_fail(nb_ctx* c, int code, const std::string& m) {
    c->err = m;
    return code;
}

#define MLP_HIP(c, expr)                                                                  \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            return mlp_fail((c), e_ == hipErrorOutOfMemory ? NB_ERR_OUT_OF_MEMORY : NB_ERR_HIP, \
                            std::string(#expr) + ": " + hipGetErrorString(e_));           \
    } while (0)

// Packs the FP32 canonical parameters into the MFMA A tiles (BF16, round to
// nearest even) in the lane layouts of mlp_forward_tiles, and the compact
// bias vector.
void pack_tiles(const float* p, std::vector<uint8_t>& tiles, std::vector<float>& bias) {
    using O = MlpOffsets;
    tiles.assign(kMlpTileBytes, 0);
    bias.assign(kMlpBiasLen, 0.f);
    uint16_t* t1 = reinterpret_cast<uint16_t*>(tiles.data());                 // [4][64][4]
    uint16_t* t23 = t1 + kMlpTiles1 * 64 * 4;                                   // [10][64][8]
    for (int t = 0; t < kMlpTiles1; t++)
        for (int l = 0; l < 64; l++)
            for (int v = 0; v < 4; v++) {
                const int m = l % 16, kk = 4 * (l / 16) + v;
                const float x = kk < kMlpIn ? p[O::W1 + (16 * t + m) * kMlpIn + kk] : 0.f;
                t1[(t * 64 + l) * 4 + v] = f32_to_bf16(x);
            }
    for (int t = 0; t < kMlpTiles23; t++)
        for (int l = 0; l < 64; l++)
            for (int v = 0; v < 8; v++) {
                const int m = l % 16, j = 8 * (l / 16) + v;
                float x;
                if (t < 8) {  // layer 2: output tile t/2, K block t%2
                    x = p[O::W2 + (16 * (t / 2) + m) * 64 + mlp_k_neuron(t % 2, j)];
                } else {      // layer 3: K block t-8
                    x = m < 2 ? p[O::W3 + m * 64 + mlp_k_neuron(t - 8, j)] : 0.f;
                }
                t23[(t * 64 + l) * 8 + v] = f32_to_bf16(x);
            }
    for (int i = 0; i < 64; i++) bias[i] = p[O::b1 + i];
    for (int i = 0; i < 64; i++) bias[64 + i] = p[O::b2 + i];
    bias[128] = p[O::b3];
    bias[129] = p[O::b3 + 1];
}

int ensure_mlp(nb_ctx* c) {
    Buffers& b = c->b;
    if (b.mlp_w) return NB_OK;
    MLP_HIP(c, hipMalloc((void**)&b.mlp_w, kMlpParams * sizeof(float)));
    MLP_HIP(c, hipMalloc((void**)&b.mlp_m, kMlpParams * sizeof(float)));
    MLP_HIP(c, hipMalloc((void**)&b.mlp_v, kMlpParams * sizeof(float)));
    MLP_HIP(c, hipMalloc((void**)&b.mlp_tiles, kMlpTileBytes));
    MLP_HIP(c, hipMalloc((void**)&b.mlp_bias, kMlpBiasLen * sizeof(float)));
    MLP_HIP(c, hipMemsetAsync(b.mlp_w, 0, kMlpParams * sizeof(float), c->stream));
    MLP_HIP(c, hipMemsetAsync(b.mlp_tiles, 0, kMlpTileBytes, c->stream));
    MLP_HIP(c, hipMemsetAsync(b.mlp_bias, 0, kMlpBiasLen * sizeof(float), c->stream));
    MLP_HIP(c, hipStreamSynchronize(c->stream));
    c->device_bytes += kMlpParams * 3 * sizeof(float) + kMlpTileBytes +
                       kMlpBiasLen * sizeof(float);
    return NB_OK;
}

int upload_weights(nb_ctx* c, const float* p) {
    std::vector<uint8_t> tiles;
    std::vector<float> bias;
    pack_tiles(p, tiles, bias);
    MLP_HIP(c, scopy(c, c->b.mlp_w, p, kMlpParams * sizeof(float), hipMemcpyHostToDevice));
    MLP_HIP(c, scopy(c, c->b.mlp_tiles, tiles.data(), tiles.size(),
                         hipMemcpyHostToDevice));
    MLP_HIP(c, scopy(c, c->b.mlp_bias, bias.data(), bias.size() * sizeof(float),
                         hipMemcpyHostToDevice));
    return NB_OK;
}

int ensure_pairs(nb_ctx* c, uint64_t cap) {
    Buffers& b = c->b;
    if (c->pair_cap >= cap) return NB_OK;
    for (void* p : {(void*)b.pair_body, (void*)b.pair_slot, (void*)b.pair_x, (void*)b.pair_t})
        if (p) (void)hipFree(p);
    b.pair_body = b.pair_slot = nullptr;
    b.pair_x = b.pair_t = nullptr;
    c->pair_cap = 0;
    MLP_HIP(c, hipMalloc((void**)&b.pair_body, cap * sizeof(uint32_t)));
    MLP_HIP(c, hipMalloc((void**)&b.pair_slot, cap * sizeof(uint32_t)));
    MLP_HIP(c, hipMalloc((void**)&b.pair_x, cap * kMlpIn * sizeof(float)));
    MLP_HIP(c, hipMalloc((void**)&b.pair_t, cap * 2 * sizeof(float)));
    c->pair_cap = cap;
    return NB_OK;
}

int ensure_train(nb_ctx* c, uint32_t parts) {
    Buffers& b = c->b;
    if (c->train_parts >= parts) return NB_OK;
    if (b.mlp_gpart) (void)hipFree(b.mlp_gpart);
    if (b.mlp_loss) (void)hipFree(b.mlp_loss);
    b.mlp_gpart = b.mlp_loss = nullptr;
    MLP_HIP(c, hipMalloc((void**)&b.mlp_gpart, (uint64_t)parts * kMlpParams * sizeof(float)));
    MLP_HIP(c, hipMalloc((void**)&b.mlp_loss, parts * sizeof(float)));
    c->train_parts = parts;
    return NB_OK;
}

}  // namespace

namespace nb {

void free_mlp(nb_ctx* c) {
    Buffers& b = c->b;
    for (void* p : {(void*)b.mlp_w, (void*)b.mlp_m, (void*)b.mlp_v, (void*)b.mlp_tiles,
                    (void*)b.mlp_bias, (void*)b.mlp_gpart, (void*)b.mlp_loss, (void*)b.pair_body,
                    (void*)b.pair_slot, (void*)b.pair_x, (void*)b.pair_t, (void*)b.sample_off,
                    (void*)b.cell_slot, (void*)b.cell_cnt, (void*)b.subs, (void*)b.body_w})
        if (p) (void)hipFree(p);
    b.cell_slot = nullptr;
    b.cell_cnt = nullptr;
    b.subs = nullptr;
    b.body_w = nullptr;
    c->cell_cap = 0;
    b.mlp_w = b.mlp_m = b.mlp_v = b.mlp_gpart = b.mlp_loss = b.pair_x = b.pair_t = nullptr;
    b.mlp_tiles = nullptr;
    b.mlp_bias = nullptr;
    b.pair_body = b.pair_slot = nullptr;
    b.sample_off = nullptr;
}

int prepare_mlp(nb_ctx* c) { return ensure_mlp(c); }

int alloc_cells(nb_ctx* c, uint32_t cap) {
    Buffers& b = c->b;
    const uint64_t groups = c->ngroups;
    if (b.cell_slot) (void)hipFree(b.cell_slot);
    b.cell_slot = nullptr;
    if (c->cell_cap) c->device_bytes -= groups * c->cell_cap * sizeof(uint32_t);
    c->cell_cap = 0;
    MLP_HIP(c, hipMalloc((void**)&b.cell_slot, groups * cap * sizeof(uint32_t)));
    if (!b.cell_cnt) {
        MLP_HIP(c, hipMalloc((void**)&b.cell_cnt, groups * sizeof(uint32_t)));
        MLP_HIP(c, hipMalloc((void**)&b.subs, groups * 64 * sizeof(SubRec)));
        MLP_HIP(c, hipMalloc((void**)&b.body_w, c->n * sizeof(float4)));
        c->device_bytes += groups * (sizeof(uint32_t) + 64 * sizeof(SubRec)) + c->n * sizeof(float4);
    }
    c->device_bytes += groups * cap * sizeof(uint32_t);
    c->cell_cap = cap;
    return NB_OK;
}

// The traversal groups changed (nb_autotune): the per-group cell regions,
// subgroup records and counters are allocated again for the new groups.
int regroup_cells(nb_ctx* c, uint64_t old_groups) {
    Buffers& b = c->b;
    if (!b.cell_cnt) return NB_OK;  // no learned far field
    c->device_bytes -= old_groups * (sizeof(uint32_t) + 64 * sizeof(SubRec) +
                                     (uint64_t)c->cell_cap * sizeof(uint32_t)) +
                       c->n * sizeof(float4);
    for (void* q : {(void*)b.cell_slot, (void*)b.cell_cnt, (void*)b.subs, (void*)b.body_w})
        (void)hipFree(q);
    b.cell_slot = nullptr;
    b.cell_cnt = nullptr;
    b.subs = nullptr;
    b.body_w = nullptr;
    const uint32_t cap = c->cell_cap;
    c->cell_cap = 0;
    return alloc_cells(c, cap);
}

}  // namespace nb

extern "C" {

int nb_mlp_set_weights(nb_ctx* c, const float* params, uint64_t count) {
    if (!c || !params) return NB_ERR_INVALID_ARG;
    if (count != (uint64_t)kMlpParams)
        return mlp_fail(c, NB_ERR_INVALID_ARG, "expected NB_MLP_PARAMS parameters");
    int rc = ensure_mlp(c);
    if (rc) return rc;
    return upload_weights(c, params);
}

int nb_mlp_get_weights(nb_ctx* c, float* params, uint64_t count) {
    if (!c || !params) return NB_ERR_INVALID_ARG;
    if (count != (uint64_t)kMlpParams)
        return mlp_fail(c, NB_ERR_INVALID_ARG, "expected NB_MLP_PARAMS parameters");
    int rc = ensure_mlp(c);
    if (rc) return rc;
    MLP_HIP(c, scopy(c, params, c->b.mlp_w, kMlpParams * sizeof(float), hipMemcpyDeviceToHost));
    return NB_OK;
}

int nb_mlp_init(nb_ctx* c, uint64_t seed) {
    if (!c) return NB_ERR_INVALID_ARG;
    std::vector<float> p(kMlpParams, 0.f);
    std::mt19937_64 rng(seed);
    std::normal_distribution<float> nd(0.f, 1.f);
    using O = MlpOffsets;
    for (int i = 0; i < 64 * kMlpIn; i++) p[O::W1 + i] = nd(rng) * std::sqrt(2.f / kMlpIn);
    for (int i = 0; i < 64 * 64; i++) p[O::W2 + i] = nd(rng) * std::sqrt(2.f / 64.f);
    for (int i = 0; i < 2 * 64; i++) p[O::W3 + i] = nd(rng) * std::sqrt(1.f / 64.f);
    int rc = ensure_mlp(c);
    if (rc) return rc;
    MLP_HIP(c, hipMemsetAsync(c->b.mlp_m, 0, kMlpParams * sizeof(float), c->stream));
    MLP_HIP(c, hipMemsetAsync(c->b.mlp_v, 0, kMlpParams * sizeof(float), c->stream));
    c->train_step = 0;
    return upload_weights(c, p.data());
}

int nb_mlp_generate(nb_ctx* c, double theta, uint32_t sample_bodies, uint64_t max_pairs,
                    uint64_t seed, uint64_t* pairs_out) {
    if (!c || theta <= 0.0 || !std::isfinite(theta)) return NB_ERR_INVALID_ARG;
    if (c->p.algorithm != NB_ALGO_BARNES_HUT)
        return mlp_fail(c, NB_ERR_NOT_SUPPORTED, "pairs need a Barnes-Hut context");
    if (c->n < 2) return mlp_fail(c, NB_ERR_INVALID_ARG, "need at least two bodies");
    if (c->nsys != 1)
        return mlp_fail(c, NB_ERR_NOT_SUPPORTED, "pairs are generated from one system");
    if (max_pairs == 0 || max_pairs > 0xffffffffull)
        return mlp_fail(c, NB_ERR_INVALID_ARG, "max_pairs must be in [1, 2^32)");
    // The scan of per-sample counts uses the context's scan scratch, sized for n.
    const uint32_t samples = (uint32_t)std::min<uint64_t>(sample_bodies, c->n);
    int rc = ensure_mom3(c);  // the pairs' inputs include the cells' third moments
    if (rc) return rc;
    rc = build_tree(c);
    if (rc) return rc;
    rc = ensure_pairs(c, max_pairs);
    if (rc) return rc;
    Buffers& b = c->b;
    if (!b.sample_off) {
        MLP_HIP(c, hipMalloc((void**)&b.sample_off, (c->n + 2) * sizeof(uint32_t)));
        c->device_bytes += (c->n + 2) * sizeof(uint32_t);
    }
    uint32_t* total = b.sample_off + c->n + 1;
    MLP_HIP(c, hipMemsetAsync(total, 0, sizeof(uint32_t), c->stream));
    const uint32_t waves = std::max<uint32_t>(1, (samples + 63) / 64);
    // this call's limit: the buffer may be larger from an earlier call
    const uint32_t cap = (uint32_t)std::min<uint64_t>(c->pair_cap, max_pairs);
    hipLaunchKernelGGL(k_mlp_collect<false>, dim3(waves), dim3(64), 0, c->stream, b.nodes,
                       b.pos[c->cur], (uint32_t)c->n, samples, seed, theta * theta,
                       b.sample_off, cap, b.pair_body, b.pair_slot);
    if (samples) launch_scan_u32(c, b.sample_off, samples, total);
    hipLaunchKernelGGL(k_mlp_collect<true>, dim3(waves), dim3(64), 0, c->stream, b.nodes,
                       b.pos[c->cur], (uint32_t)c->n, samples, seed, theta * theta,
                       b.sample_off, cap, b.pair_body, b.pair_slot);
    hipLaunchKernelGGL(k_mlp_targets, dim3(2048), dim3(256), 0, c->stream, b.nodes, b.mom3,
                       b.pos[c->cur], b.mass[c->cur], b.box,
                       c->p.softening * c->p.softening, b.pair_body, b.pair_slot, total,
                       (uint64_t)cap, b.pair_x, b.pair_t);
    MLP_HIP(c, hipGetLastError());
    uint32_t cnt = 0;
    MLP_HIP(c, scopy(c, &cnt, total, sizeof(cnt), hipMemcpyDeviceToHost));
    c->pair_count = std::min<uint64_t>(cnt, cap);
    if (pairs_out) *pairs_out = c->pair_count;
    return NB_OK;
}

int nb_mlp_get_pairs(nb_ctx* c, float* features, float* targets, uint32_t* body,
                     uint32_t* cell, uint64_t cap, uint64_t* count) {
    if (!c) return NB_ERR_INVALID_ARG;
    const uint64_t k = c->pair_count < cap ? c->pair_count : cap;
    if (count) *count = c->pair_count;
    if (k == 0) return NB_OK;
    if (body)
        MLP_HIP(c, scopy(c, body, c->b.pair_body, k * sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (cell)
        MLP_HIP(c, scopy(c, cell, c->b.pair_slot, k * sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (features)
        MLP_HIP(c, scopy(c, features, c->b.pair_x, k * kMlpIn * sizeof(float),
                         hipMemcpyDeviceToHost));
    if (targets)
        MLP_HIP(c, scopy(c, targets, c->b.pair_t, k * 2 * sizeof(float), hipMemcpyDeviceToHost));
    return NB_OK;
}

int nb_mlp_set_pairs(nb_ctx* c, const float* features, const float* targets, uint64_t count) {
    if (!c || (count && (!features || !targets))) return NB_ERR_INVALID_ARG;
    int rc = ensure_pairs(c, count ? count : 1);
    if (rc) return rc;
    if (count) {
        MLP_HIP(c, scopy(c, c->b.pair_x, features, count * kMlpIn * sizeof(float),
                             hipMemcpyHostToDevice));
        MLP_HIP(c, scopy(c, c->b.pair_t, targets, count * 2 * sizeof(float),
                             hipMemcpyHostToDevice));
    }
    c->pair_count = count;
    return NB_OK;
}

int nb_mlp_train(nb_ctx* c, uint32_t steps, uint32_t batch, float lr, uint64_t seed,
                 float* last_loss) {
    if (!c || batch == 0 || lr <= 0.f) return NB_ERR_INVALID_ARG;
    if (c->pair_count == 0) return mlp_fail(c, NB_ERR_INVALID_ARG, "no training pairs");
    int rc = ensure_mlp(c);
    if (rc) return rc;
    const uint32_t parts = (batch + kTrainPerBlock - 1) / kTrainPerBlock;
    rc = ensure_train(c, parts);
    if (rc) return rc;
    Buffers& b = c->b;
    for (uint32_t s = 0; s < steps; s++) {
        c->train_step++;
        hipLaunchKernelGGL(k_mlp_grad, dim3(parts), dim3(256), 0, c->stream, b.mlp_w, b.pair_x,
                           b.pair_t, (uint64_t)c->pair_count, (const uint32_t*)nullptr, batch,
                           seed, c->train_step, b.mlp_gpart, b.mlp_loss);
        hipLaunchKernelGGL(k_mlp_adam, dim3((kMlpParams + 255) / 256), dim3(256), 0, c->stream,
                           b.mlp_w, b.mlp_m, b.mlp_v, b.mlp_gpart, parts, lr, 0.9f, 0.999f,
                           1e-8f, c->train_step, (float*)nullptr);
    }
    MLP_HIP(c, hipGetLastError());
    std::vector<float> loss(parts);
    MLP_HIP(c, hipMemcpyAsync(loss.data(), b.mlp_loss, parts * sizeof(float),
                              hipMemcpyDeviceToHost, c->stream));
    std::vector<float> w(kMlpParams);
    MLP_HIP(c, hipMemcpyAsync(w.data(), b.mlp_w, kMlpParams * sizeof(float),
                              hipMemcpyDeviceToHost, c->stream));
    MLP_HIP(c, hipStreamSynchronize(c->stream));
    double l = 0.0;
    for (float v : loss) l += v;
    if (last_loss) *last_loss = (float)(l / (double)batch);
    return upload_weights(c, w.data());
}

int nb_mlp_grad(nb_ctx* c, const float* features, const float* targets, uint64_t count,
                float* grad, float* loss_out) {
    if (!c || !grad || count == 0 || count > 0x7fffffffull) return NB_ERR_INVALID_ARG;
    int rc = nb_mlp_set_pairs(c, features, targets, count);
    if (rc) return rc;
    const uint32_t parts = (uint32_t)((count + kTrainPerBlock - 1) / kTrainPerBlock);
    rc = ensure_train(c, parts);
    if (rc) return rc;
    Buffers& b = c->b;
    std::vector<uint32_t> idx(count);
    for (uint64_t i = 0; i < count; i++) idx[i] = (uint32_t)i;
    uint32_t* didx = nullptr;
    MLP_HIP(c, hipMalloc((void**)&didx, count * sizeof(uint32_t)));
    float* dgrad = nullptr;
    hipError_t e = hipMalloc((void**)&dgrad, kMlpParams * sizeof(float));
    if (e == hipSuccess)
        e = scopy(c, didx, idx.data(), count * sizeof(uint32_t), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_mlp_grad, dim3(parts), dim3(256), 0, c->stream, b.mlp_w, b.pair_x,
                           b.pair_t, count, didx, (uint32_t)count, 0ull, 0u, b.mlp_gpart,
                           b.mlp_loss);
        hipLaunchKernelGGL(k_mlp_adam, dim3((kMlpParams + 255) / 256), dim3(256), 0, c->stream,
                           b.mlp_w, b.mlp_m, b.mlp_v, b.mlp_gpart, parts, 0.f, 0.f, 0.f, 0.f,
                           1u, dgrad);
        e = hipGetLastError();
    }
    std::vector<float> loss(parts);
    if (e == hipSuccess)
        e = scopy(c, grad, dgrad, kMlpParams * sizeof(float), hipMemcpyDeviceToHost);
    if (e == hipSuccess)
        e = scopy(c, loss.data(), b.mlp_loss, parts * sizeof(float), hipMemcpyDeviceToHost);
    (void)hipFree(didx);
    if (dgrad) (void)hipFree(dgrad);
    MLP_HIP(c, e);
    if (loss_out) {
        double l = 0.0;
        for (float v : loss) l += v;
        *loss_out = (float)(l / (double)count);
    }
    return NB_OK;
}

int nb_debug_far_field(nb_ctx* c, uint32_t* nsub, double* sub_centre, uint32_t* sub_range,
                       double* sub_thr, uint32_t* cell_slots, int32_t* body_sub,
                       float* body_off, uint32_t* cell_cap_out, uint32_t* groups_out) {
    if (!c) return NB_ERR_INVALID_ARG;
    if (!c->b.cell_slot)
        return mlp_fail(c, NB_ERR_NOT_SUPPORTED, "no learned far field in this context");
    const uint64_t n = c->n, groups = c->ngroups;
    if (cell_cap_out) *cell_cap_out = c->cell_cap;
    if (groups_out) *groups_out = (uint32_t)groups;
    if (nsub) MLP_HIP(c, scopy(c, nsub, c->b.cell_cnt, groups * 4, hipMemcpyDeviceToHost));
    if (sub_centre || sub_range || sub_thr) {
        std::vector<SubRec> h(groups * 64);
        MLP_HIP(c, scopy(c, h.data(), c->b.subs, h.size() * sizeof(SubRec),
                         hipMemcpyDeviceToHost));
        for (uint64_t k = 0; k < h.size(); k++) {
            if (sub_centre) {
                sub_centre[2 * k] = h[k].ox;
                sub_centre[2 * k + 1] = h[k].oy;
            }
            if (sub_range) {
                sub_range[2 * k] = h[k].start;
                sub_range[2 * k + 1] = h[k].count;
            }
            if (sub_thr) sub_thr[k] = h[k].thr;
        }
    }
    if (cell_slots)
        MLP_HIP(c, scopy(c, cell_slots, c->b.cell_slot, groups * c->cell_cap * 4,
                         hipMemcpyDeviceToHost));
    if (body_sub || body_off) {
        std::vector<float4> h(n);
        MLP_HIP(c, scopy(c, h.data(), c->b.body_w, n * sizeof(float4), hipMemcpyDeviceToHost));
        for (uint64_t i = 0; i < n; i++) {
            if (body_sub) std::memcpy(&body_sub[i], &h[i].z, sizeof(int32_t));
            if (body_off) {
                body_off[2 * i] = h[i].x;
                body_off[2 * i + 1] = h[i].y;
            }
        }
    }
    return NB_OK;
}

int nb_mlp_forward(nb_ctx* c, const float* features, uint64_t count, float* out) {
    if (!c || (count && (!features || !out))) return NB_ERR_INVALID_ARG;
    int rc = ensure_mlp(c);
    if (rc || count == 0) return rc;
    float *dx = nullptr, *dy = nullptr;
    MLP_HIP(c, hipMalloc((void**)&dx, count * kMlpIn * sizeof(float)));
    hipError_t e = hipMalloc((void**)&dy, count * 2 * sizeof(float));
    if (e == hipSuccess)
        e = scopy(c, dx, features, count * kMlpIn * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        const uint64_t batches = (count + 63) / 64;
        const uint64_t blocks = std::min<uint64_t>((batches + 3) / 4, 2048);
        hipLaunchKernelGGL(k_mlp_forward, dim3((uint32_t)blocks), dim3(256), 0,
                           c->stream, dx, count, c->b.mlp_tiles, c->b.mlp_bias, dy);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = scopy(c, out, dy, count * 2 * sizeof(float), hipMemcpyDeviceToHost);
    (void)hipFree(dx);
    if (dy) (void)hipFree(dy);
    MLP_HIP(c, e);
    return NB_OK;
}

}  // extern "C"
