// nb_mlp.hip — the learned far-field correction: training pairs generated by
// exact direct sums on the GPU, hand-written training kernels (forward +
// back-propagation in FP32 with the network's BF16 rounding, Adam), MFMA
// inference (nb_mlp.h), and the C ABI around them.
//
// Pair generation: sample bodies walk the tree with the per-body opening test
// at theta (cells holding the body itself are never accepted); every accepted
// internal cell with at most kPairMaxBodies bodies becomes a pair. Its target
// is the exact FP64 Plummer acceleration of the cell's bodies at the sample
// body minus the cell's monopole, in units of M s^2 / r^4.
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
// Pair collection: wave-cooperative masked walk (as the EXACT traversal).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_mlp_collect(const NodeRec* __restrict__ nodes,
                                                    const double2* __restrict__ pos, uint32_t n,
                                                    uint32_t samples, uint64_t seed,
                                                    double theta_sq, uint64_t cap,
                                                    uint32_t* __restrict__ pair_body,
                                                    uint32_t* __restrict__ pair_slot,
                                                    unsigned long long* __restrict__ count) {
    __shared__ int32_t s_node[kStackDepth];
    __shared__ uint64_t s_mask[kStackDepth];
    const int l = lane_id();
    const uint32_t k = blockIdx.x * 64 + l;
    const bool valid = k < samples && n > 1;
    const int32_t me = valid ? (int32_t)(hash_u32(seed * 0x9e3779b97f4a7c15ull + k) % n) : -1;
    const double2 p = valid ? pos[me] : make_double2(0.0, 0.0);
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
        const bool take = accept && (uint32_t)(be - bs + 1) <= kPairMaxBodies;
        const uint64_t tm = ballot64(take);
        if (tm) {
            unsigned long long base = 0;
            if (l == __ffsll((unsigned long long)tm) - 1)
                base = atomicAdd(count, (unsigned long long)__popcll(tm));
            base = __shfl(base, __ffsll((unsigned long long)tm) - 1, 64);
            const uint64_t at = base + __popcll(tm & lanemask_lt());
            if (take && at < cap) {
                pair_body[at] = (uint32_t)me;
                pair_slot[at] = (uint32_t)node;
            }
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
}

// Features and exact-minus-monopole targets, one wave per pair.
__global__ __launch_bounds__(256) void k_mlp_targets(
    const NodeRec* __restrict__ nodes, const double2* __restrict__ pos,
    const double* __restrict__ mass, const double* __restrict__ box, double eps2,
    const uint32_t* __restrict__ pair_body, const uint32_t* __restrict__ pair_slot,
    const unsigned long long* __restrict__ count, uint64_t cap, float* __restrict__ X,
    float* __restrict__ T) {
    const uint64_t npairs = *count < cap ? *count : cap;
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
            float f[8];
            const float scale = mlp_features((float)dx, (float)dy, (float)sqrt(r->wsq), r->mf,
                                             r->ixx, r->ixy, r->iyy, eps2f, inv_mroot, inv_ell, f);
#pragma unroll
            for (int i = 0; i < 8; i++) X[q * 8 + i] = f[i];
            T[q * 2 + 0] = (float)((ax - dx * w) / (double)scale);
            T[q * 2 + 1] = (float)((ay - dy * w) / (double)scale);
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
    static constexpr int W1 = 0, b1 = 512, W2 = 576, b2 = 4672, W3 = 4736, b3 = 4864;
};

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
        for (int k = 0; k < kMlpIn; k++) sx[i][k] = ok ? bf16_round(X[src * 8 + k]) : 0.f;
        st[i][0] = ok ? T[src * 2 + 0] : 0.f;
        st[i][1] = ok ? T[src * 2 + 1] : 0.f;
    }
    __syncthreads();
    const int s = t >> 2, j0 = (t & 3) * 16;
    // layer 1
    for (int j = j0; j < j0 + 16; j++) {
        float a = sw[O::b1 + j];
#pragma unroll
        for (int k = 0; k < kMlpIn; k++) a = fmaf(sw[O::W1 + j * 8 + k], sx[s][k], a);
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
    // dW1 (2 outputs per thread), db1
    for (int q = 0; q < 2; q++) {
        const int idx = t * 2 + q, j = idx >> 3, k = idx & 7;
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

// Batch inference on MFMA: each wave loads the weights once and runs 16-row
// tiles of X.
__global__ __launch_bounds__(256) void k_mlp_forward(const float* __restrict__ X, uint64_t count,
                                                     const uint2* __restrict__ wt,
                                                     const float* __restrict__ bias,
                                                     float* __restrict__ Y) {
    __shared__ __attribute__((aligned(16))) float sb[kMlpBiasLen];
    for (int i = threadIdx.x; i < kMlpBiasLen; i += 256) sb[i] = bias[i];
    __syncthreads();
    const int l = lane_id();
    MlpWeights w;
    mlp_load_weights(w, wt, l);
    const uint64_t tiles = (count + 15) / 16;
    const uint64_t wave = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const uint64_t waves = (uint64_t)gridDim.x * 4;
    const int part = l >> 4;  // 0: f0..f3, 1: f4..f7, 2-3: zero padding
    for (uint64_t tile = wave; tile < tiles; tile += waves) {
        const uint64_t col = tile * 16 + (l & 15);
        f32x4 x = {0.f, 0.f, 0.f, 0.f};
        if (col < count && part < 2) {
            x.x = X[col * 8 + part * 4 + 0];
            x.y = X[col * 8 + part * 4 + 1];
            x.z = X[col * 8 + part * 4 + 2];
            x.w = X[col * 8 + part * 4 + 3];
        }
        const f32x4 y = mlp_tile_forward(x, w, sb, l);
        if (l < 16 && col < count) {
            Y[col * 2 + 0] = y.x;
            Y[col * 2 + 1] = y.y;
        }
    }
}

// ---------------------------------------------------------------------------
// Far-field network in the step: one wavefront per group of 64 bodies. The
// group's accepted cells (written by the walk) run through the network in
// 16-cell tiles at the three points of their subgroup; the corrections are
// summed per subgroup in LDS in a fixed order, then every body adds the
// first-order interpolation at its own offset to the walk's acceleration and
// kicks (or stores G a).
// ---------------------------------------------------------------------------
constexpr int kApplyWaves = 4;

template <bool KICK>
__global__ __launch_bounds__(64 * kApplyWaves) void k_mlp_apply(
    const CellRec* __restrict__ cells, const uint32_t* __restrict__ cell_cnt, uint32_t cell_cap,
    const float2* __restrict__ sub_h, const float4* __restrict__ body_w,
    double2* __restrict__ acc, double2* __restrict__ vel, uint32_t n,
    const StepConsts* __restrict__ k, uint32_t* __restrict__ meta,
    const NodeRec* __restrict__ nodes, const double* __restrict__ box,
    const uint2* __restrict__ tiles, const float* __restrict__ bias, int count_step) {
    if (meta[META_FLAGS] & kHaltFlags) return;
    __shared__ __attribute__((aligned(16))) float sb[kMlpBiasLen];
    __shared__ float su[kApplyWaves][64 * 6];
    __shared__ float2 sh[kApplyWaves][64];
    for (int i = threadIdx.x; i < kMlpBiasLen; i += 64 * kApplyWaves) sb[i] = bias[i];
    if (count_step && blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&meta[META_STEPS_DONE], 1u);
    __syncthreads();
    const int l = lane_id(), wv = threadIdx.x >> 6;
    const uint32_t groups = (n + 63) / 64;
    const uint32_t g = blockIdx.x * kApplyWaves + wv;
    if (g >= groups) return;
    float* u = su[wv];
#pragma unroll
    for (int q = 0; q < 6; q++) u[q * 64 + l] = 0.f;
    sh[wv][l] = sub_h[(uint64_t)g * 64 + l];
    __builtin_amdgcn_wave_barrier();
    const float eps2f = (float)k->eps2;
    const float inv_mroot = (float)(1.0 / nodes[0].m);
    const float inv_ell = mlp_inv_ell(eps2f, box[4]);
    MlpWeights w;
    mlp_load_weights(w, tiles, l);
    const uint32_t cnt = min(cell_cnt[g], cell_cap);
    const CellRec* cg = cells + (uint64_t)g * cell_cap;
    const int col = l & 15, part = l >> 4;
    for (uint32_t t0 = 0; t0 < cnt; t0 += 16) {
        const uint32_t cell = t0 + col;
        const bool ok = cell < cnt;
        CellRec cr;
        cr.a = make_float4(1.f, 0.f, 1.f, 1.f);
        cr.b = make_float4(0.f, 0.f, 0.f, 0.f);
        if (ok) cr = cg[cell];
        const int sub = ok ? (__float_as_int(cr.b.w) & 63) : 0;
        const float2 h = sh[wv][sub];
        float cx[3], cy[3];
#pragma unroll
        for (int q = 0; q < 3; q++) {
            const float ox = q == 1 ? h.x : 0.f, oy = q == 2 ? h.y : 0.f;
            float f[8];
            const float scale = mlp_features(cr.a.x - ox, cr.a.y - oy, cr.a.w, cr.a.z, cr.b.x,
                                             cr.b.y, cr.b.z, eps2f, inv_mroot, inv_ell, f);
            f32x4 x = {0.f, 0.f, 0.f, 0.f};
            if (ok && part < 2) {
                x.x = part ? f[4] : f[0];
                x.y = part ? f[5] : f[1];
                x.z = part ? f[6] : f[2];
                x.w = part ? f[7] : f[3];
            }
            const f32x4 y = mlp_tile_forward(x, w, sb, l);
            const bool use = ok && part == 0;
            cx[q] = use ? y.x * scale : 0.f;
            cy[q] = use ? y.y * scale : 0.f;
        }
        // Add the 16 columns to their subgroups' sums, in column order.
        const int s0 = __shfl(sub, 0, 64);
        if (ballot64(l < 16 && ok && sub != s0) == 0) {
#pragma unroll
            for (int q = 0; q < 3; q++) {
#pragma unroll
                for (int d = 1; d < 16; d <<= 1) {
                    cx[q] += __shfl_xor(cx[q], d, 64);
                    cy[q] += __shfl_xor(cy[q], d, 64);
                }
            }
            if (l == 0) {
#pragma unroll
                for (int q = 0; q < 3; q++) {
                    u[(2 * q) * 64 + s0] += cx[q];
                    u[(2 * q + 1) * 64 + s0] += cy[q];
                }
            }
        } else {
            for (int j = 0; j < 16; j++) {
                if (l == j && ok) {
#pragma unroll
                    for (int q = 0; q < 3; q++) {
                        u[(2 * q) * 64 + sub] += cx[q];
                        u[(2 * q + 1) * 64 + sub] += cy[q];
                    }
                }
                __builtin_amdgcn_wave_barrier();
            }
        }
    }
    __builtin_amdgcn_wave_barrier();
    const uint32_t i = g * 64 + l;
    if (i >= n) return;
    const float4 bw = body_w[i];
    const int sub = __float_as_int(bw.z) & 63;
    const float u0 = u[0 * 64 + sub], u1 = u[1 * 64 + sub];
    const float ax = u0 + bw.x * (u[2 * 64 + sub] - u0) + bw.y * (u[4 * 64 + sub] - u0);
    const float ay = u1 + bw.x * (u[3 * 64 + sub] - u1) + bw.y * (u[5 * 64 + sub] - u1);
    const double2 a0 = acc[i];
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
    const uint32_t groups = (n + 63) / 64;
    const uint32_t blocks = (groups + kApplyWaves - 1) / kApplyWaves;
    const int count = count_step ? 1 : 0;
#define NB_APPLY(K)                                                                            \
    hipLaunchKernelGGL((k_mlp_apply<K>), dim3(blocks), dim3(64 * kApplyWaves), 0, c->stream,   \
                       c->b.cells, c->b.cell_cnt, c->cell_cap, c->b.sub_h, c->b.body_w,        \
                       c->b.acc, c->b.vel[src], n, c->b.consts, c->b.meta, c->b.nodes,         \
                       c->b.box, c->b.mlp_tiles, c->b.mlp_bias, count)
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

// Packs the FP32 canonical parameters into the MFMA A tiles (BF16, per-lane
// layout: lane l of tile T holds A[l%16][4(l/16) + v], v = 0..3) and the
// compact bias vector.
void pack_tiles(const float* p, std::vector<uint2>& tiles, std::vector<float>& bias) {
    using O = MlpOffsets;
    tiles.assign(kMlpTiles * 64, uint2{0, 0});
    bias.assign(kMlpBiasLen, 0.f);
    auto W = [&](int T, int row, int col) -> float {
        if (T < 4) {  // W1[64][8], k >= 8 zero
            const int n = 16 * T + row;
            return col < 8 ? p[O::W1 + n * 8 + col] : 0.f;
        }
        if (T < 20) {
            const int nt = (T - 4) / 4, kt = (T - 4) % 4;
            return p[O::W2 + (16 * nt + row) * 64 + 16 * kt + col];
        }
        const int kt = T - 20;
        return row < 2 ? p[O::W3 + row * 64 + 16 * kt + col] : 0.f;
    };
    for (int T = 0; T < kMlpTiles; T++)
        for (int l = 0; l < 64; l++) {
            uint16_t h[4];
            for (int v = 0; v < 4; v++) h[v] = f32_to_bf16(W(T, l % 16, 4 * (l / 16) + v));
            tiles[T * 64 + l] = uint2{(uint32_t)h[0] | ((uint32_t)h[1] << 16),
                                      (uint32_t)h[2] | ((uint32_t)h[3] << 16)};
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
    MLP_HIP(c, hipMalloc((void**)&b.mlp_tiles, kMlpTiles * 64 * sizeof(uint2)));
    MLP_HIP(c, hipMalloc((void**)&b.mlp_bias, kMlpBiasLen * sizeof(float)));
    MLP_HIP(c, hipMemsetAsync(b.mlp_w, 0, kMlpParams * sizeof(float), c->stream));
    MLP_HIP(c, hipMemsetAsync(b.mlp_tiles, 0, kMlpTiles * 64 * sizeof(uint2), c->stream));
    MLP_HIP(c, hipMemsetAsync(b.mlp_bias, 0, kMlpBiasLen * sizeof(float), c->stream));
    MLP_HIP(c, hipStreamSynchronize(c->stream));
    c->device_bytes += kMlpParams * 3 * sizeof(float) + kMlpTiles * 64 * sizeof(uint2) +
                       kMlpBiasLen * sizeof(float);
    return NB_OK;
}

int upload_weights(nb_ctx* c, const float* p) {
    std::vector<uint2> tiles;
    std::vector<float> bias;
    pack_tiles(p, tiles, bias);
    MLP_HIP(c, scopy(c, c->b.mlp_w, p, kMlpParams * sizeof(float), hipMemcpyHostToDevice));
    MLP_HIP(c, scopy(c, c->b.mlp_tiles, tiles.data(), tiles.size() * sizeof(uint2),
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
    MLP_HIP(c, hipMalloc((void**)&b.pair_x, cap * 8 * sizeof(float)));
    MLP_HIP(c, hipMalloc((void**)&b.pair_t, cap * 2 * sizeof(float)));
    if (!b.pair_count) MLP_HIP(c, hipMalloc((void**)&b.pair_count, sizeof(unsigned long long)));
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
                    (void*)b.pair_slot, (void*)b.pair_x, (void*)b.pair_t, (void*)b.pair_count,
                    (void*)b.cells, (void*)b.cell_cnt, (void*)b.sub_h, (void*)b.body_w})
        if (p) (void)hipFree(p);
    b.cells = nullptr;
    b.cell_cnt = nullptr;
    b.sub_h = nullptr;
    b.body_w = nullptr;
    c->cell_cap = 0;
    b.mlp_w = b.mlp_m = b.mlp_v = b.mlp_gpart = b.mlp_loss = b.pair_x = b.pair_t = nullptr;
    b.mlp_tiles = nullptr;
    b.mlp_bias = nullptr;
    b.pair_body = b.pair_slot = nullptr;
    b.pair_count = nullptr;
}

int prepare_mlp(nb_ctx* c) { return ensure_mlp(c); }

int alloc_cells(nb_ctx* c, uint32_t cap) {
    Buffers& b = c->b;
    const uint64_t groups = (c->n + 63) / 64;
    if (b.cells) (void)hipFree(b.cells);
    b.cells = nullptr;
    if (c->cell_cap) c->device_bytes -= groups * c->cell_cap * sizeof(CellRec);
    c->cell_cap = 0;
    MLP_HIP(c, hipMalloc((void**)&b.cells, groups * cap * sizeof(CellRec)));
    if (!b.cell_cnt) {
        MLP_HIP(c, hipMalloc((void**)&b.cell_cnt, groups * sizeof(uint32_t)));
        MLP_HIP(c, hipMalloc((void**)&b.sub_h, groups * 64 * sizeof(float2)));
        MLP_HIP(c, hipMalloc((void**)&b.body_w, c->n * sizeof(float4)));
        c->device_bytes += groups * (sizeof(uint32_t) + 64 * sizeof(float2)) + c->n * sizeof(float4);
    }
    c->device_bytes += groups * cap * sizeof(CellRec);
    c->cell_cap = cap;
    return NB_OK;
}

}  // namespace nb

extern "C" {

int nb_mlp_set_weights(nb_ctx* c, const float* params, uint64_t count) {
    if (!c || !params) return NB_ERR_INVALID_ARG;
    if (count != (uint64_t)kMlpParams)
        return mlp_fail(c, NB_ERR_INVALID_ARG, "expected 4866 parameters");
    int rc = ensure_mlp(c);
    if (rc) return rc;
    return upload_weights(c, params);
}

int nb_mlp_get_weights(nb_ctx* c, float* params, uint64_t count) {
    if (!c || !params) return NB_ERR_INVALID_ARG;
    if (count != (uint64_t)kMlpParams)
        return mlp_fail(c, NB_ERR_INVALID_ARG, "expected 4866 parameters");
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
    for (int i = 0; i < 64 * 8; i++) p[O::W1 + i] = nd(rng) * std::sqrt(2.f / 8.f);
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
    int rc = build_tree(c);
    if (rc) return rc;
    rc = ensure_pairs(c, max_pairs ? max_pairs : 1);
    if (rc) return rc;
    Buffers& b = c->b;
    MLP_HIP(c, hipMemsetAsync(b.pair_count, 0, sizeof(unsigned long long), c->stream));
    const uint32_t waves = (sample_bodies + 63) / 64;
    hipLaunchKernelGGL(k_mlp_collect, dim3(waves ? waves : 1), dim3(64), 0, c->stream, b.nodes,
                       b.pos[c->cur], (uint32_t)c->n, sample_bodies, seed, theta * theta,
                       (uint64_t)c->pair_cap, b.pair_body, b.pair_slot, b.pair_count);
    hipLaunchKernelGGL(k_mlp_targets, dim3(2048), dim3(256), 0, c->stream, b.nodes,
                       b.pos[c->cur], b.mass[c->cur], b.box,
                       c->p.softening * c->p.softening, b.pair_body, b.pair_slot, b.pair_count,
                       (uint64_t)c->pair_cap, b.pair_x, b.pair_t);
    MLP_HIP(c, hipGetLastError());
    unsigned long long cnt = 0;
    MLP_HIP(c, hipMemcpyAsync(&cnt, b.pair_count, sizeof(cnt), hipMemcpyDeviceToHost, c->stream));
    MLP_HIP(c, hipStreamSynchronize(c->stream));
    c->pair_count = cnt < c->pair_cap ? cnt : c->pair_cap;
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
        MLP_HIP(c, scopy(c, features, c->b.pair_x, k * 8 * sizeof(float), hipMemcpyDeviceToHost));
    if (targets)
        MLP_HIP(c, scopy(c, targets, c->b.pair_t, k * 2 * sizeof(float), hipMemcpyDeviceToHost));
    return NB_OK;
}

int nb_mlp_set_pairs(nb_ctx* c, const float* features, const float* targets, uint64_t count) {
    if (!c || (count && (!features || !targets))) return NB_ERR_INVALID_ARG;
    int rc = ensure_pairs(c, count ? count : 1);
    if (rc) return rc;
    if (count) {
        MLP_HIP(c, scopy(c, c->b.pair_x, features, count * 8 * sizeof(float),
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

int nb_mlp_forward(nb_ctx* c, const float* features, uint64_t count, float* out) {
    if (!c || (count && (!features || !out))) return NB_ERR_INVALID_ARG;
    int rc = ensure_mlp(c);
    if (rc || count == 0) return rc;
    float *dx = nullptr, *dy = nullptr;
    MLP_HIP(c, hipMalloc((void**)&dx, count * 8 * sizeof(float)));
    hipError_t e = hipMalloc((void**)&dy, count * 2 * sizeof(float));
    if (e == hipSuccess)
        e = scopy(c, dx, features, count * 8 * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        const uint64_t tiles = (count + 15) / 16;
        const uint64_t blocks = std::min<uint64_t>((tiles + 3) / 4, 2048);
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
