// nb_mlp.h — the far-field correction network on MFMA (device side).
//
// Network: 8 -> 64 -> 64 -> 2, ReLU, BF16 weights and activations, FP32
// accumulation. Parameters in the canonical order W1[64][8], b1[64],
// W2[64][64], b2[64], W3[2][64], b3[2] (4,866 values).
//
// MFMA formulation (v_mfma_f32_16x16x16bf16_1k, D = A B + C, wave64):
//   lane l holds A[m = l%16][k = 4(l/16)+v], B[k = 4(l/16)+v][n = l%16] and
//   D[m = 4(l/16)+v][n = l%16], v = 0..3.
// The weights are A (rows = output neurons, cols = inputs) and the batch is
// B (16 evaluation points per tile, one per column), so a layer's output
// tile t (neurons 16t..16t+15) is, lane for lane, the B operand of k-block t
// of the next layer: no shuffle or LDS round trip between layers. Biases are
// the accumulator's initial value.
//
// Inputs of a (cell, point) pair, d = COM - point, r = |d|, R^2 = r^2 + eps^2,
// s = cell side, M = cell mass, I = second moments about the COM:
//   f0,f1 = d/r   f2 = s/R   f3 = log2(M/M_root)/16
//   f4,f5,f6 = (Ixx, Ixy, Iyy) / (M s^2)   f7 = log2(r/ell)/8
// with ell = eps (or the root cell side when eps = 0). Output y (2 values):
// the correction a_exact - a_monopole of the Plummer kernel in units of
// M s^2 r / R^5, the scale of the softened quadrupole term, whose value in
// these units is (3/2)[-tr(I') u - 2 I' u + 5 (r/R)^2 (u.I'u) u] with
// u = (f0, f1), I' = (f4..f6): an O(1) function of the inputs.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nb {

constexpr int kMlpIn = 8;
constexpr int kMlpHidden = 64;
constexpr int kMlpOut = 2;
constexpr int kMlpParams = kMlpHidden * kMlpIn + kMlpHidden + kMlpHidden * kMlpHidden +
                           kMlpHidden + kMlpOut * kMlpHidden + kMlpOut;  // 4866
constexpr int kMlpTiles = 24;   // A tiles: W1 0..3, W2 4..19 (4 n x 4 k), W3 20..23

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// Round to nearest even, as bfloat16 (the same rounding the host packer and
// the training kernel use).
__host__ __device__ inline uint16_t f32_to_bf16(float f) {
    union {
        float f;
        uint32_t u;
    } c;
    c.f = f;
    uint32_t u = c.u;
    if ((u & 0x7f800000u) != 0x7f800000u) u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__host__ __device__ inline float bf16_to_f32(uint16_t h) {
    union {
        uint32_t u;
        float f;
    } c;
    c.u = (uint32_t)h << 16;
    return c.f;
}

__host__ __device__ inline float bf16_round(float f) { return bf16_to_f32(f32_to_bf16(f)); }

__device__ inline bf16x4 pack_bf16(f32x4 v) {
    bf16x4 r;
    r.x = (short)f32_to_bf16(v.x);
    r.y = (short)f32_to_bf16(v.y);
    r.z = (short)f32_to_bf16(v.z);
    r.w = (short)f32_to_bf16(v.w);
    return r;
}

__device__ inline f32x4 relu4(f32x4 v) {
    f32x4 r;
    r.x = fmaxf(v.x, 0.f);
    r.y = fmaxf(v.y, 0.f);
    r.z = fmaxf(v.z, 0.f);
    r.w = fmaxf(v.w, 0.f);
    return r;
}

// The 8 inputs of one (cell, point) pair and the output scale.
__device__ inline float mlp_features(float dx, float dy, float s, float M, float ixx, float ixy,
                                     float iyy, float eps2, float inv_mroot, float inv_ell,
                                     float f[8]) {
    const float r2 = fmaf(dx, dx, dy * dy);
    const float inv_r = __builtin_amdgcn_rsqf(r2);
    const float r = r2 * inv_r;
    const float inv_R = __builtin_amdgcn_rsqf(r2 + eps2);
    const float ms2 = M * s * s;
    const float k = 1.0f / ms2;
    f[0] = dx * inv_r;
    f[1] = dy * inv_r;
    f[2] = s * inv_R;
    f[3] = __log2f(M * inv_mroot) * 0.0625f;
    f[4] = ixx * k;
    f[5] = ixy * k;
    f[6] = iyy * k;
    f[7] = __log2f(r * inv_ell) * 0.125f;
    const float iR2 = inv_R * inv_R;
    return ms2 * r * (iR2 * iR2 * inv_R);
}

// 1 / ell for the distance input: eps, or the root cell side when eps = 0.
__device__ inline float mlp_inv_ell(float eps2f, double root_wsq) {
    return eps2f > 0.f ? __builtin_amdgcn_rsqf(eps2f) : (float)(1.0 / sqrt(root_wsq));
}

// Bias vector in LDS: b1[64], b2[64], b3[2], zero padding to kMlpBiasLen.
constexpr int kMlpBiasLen = 144;

// The 24 A tiles of this lane, loaded once per batch of tiles.
struct MlpWeights {
    bf16x4 a[kMlpTiles];
};

__device__ inline void mlp_load_weights(MlpWeights& w, const uint2* __restrict__ wt, int lane) {
#pragma unroll
    for (int t = 0; t < kMlpTiles; t++) {
        const uint2 u = wt[t * 64 + lane];
        w.a[t].x = (short)(u.x & 0xffff);
        w.a[t].y = (short)(u.x >> 16);
        w.a[t].z = (short)(u.y & 0xffff);
        w.a[t].w = (short)(u.y >> 16);
    }
}

__device__ inline f32x4 mlp_bias4(const float* sb, int off, int lane) {
    return *reinterpret_cast<const f32x4*>(sb + off + 4 * (lane >> 4));
}

// Forward pass of one 16-column tile. `x` holds this lane's four inputs
// (lanes 0-15: f0..f3 of column l, lanes 16-31: f4..f7 of column l-16,
// lanes 32-63: zero). Returns this lane's D rows of the last layer: lanes
// 0-15 hold (y0, y1, 0, 0) of column l.
__device__ inline f32x4 mlp_tile_forward(f32x4 x, const MlpWeights& w, const float* sb,
                                         int lane) {
    const bf16x4 b0 = pack_bf16(x);
    bf16x4 h1[4];
#pragma unroll
    for (int t = 0; t < 4; t++)
        h1[t] = pack_bf16(relu4(__builtin_amdgcn_mfma_f32_16x16x16bf16_1k(
            w.a[t], b0, mlp_bias4(sb, 16 * t, lane), 0, 0, 0)));
    bf16x4 h2[4];
#pragma unroll
    for (int nt = 0; nt < 4; nt++) {
        f32x4 c = mlp_bias4(sb, 64 + 16 * nt, lane);
#pragma unroll
        for (int kt = 0; kt < 4; kt++)
            c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(w.a[4 + 4 * nt + kt], h1[kt], c, 0, 0,
                                                          0);
        h2[nt] = pack_bf16(relu4(c));
    }
    f32x4 y = mlp_bias4(sb, 128, lane);
#pragma unroll
    for (int kt = 0; kt < 4; kt++)
        y = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(w.a[20 + kt], h2[kt], y, 0, 0, 0);
    return y;
}

// Corrections of `count` accepted cells (LDS: ca = {dx, dy, M, s} with
// d = COM - group centre, cb = {Ixx, Ixy, Iyy, -}) at three points: the group
// centre, centre + (hx, 0) and centre + (0, hy). Lanes 0-15 add the sums over
// cells to u[2q], u[2q+1] for point q; a member at offset (bx, by) from the
// centre then takes u0 + (bx/hx)(u1 - u0) + (by/hy)(u2 - u0).
__device__ inline void mlp_eval_cells(const float4* ca, const float4* cb, int count, float hx,
                                      float hy, float eps2f, float inv_mroot, float inv_ell,
                                      const MlpWeights& w, const float* sb, int lane,
                                      float u[6]) {
    const int col = lane & 15, part = lane >> 4;
    const int tiles = (count + 15) >> 4;
#pragma nounroll
    for (int q = 0; q < 3; q++) {
        const float ox = q == 1 ? hx : 0.f, oy = q == 2 ? hy : 0.f;
        for (int t = 0; t < tiles; t++) {
            const int cell = 16 * t + col;
            const bool ok = cell < count && part < 2;
            f32x4 x = {0.f, 0.f, 0.f, 0.f};
            float scale = 0.f;
            if (ok) {
                const float4 a = ca[cell], b = cb[cell];
                float f[8];
                scale = mlp_features(a.x - ox, a.y - oy, a.w, a.z, b.x, b.y, b.z, eps2f,
                                     inv_mroot, inv_ell, f);
                x.x = part ? f[4] : f[0];
                x.y = part ? f[5] : f[1];
                x.z = part ? f[6] : f[2];
                x.w = part ? f[7] : f[3];
            }
            const f32x4 y = mlp_tile_forward(x, w, sb, lane);
            if (ok && part == 0) {
                u[2 * q] = fmaf(y.x, scale, u[2 * q]);
                u[2 * q + 1] = fmaf(y.y, scale, u[2 * q + 1]);
            }
        }
    }
}

}  // namespace nb
