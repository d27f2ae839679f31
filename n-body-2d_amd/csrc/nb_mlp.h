// nb_mlp.h — the far-field correction network on MFMA (device side).
//
// Network: 12 -> 64 -> 64 -> 2, ReLU, BF16 weights and activations, FP32
// accumulation. Parameters in the canonical order W1[64][12], b1[64],
// W2[64][64], b2[64], W3[2][64], b3[2] (5,122 values).
//
// MFMA formulation (D = A B + C, wave64). The weights are A (rows = output
// neurons), the batch is B (16 evaluation points per tile, one per column);
// biases are the accumulator's initial value.
//   layer 1, v_mfma_f32_16x16x16_bf16: lane l holds A[l%16][4(l/16)+v],
//     B[4(l/16)+v][l%16] (v = 0..3): inputs f0..f3 in lanes 0-15, f4..f7 in
//     lanes 16-31, f8..f11 in lanes 32-47, zero padding (k = 12..15) in
//     lanes 48-63.
//   layers 2, 3, v_mfma_f32_16x16x32_bf16: lane l holds A[l%16][8(l/16)+v],
//     B[8(l/16)+v][l%16] (v = 0..7).
//   every layer's D: lane l holds D[4(l/16)+v][l%16] (v = 0..3).
// Output tile t of a layer (neurons 16t..16t+15) gives lane l the neurons
// 16t + 4(l/16) + v. The K = 32 block kb of the next layer takes tiles 2kb
// and 2kb+1 as they are in the registers: its k index 8g + v (g = l/16)
// stands for neuron 32kb + 4g + v (v < 4) or 32kb + 16 + 4g + v - 4
// (v >= 4), and the packed weights use the same order (mlp_k_neuron), so
// no shuffle or LDS round trip happens between layers.
//
// Inputs of a (cell, point) pair, d = COM - point, r = |d|, u = d/r,
// w = (-u_y, u_x), R^2 = r^2 + eps^2, s = cell side, M = cell mass, I =
// second and T = third moments about the COM, taken in the frame (u, w):
//   f0,f1 = u   f2 = s/R   f3 = log2(M/M_root)/16
//   f4,f5,f6 = (I_uu, I_uw, I_ww) / (M s^2)   f7 = log2(r/ell)/8
//   f8..f11 = (T_uuu, T_uuw, T_uww, T_www) / (M s^3)
// with ell = eps (or the root cell side when eps = 0). Units of the
// correction a_exact - a_monopole of the Plummer kernel along (u, w):
// M s^2 r / R^5, the scale of the softened quadrupole term, which in these
// units is q = (3/2)(5 rho f4 - 3 f4 - f6, -2 f5), rho = r^2/R^2. The
// network output (y0, y1) is trained on the rest, a_exact - a_monopole - q,
// over f2 = s/R. The rest is led by the octupole, of order s/R in these
// units and linear in the third moments f8..f11 (with coefficients that
// depend on rho), so every target is of order one and a far cell's share of
// the network's error is small. In a step the walk adds the softened
// quadrupole term of each cell near the body's subgroup at the body's own
// position; the farther cells' terms (moved to the body by their Jacobian,
// quad_jacobian) and f2 (y0, y1) are added at the centre of the subgroup.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nb {

constexpr int kMlpIn = 12;
constexpr int kMlpHidden = 64;
constexpr int kMlpOut = 2;
constexpr int kMlpParams = kMlpHidden * kMlpIn + kMlpHidden + kMlpHidden * kMlpHidden +
                           kMlpHidden + kMlpOut * kMlpHidden + kMlpOut;  // 5122
// Packed A tiles: layer 1 as 4 x 64 uint2 (4 BF16 per lane), then layer 2
// (4 output tiles x 2 K blocks) and layer 3 (2 K blocks) as 10 x 64 uint4
// (8 BF16 per lane).
constexpr int kMlpTiles1 = 4;
constexpr int kMlpTiles23 = 10;
constexpr int kMlpTileBytes = kMlpTiles1 * 64 * 8 + kMlpTiles23 * 64 * 16;
// Bias vector in LDS: b1[64], b2[64], b3[2], zero padding.
constexpr int kMlpBiasLen = 144;

typedef short s16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// Neuron that k index j (0..31) of K block kb stands for (see above).
__host__ __device__ constexpr int mlp_k_neuron(int kb, int j) {
    return 32 * kb + ((j & 7) < 4 ? 4 * (j >> 3) + (j & 7) : 16 + 4 * (j >> 3) + (j & 7) - 4);
}

// Round to nearest even, as bfloat16 (host packing and the training kernel;
// the inference path uses the hardware conversion, which rounds the same way).
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

// Two floats to two BF16 in one register (v_cvt_pk_bf16_f32).
__device__ inline uint32_t pk2(float a, float b) {
    bf16x2 v;
    v.x = (__bf16)a;
    v.y = (__bf16)b;
    return __builtin_bit_cast(uint32_t, v);
}

__device__ inline s16x4 pack4(f32x4 v) {
    return __builtin_bit_cast(s16x4, make_uint2(pk2(v.x, v.y), pk2(v.z, v.w)));
}

__device__ inline f32x4 relu4(f32x4 v) {
    f32x4 r;
    r.x = fmaxf(v.x, 0.f);
    r.y = fmaxf(v.y, 0.f);
    r.z = fmaxf(v.z, 0.f);
    r.w = fmaxf(v.w, 0.f);
    return r;
}

// ReLU of two packed BF16: a BF16 with its sign bit set is a negative int16,
// so max(int16, 0) per half (v_pk_max_i16) zeroes exactly the negative values
// (and -0) and leaves the rest: one instruction for two activations, instead
// of an FP32 max (with its canonicalisation) per activation before packing.
__device__ inline uint32_t relu_bf16x2(uint32_t v) {
    typedef short s16x2 __attribute__((ext_vector_type(2)));
    const s16x2 x = __builtin_elementwise_max(__builtin_bit_cast(s16x2, v), (s16x2){0, 0});
    return __builtin_bit_cast(uint32_t, x);
}

// Two output tiles of a layer (ReLU applied) as the B operand of one K block.
__device__ inline bf16x8 pack8_relu(f32x4 a, f32x4 b) {
    return __builtin_bit_cast(bf16x8, make_uint4(relu_bf16x2(pk2(a.x, a.y)),
                                                 relu_bf16x2(pk2(a.z, a.w)),
                                                 relu_bf16x2(pk2(b.x, b.y)),
                                                 relu_bf16x2(pk2(b.z, b.w))));
}

// The kMlpIn inputs of one (cell, point) pair; returns the output scale and
// the unit offset (ux, uy) that rotates the output back. t3 = third moments
// about the COM {T_xxx, T_xxy, T_xyy, T_yyy}.
__device__ inline float mlp_features(float dx, float dy, float s, float M, float ixx, float ixy,
                                     float iyy, float4 t3, float eps2, float inv_mroot,
                                     float inv_ell, float f[kMlpIn], float& ux, float& uy,
                                     float q[2]) {
    const float r2 = fmaf(dx, dx, dy * dy);
    const float inv_r = __builtin_amdgcn_rsqf(r2);
    const float r = r2 * inv_r;
    const float inv_R = __builtin_amdgcn_rsqf(r2 + eps2);
    const float ms2 = M * s * s;
    const float k = __builtin_amdgcn_rcpf(ms2);
    ux = dx * inv_r;
    uy = dy * inv_r;
    // moments in the frame (u, w), w = (-uy, ux)
    const float iu_x = fmaf(ixx, ux, ixy * uy), iu_y = fmaf(ixy, ux, iyy * uy);
    const float uiu = fmaf(ux, iu_x, uy * iu_y);
    const float uiw = fmaf(-uy, iu_x, ux * iu_y);
    const float wiw = (ixx + iyy) - uiu;
    f[0] = ux;
    f[1] = uy;
    f[2] = s * inv_R;
    f[3] = __log2f(M * inv_mroot) * 0.0625f;
    f[4] = uiu * k;
    f[5] = uiw * k;
    f[6] = wiw * k;
    f[7] = __log2f(r * inv_ell) * 0.125f;
    // third moments in the frame (u, w): T_abc u or w in each slot
    const float k3 = k * __builtin_amdgcn_rcpf(s);
    const float wx = -uy, wy = ux;
    // contract one index with u: the 2x2 symmetric matrix A_bc = u_a T_abc
    const float axx = fmaf(ux, t3.x, uy * t3.y), axy = fmaf(ux, t3.y, uy * t3.z),
                ayy = fmaf(ux, t3.z, uy * t3.w);
    // ... or with w: B_bc = w_a T_abc
    const float bxx = fmaf(wx, t3.x, wy * t3.y), bxy = fmaf(wx, t3.y, wy * t3.z),
                byy = fmaf(wx, t3.z, wy * t3.w);
    const float uau = fmaf(ux, fmaf(ux, axx, uy * axy), uy * fmaf(ux, axy, uy * ayy));
    const float uaw = fmaf(wx, fmaf(ux, axx, uy * axy), wy * fmaf(ux, axy, uy * ayy));
    const float wbw = fmaf(wx, fmaf(wx, bxx, wy * bxy), wy * fmaf(wx, bxy, wy * byy));
    const float wbu = fmaf(wx, fmaf(ux, bxx, uy * bxy), wy * fmaf(ux, bxy, uy * byy));
    f[8] = uau * k3;   // T_uuu
    f[9] = uaw * k3;   // T_uuw
    f[10] = wbu * k3;  // T_uww
    f[11] = wbw * k3;  // T_www
    const float iR2 = inv_R * inv_R;
    // the softened quadrupole term in these units
    const float rho = r2 * iR2;
    q[0] = 1.5f * fmaf(5.f * rho, f[4], -fmaf(3.f, f[4], f[6]));
    q[1] = -3.f * f[5];
    return ms2 * r * (iR2 * iR2 * inv_R);
}

// Jacobian dQ/dd of the softened quadrupole term Q(d) = (3/(2R^5))
// [-tr(I) d - 2 I d + 5 (d.Id) d / R^2] (d = COM - point, R^2 = |d|^2 + eps^2),
// FP32: J = {dQx/ddx, dQx/ddy, dQy/ddx, dQy/ddy}. Moving the point by b
// changes d by -b, so Q(point + b) ~ Q - J b.
__device__ inline void quad_jacobian(float dx, float dy, float ixx, float ixy, float iyy,
                                     float eps2, float J[4]) {
    const float R2 = fmaf(dx, dx, fmaf(dy, dy, eps2));
    const float iR = __builtin_amdgcn_rsqf(R2);
    const float iR2 = iR * iR;
    const float iR5 = iR2 * iR2 * iR;
    const float t = ixx + iyy;
    const float idx = fmaf(ixx, dx, ixy * dy), idy = fmaf(ixy, dx, iyy * dy);
    const float S = fmaf(dx, idx, dy * idy);
    const float px = fmaf(5.f * S * iR2, dx, -fmaf(t, dx, 2.f * idx));
    const float py = fmaf(5.f * S * iR2, dy, -fmaf(t, dy, 2.f * idy));
    const float k = 1.5f * iR5;
    const float diag = 5.f * S * iR2 - t;
    const float a = 5.f * iR2, b = 10.f * iR2, c = 10.f * S * iR2 * iR2;
    // J_ij = k [ -a d_j P_i - t d_ij - 2 I_ij + b (Id)_j d_i + 5 S R^-2 d_ij - c d_i d_j ]
    J[0] = k * (-a * dx * px + diag - 2.f * ixx + b * idx * dx - c * dx * dx);
    J[1] = k * (-a * dy * px - 2.f * ixy + b * idy * dx - c * dx * dy);
    J[2] = k * (-a * dx * py - 2.f * ixy + b * idx * dy - c * dy * dx);
    J[3] = k * (-a * dy * py + diag - 2.f * iyy + b * idy * dy - c * dy * dy);
}

// 1 / ell for the distance input: eps, or the root cell side when eps = 0.
__device__ inline float mlp_inv_ell(float eps2f, double root_wsq) {
    return eps2f > 0.f ? __builtin_amdgcn_rsqf(eps2f) : (float)(1.0 / sqrt(root_wsq));
}

// This lane's A tiles, loaded once per batch of tiles (48 VGPRs).
struct MlpWeights {
    s16x4 a1[kMlpTiles1];
    bf16x8 a23[kMlpTiles23];
};

__device__ inline void mlp_load_weights(MlpWeights& w, const void* __restrict__ tiles,
                                        int lane) {
    const uint2* t1 = reinterpret_cast<const uint2*>(tiles);
    const uint4* t23 = reinterpret_cast<const uint4*>(t1 + kMlpTiles1 * 64);
#pragma unroll
    for (int t = 0; t < kMlpTiles1; t++)
        w.a1[t] = __builtin_bit_cast(s16x4, t1[t * 64 + lane]);
#pragma unroll
    for (int t = 0; t < kMlpTiles23; t++)
        w.a23[t] = __builtin_bit_cast(bf16x8, t23[t * 64 + lane]);
}

__device__ inline f32x4 mlp_bias4(const float* sb, int off, int lane) {
    return *reinterpret_cast<const f32x4*>(sb + off + 4 * (lane >> 4));
}

// Forward pass of T independent 16-column tiles (interleaved for the MFMA
// pipeline). x[t] holds this lane's four inputs of tile t (lanes 0-15: f0..f3
// of column l, lanes 16-31: f4..f7 of column l-16, lanes 32-47: f8..f11 of
// column l-32, lanes 48-63: zero).
// y[t]: lanes 0-15 hold (y0, y1, 0, 0) of column l.
template <int T>
__device__ inline void mlp_forward_tiles(const f32x4* x, const MlpWeights& w, const float* sb,
                                         int lane, f32x4* y) {
    s16x4 b0[T];
#pragma unroll
    for (int t = 0; t < T; t++) b0[t] = pack4(x[t]);
    f32x4 h1[T][4];
#pragma unroll
    for (int n = 0; n < 4; n++) {
        const f32x4 c = mlp_bias4(sb, 16 * n, lane);
#pragma unroll
        for (int t = 0; t < T; t++)
            h1[t][n] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(w.a1[n], b0[t], c, 0, 0, 0);
    }
    bf16x8 k1[T][2];
#pragma unroll
    for (int t = 0; t < T; t++) {
        k1[t][0] = pack8_relu(h1[t][0], h1[t][1]);
        k1[t][1] = pack8_relu(h1[t][2], h1[t][3]);
    }
    f32x4 h2[T][4];
#pragma unroll
    for (int n = 0; n < 4; n++) {
        const f32x4 c = mlp_bias4(sb, 64 + 16 * n, lane);
#pragma unroll
        for (int t = 0; t < T; t++) {
            f32x4 a = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w.a23[2 * n], k1[t][0], c, 0, 0, 0);
            h2[t][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w.a23[2 * n + 1], k1[t][1], a, 0,
                                                               0, 0);
        }
    }
    const f32x4 c3 = mlp_bias4(sb, 128, lane);
#pragma unroll
    for (int t = 0; t < T; t++) {
        const bf16x8 k20 = pack8_relu(h2[t][0], h2[t][1]);
        const bf16x8 k21 = pack8_relu(h2[t][2], h2[t][3]);
        const f32x4 a = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w.a23[8], k20, c3, 0, 0, 0);
        y[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w.a23[9], k21, a, 0, 0, 0);
    }
}

}  // namespace nb
