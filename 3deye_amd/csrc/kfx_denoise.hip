// kfx_denoise.hip — learned residual depth denoiser (no reference
// counterpart): forward, backward and Adam, all hand-written for gfx950.
//
// Network: input x0 = depth in metres (0 where invalid), x1 = valid mask;
// four 3x3 convolutions with zero padding, 2 -> 16 -> 16 -> 16 -> 1
// channels, ReLU after the first three; output y = x0 + conv4 on valid pixels.
// 5,089 parameters, flat fp32 (per layer weights [co][ci][ky][kx], then
// biases [co]).  Weights are used in bfloat16, accumulation is fp32.
//
// Layout: hidden activations and their gradients are bfloat16 NHWC with 16
// channels (32 bytes per pixel).  Forward, every layer is an implicit GEMM on
// the matrix cores, one wave computing 16 output channels x 16 consecutive
// pixels: the 16 -> 16 layers and the 16 -> 1 layer (one useful output row)
// with v_mfma_f32_16x16x32_bf16, two 3x3 taps per MFMA (K = 2 taps x 16
// channels; five MFMAs, the last with one tap), the B operand of a lane 8
// channels of one pixel (one 16-byte load); the 2 -> 16 layer with
// v_mfma_f32_16x16x4_f32 on its single-precision inputs (K = 2 channels x 9
// taps, five MFMAs).  Backward, the input gradients of the hidden layers use
// v_mfma_f32_16x16x16_bf16, one MFMA per tap; weight gradients are GEMMs
// over pixels (K = pixels) accumulated per wave, then reduced in a fixed
// order.
#include <math.h>

#include <algorithm>

#include "kfx_common.h"

namespace kfx {

constexpr int kDnC = 16;
constexpr int kDnParams = 5089;
constexpr int kOffW1 = 0, kOffB1 = 288;
constexpr int kOffW2 = 304, kOffB2 = 2608;
constexpr int kOffW3 = 2624, kOffB3 = 4928;
constexpr int kOffW4 = 4944, kOffB4 = 5088;

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

struct DnPacked {
    float w1[16 * 2 * 9];        // [co][ci][tap], bf16-rounded values
    float b1[16];
    uint16_t w2[9][16][16];      // [tap][co][ci] bf16 bits (forward A operand)
    uint16_t w2t[9][16][16];     // [tap][ci][co] (input-gradient A operand)
    float b2[16];
    uint16_t w3[9][16][16];
    uint16_t w3t[9][16][16];
    float b3[16];
    float w4[16 * 9];            // [ci][tap], bf16-rounded values
    float b4;
    // Forward A operands for tap pairs (2 P, 2 P + 1), P < 5: [P][co][k] with
    // k = 16 (tap - 2 P) + ci, bf16 bits, zero for tap 9.
    alignas(16) uint16_t w2p[5][16][32];
    alignas(16) uint16_t w3p[5][16][32];
    alignas(16) uint16_t w4p[5][32];
};
constexpr int kDnPackThreads = 5 * 16 * 32;

__host__ __device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t b = __float_as_uint(f);
    b += 0x7FFFu + ((b >> 16) & 1u);
    return (uint16_t)(b >> 16);
}
__host__ __device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__global__ void dn_pack_kernel(const float* __restrict__ m, DnPacked* __restrict__ pk) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 288) pk->w1[i] = bf2f(f2bf(m[kOffW1 + i]));
    if (i < 16) {
        pk->b1[i] = m[kOffB1 + i];
        pk->b2[i] = m[kOffB2 + i];
        pk->b3[i] = m[kOffB3 + i];
    }
    if (i < 144) pk->w4[i] = bf2f(f2bf(m[kOffW4 + i]));
    if (i == 0) pk->b4 = m[kOffB4];
    if (i < 2304) {
        // i = ((co * 16 + ci) * 9 + tap)
        const int tap = i % 9, ci = (i / 9) % 16, co = i / 144;
        pk->w2[tap][co][ci] = f2bf(m[kOffW2 + i]);
        pk->w2t[tap][ci][co] = f2bf(m[kOffW2 + i]);
        pk->w3[tap][co][ci] = f2bf(m[kOffW3 + i]);
        pk->w3t[tap][ci][co] = f2bf(m[kOffW3 + i]);
    }
    if (i < kDnPackThreads) {
        // i = (P * 16 + co) * 32 + k
        const int k = i % 32, co = (i / 32) % 16, P = i / 512;
        const int tap = 2 * P + k / 16, ci = k % 16;
        const bool real = tap < 9;
        pk->w2p[P][co][k] = real ? f2bf(m[kOffW2 + (co * 16 + ci) * 9 + tap]) : (uint16_t)0;
        pk->w3p[P][co][k] = real ? f2bf(m[kOffW3 + (co * 16 + ci) * 9 + tap]) : (uint16_t)0;
        if (co == 0) pk->w4p[P][k] = real ? f2bf(m[kOffW4 + ci * 9 + tap]) : (uint16_t)0;
    }
}

__device__ __forceinline__ float depth_m(uint16_t d) { return (float)d * 0.001f; }

// Input normalisation, in inverse depth w = 1/z (planes are affine in pixel
// coordinates there, so a smoothing filter of w is unbiased on them, unlike
// one of z): tap q enters as (w_q - w_p) / sigma_w(p), clamped to
// +-kDnRelClip, where p is the output pixel and sigma_w = sigma_z / z^2 is
// the inverse-depth noise of the Kinect model sigma_z = 1.2 + 1.9 (z - 0.4)^2
// mm; 0 where the tap or the centre has no depth.  The output is
// 1 / (w_p + sigma_w(p) s) for the network's residual s (the inverse depth
// kept above w_p / 4), so the network is the same at every distance.
constexpr float kDnRelClip = 8.0f;
__device__ __forceinline__ float noise_sigma_m(float z) {
    const float dz = z - 0.4f;
    return (1.2f + 1.9f * (dz * dz)) * 0.001f;
}
// 1000 / d, correctly rounded without a division: RN(1/d) from rcp_rn and
// Markstein's correction (equal to the IEEE quotient for every d in
// [1, 65535], checked on the CPU); d = 0 gives +inf as the quotient would.
__device__ __forceinline__ float inv_depth(uint16_t d) {
    const float b = (float)d, y = rcp_rn(b), q0 = 1000.0f * y;
    return d ? fmaf(fmaf(-q0, b, 1000.0f), y, q0) : __builtin_inff();
}
// sigma_w(p) = sigma_z(z_p) w_p^2.
__device__ __forceinline__ float noise_sigma_inv(uint16_t d) {
    const float w = inv_depth(d);
    return noise_sigma_m(depth_m(d)) * (w * w);
}
__device__ __forceinline__ float rel_depth(uint16_t dq, float wp, float inv_sigma_w) {
    if (!dq) return 0.0f;
    const float r = (inv_depth(dq) - wp) * inv_sigma_w;
    return fminf(fmaxf(r, -kDnRelClip), kDnRelClip);
}
// Output depth of a pixel with raw depth d > 0 and residual s.
__device__ __forceinline__ float dn_output(uint16_t d, float s) {
    const float wp = inv_depth(d);
    const float w = fmaxf(wp + noise_sigma_inv(d) * s, 0.25f * wp);
    return rcp_rn(w);  // w >= wp / 4 >= 0.0038: inside rcp_rn's exact range
}

// The forward layers on groups of 16 positions: lane l (n = l % 16, q =
// l / 16) supplies position n's operands and ends with the four output
// channels 4 q .. 4 q + 3 of position n (the MFMA D layout).  The per-layer
// kernels and the one-launch inference kernel both evaluate every layer
// through these, so their outputs are equal bit for bit.

// Per-pixel inputs of layer 1: the inverse depth w = 1000 / d (0 where the
// pixel has no depth) and 1 / sigma_w (0 likewise).
__device__ __forceinline__ float dn_w(uint16_t d) { return d ? inv_depth(d) : 0.0f; }
__device__ __forceinline__ float dn_inv_sigma(uint16_t d) { return d ? rcp_rn(noise_sigma_inv(d)) : 0.0f; }

// Layer 1 (2 -> 16): D[co][n] = sum_k W1[co][k] X[k][n], k = ci * 9 + tap,
// with X[ci = 0] the tap's inverse depth relative to the centre's in units
// of the centre's sigma_w, clamped (rel_depth; 0 where the tap or the
// centre has no depth), and X[ci = 1] the tap's valid mask.  w_at(tap)
// gives position n's tap inverse depths (dn_w; tap 4 the centre; 0 outside
// the image), inv_sigma_c the centre's dn_inv_sigma.  Lane l supplies
// k = 4 j + q of MFMA j.
// Lane l's A operands of layer 1 (W1[n][4 j + q], 0 past k = 17).
__device__ __forceinline__ void load_w1(const DnPacked* __restrict__ pk, int lane, float (&a)[5]) {
    const int n = lane & 15, q = lane >> 4;
#pragma unroll
    for (int j = 0; j < 5; ++j) {
        const int k = 4 * j + q;
        a[j] = k < 18 ? pk->w1[n * 18 + k] : 0.0f;
    }
}

template <class WAt>
__device__ __forceinline__ v4f l1_group(WAt w_at, float inv_sigma_c, const float (&a)[5], int lane) {
    const int q = lane >> 4;
    const float wc = w_at(4);
    v4f acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int j = 0; j < 5; ++j) {
        const int k = 4 * j + q;
        float x = 0.0f;
        if (k < 18) {
            const float w = w_at(k % 9);
            if (k < 9) {
                const float r = (w - wc) * inv_sigma_c;
                x = (w != 0.0f && wc != 0.0f) ? fminf(fmaxf(r, -kDnRelClip), kDnRelClip) : 0.0f;
            } else {
                x = w != 0.0f ? 1.0f : 0.0f;
            }
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a[j], x, acc, 0, 0, 0);
    }
    return acc;
}

template <class WAt>
__device__ __forceinline__ v4f l1_group(WAt w_at, float inv_sigma_c, const DnPacked* __restrict__ pk, int lane) {
    float a[5];
    load_w1(pk, lane, a);
    return l1_group(w_at, inv_sigma_c, a, lane);
}

typedef __bf16 v4bf __attribute__((ext_vector_type(4)));
typedef float v4fp __attribute__((ext_vector_type(4)));

// Bias, ReLU and bfloat16 (round to nearest even, v_cvt_pk_bf16_f32) of a
// group result: four channels 4 q .. 4 q + 3.
__device__ __forceinline__ uint2 relu_pack4(const v4f& acc, const float (&bq)[4]) {
    v4fp v;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const float z = acc[i] + bq[i];
        v[i] = z > 0.0f ? z : 0.0f;
    }
    return __builtin_bit_cast(uint2, __builtin_convertvector(v, v4bf));
}
__device__ __forceinline__ void load_bias4(const float* __restrict__ b, int q, float (&bq)[4]) {
#pragma unroll
    for (int i = 0; i < 4; ++i) bq[i] = b[4 * q + i];
}
__device__ __forceinline__ uint2 relu_pack(const v4f& acc, const float* __restrict__ b, int q) {
    float bq[4];
    load_bias4(b, q, bq);
    return relu_pack4(acc, bq);
}

__device__ __forceinline__ v8bf as_v8bf(uint4 u) { return __builtin_bit_cast(v8bf, u); }

// A 16 -> 16 layer or the 16 -> 1 layer: D[co][n] = sum over tap pairs P of
// sum_k Wp[P][co][k] in[n + off(tap)][ci], k = 16 (tap - 2 P) + ci.
// wa[P]: lane l's A operand (Wp[P][n][8 q .. 8 q + 7]); in_at(tap, c0): the
// 8 channels c0 .. c0 + 7 of position n's input at the tap (zero outside
// the image).
template <class InAt>
__device__ __forceinline__ v4f conv_group(InAt in_at, const v8bf (&wa)[5], int lane) {
    const int q = lane >> 4;
    v4f acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int P = 0; P < 5; ++P) {
        const int tap = 2 * P + (q >> 1);
        const uint4 u = tap < 9 ? in_at(tap, 8 * (q & 1)) : make_uint4(0, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa[P], as_v8bf(u), acc, 0, 0, 0);
    }
    return acc;
}

// A operands of a 16 -> 16 layer (wp = w2p or w3p) or of the 16 -> 1 layer
// (wp4 = w4p: row co = 0 only, the other rows zero).
__device__ __forceinline__ void load_wa(const uint16_t (*wp)[16][32], int lane, v8bf (&wa)[5]) {
    const int n = lane & 15, q = lane >> 4;
#pragma unroll
    for (int P = 0; P < 5; ++P) wa[P] = as_v8bf(*reinterpret_cast<const uint4*>(&wp[P][n][8 * q]));
}
__device__ __forceinline__ void load_wa4(const uint16_t (*wp4)[32], int lane, v8bf (&wa)[5]) {
    const int n = lane & 15, q = lane >> 4;
#pragma unroll
    for (int P = 0; P < 5; ++P)
        wa[P] = as_v8bf(n == 0 ? *reinterpret_cast<const uint4*>(&wp4[P][8 * q]) : make_uint4(0, 0, 0, 0));
}

// Output depth of the residual s of a pixel with raw depth d: metres
// (ypred) and millimetres (out), both optional.
__device__ __forceinline__ void dn_store_output(uint16_t d, float s, int p, float* __restrict__ ypred,
                                                uint16_t* __restrict__ out) {
    const float y = d ? dn_output(d, s) : 0.0f;
    if (ypred) ypred[p] = y;
    if (out) {
        int mm = 0;
        if (d) mm = (int)fminf(fmaxf(rintf(y * 1000.0f), 1.0f), 65535.0f);
        out[p] = (uint16_t)mm;
    }
}

// The 8 bf16 channels c0 .. c0 + 7 of pixel (qx, qy) of an NHWC activation;
// zero outside the image.
__device__ __forceinline__ uint4 load8(const uint16_t* __restrict__ act, int W, int H, int qx, int qy, int c0) {
    if (qx < 0 || qy < 0 || qx >= W || qy >= H) return make_uint4(0, 0, 0, 0);
    return *reinterpret_cast<const uint4*>(act + (size_t)(qy * W + qx) * kDnC + c0);
}

// 4 bf16 channels [c0, c0+4) of pixel (qx, qy) as a 16x16x16 MFMA operand
// (the backward input gradients); zero outside the image.
__device__ __forceinline__ v4s load4(const uint16_t* __restrict__ act, int W, int H, int qx, int qy, int c0) {
    v4s r = {0, 0, 0, 0};
    if (qx >= 0 && qy >= 0 && qx < W && qy < H) {
        const uint2 u = *reinterpret_cast<const uint2*>(act + (size_t)(qy * W + qx) * kDnC + c0);
        r[0] = (short)(u.x & 0xffffu);
        r[1] = (short)(u.x >> 16);
        r[2] = (short)(u.y & 0xffffu);
        r[3] = (short)(u.y >> 16);
    }
    return r;
}

__device__ __forceinline__ v4s wload4(const uint16_t* w) {
    const uint2 u = *reinterpret_cast<const uint2*>(w);
    v4s r;
    r[0] = (short)(u.x & 0xffffu);
    r[1] = (short)(u.x >> 16);
    r[2] = (short)(u.y & 0xffffu);
    r[3] = (short)(u.y >> 16);
    return r;
}

// Per-layer forward kernels (training): a wave computes 16 consecutive
// pixels of a row, four waves per workgroup.
__global__ void __launch_bounds__(256) dn_l1_fwd(const uint16_t* __restrict__ depth, int W, int H,
                                                 const DnPacked* __restrict__ pk, uint16_t* __restrict__ a1) {
    const int lane = threadIdx.x & 63;
    const int x0 = (blockIdx.x * 4 + (threadIdx.x >> 6)) * 16;
    const int y = blockIdx.y;
    if (x0 >= W) return;
    const int n = lane & 15, q = lane >> 4;
    const int x = x0 + n;
    auto w_at = [&](int t) -> float {
        const int qx = x + t % 3 - 1, qy = y + t / 3 - 1;
        return dn_w((qx >= 0 && qy >= 0 && qx < W && qy < H) ? depth[qy * W + qx] : (uint16_t)0);
    };
    const uint16_t dc = (x < W) ? depth[y * W + x] : (uint16_t)0;
    const v4f acc = l1_group(w_at, dn_inv_sigma(dc), pk, lane);
    if (x < W) *reinterpret_cast<uint2*>(a1 + (size_t)(y * W + x) * kDnC + 4 * q) = relu_pack(acc, pk->b1, q);
}

// Hidden layer (16 -> 16): out[p][co] = relu(b[co] + sum_tap sum_ci
// W[tap][co][ci] in[p + off(tap)][ci]).
__global__ void __launch_bounds__(256) dn_mid_fwd(const uint16_t* __restrict__ in, uint16_t* __restrict__ out, int W,
                                                  int H, const uint16_t (*wp)[16][32], const float* __restrict__ b) {
    const int lane = threadIdx.x & 63;
    const int x0 = (blockIdx.x * 4 + (threadIdx.x >> 6)) * 16;
    const int y = blockIdx.y;
    if (x0 >= W) return;
    const int n = lane & 15, q = lane >> 4;
    v8bf wa[5];
    load_wa(wp, lane, wa);
    auto in_at = [&](int t, int c0) { return load8(in, W, H, x0 + n + t % 3 - 1, y + t / 3 - 1, c0); };
    const v4f acc = conv_group(in_at, wa, lane);
    const int px = x0 + n;
    if (px < W) *reinterpret_cast<uint2*>(out + (size_t)(y * W + px) * kDnC + 4 * q) = relu_pack(acc, b, q);
}

// Layer 4 (16 -> 1) and the output: s = b4 + sum w4 a3 (row 0 of the
// group result, held by lanes 0..15).
__global__ void __launch_bounds__(256) dn_l4_fwd(const uint16_t* __restrict__ depth, const uint16_t* __restrict__ a3,
                                                 int W, int H, const DnPacked* __restrict__ pk,
                                                 float* __restrict__ ypred, uint16_t* __restrict__ out) {
    const int lane = threadIdx.x & 63;
    const int x0 = (blockIdx.x * 4 + (threadIdx.x >> 6)) * 16;
    const int y = blockIdx.y;
    if (x0 >= W) return;
    const int n = lane & 15;
    v8bf wa[5];
    load_wa4(pk->w4p, lane, wa);
    auto in_at = [&](int t, int c0) { return load8(a3, W, H, x0 + n + t % 3 - 1, y + t / 3 - 1, c0); };
    const v4f acc = conv_group(in_at, wa, lane);
    const int px = x0 + n;
    if (lane < 16 && px < W) dn_store_output(depth[y * W + px], acc[0] + pk->b4, y * W + px, ypred, out);
}

// ---------------------------------------------------------------------------
// Inference: the four layers in one launch.  A workgroup owns a tile of
// kDfTX x kDfTY output pixels and computes, in LDS, layer 1 on the tile and
// a 3-pixel halo, layer 2 on a 2-pixel halo, layer 3 on a 1-pixel halo, then
// layer 4 and the output depth; the hidden activations never leave the CU.
// Every layer goes through the per-layer kernels' group functions, and
// positions outside the image hold 0 (the zero padding), so the output
// equals the four-launch forward bit for bit.  Tiles of 32 x 20 with 512
// threads (61 KB of LDS, two workgroups per CU): 480 workgroups at 640x480,
// so no CU holds more than two (32 x 16 tiles made 600, and the 88 CUs that
// held three set the time).
// ---------------------------------------------------------------------------
constexpr int kDfTX = 32, kDfTY = 20, kDfThreads = 512;
constexpr int kDfDX = kDfTX + 8, kDfDY = kDfTY + 8;   // depth tile (4-pixel halo)
constexpr int kDfA1X = kDfTX + 6, kDfA1Y = kDfTY + 6; // layer-1 outputs
constexpr int kDfA2X = kDfTX + 4, kDfA2Y = kDfTY + 4; // layer-2 outputs
constexpr int kDfA3X = kDfTX + 2, kDfA3Y = kDfTY + 2; // layer-3 outputs (in the layer-1 buffer)
static_assert(kDfA3X * kDfA3Y <= kDfA1X * kDfA1Y, "layer 3 reuses the layer-1 buffer");

// A 16 -> 16 layer on the tile: out (ox x oy positions, image origin gx0,
// gy0) from in (ox + 2 wide), in groups of 16 consecutive positions of the
// flattened ox x oy tile (a group may span two rows: each lane addresses its
// own position).
__device__ __forceinline__ void df_mid_layer(const uint16_t* __restrict__ in, uint16_t* __restrict__ out, int ox,
                                             int oy, int gx0, int gy0, int W, int H, const v8bf (&wa)[5],
                                             const float (&bq)[4]) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int n = lane & 15, q = lane >> 4;
    const int npos = ox * oy;
    const int ix = ox + 2;
    for (int g = wave; g < (npos + 15) / 16; g += kDfThreads / 64) {
        const int i = min(16 * g + n, npos - 1);  // the partial last group reads a valid position
        const int ly = i / ox, lx = i - ly * ox;
        auto in_at = [&](int t, int c0) {
            return *reinterpret_cast<const uint4*>(in + ((ly + t / 3) * ix + lx + t % 3) * kDnC + c0);
        };
        const v4f acc = conv_group(in_at, wa, lane);
        if (16 * g + n >= npos) continue;
        const int gx = gx0 + lx, gy = gy0 + ly;
        const uint2 v = (gx >= 0 && gy >= 0 && gx < W && gy < H) ? relu_pack4(acc, bq) : make_uint2(0, 0);
        *reinterpret_cast<uint2*>(out + i * kDnC + 4 * q) = v;
    }
}

__global__ void __launch_bounds__(kDfThreads) dn_fused_fwd(const uint16_t* __restrict__ depth, int W, int H,
                                                           const DnPacked* __restrict__ pk,
                                                           float* __restrict__ ypred, uint16_t* __restrict__ out) {
    __shared__ uint16_t sd[kDfDY * kDfDX];
    __shared__ __align__(16) uint16_t s1[kDfA1Y * kDfA1X * kDnC];
    __shared__ __align__(16) uint16_t s2[kDfA2Y * kDfA2X * kDnC];
    const int x0 = blockIdx.x * kDfTX, y0 = blockIdx.y * kDfTY;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int n = lane & 15, q = lane >> 4;
    // The depth tile, and per tile pixel dn_w and dn_inv_sigma in the
    // layer-2 buffer (free until layer 2).
    float* sw = reinterpret_cast<float*>(s2);
    float* sis = sw + kDfDX * kDfDY;
    static_assert(2 * kDfDX * kDfDY * sizeof(float) <= sizeof(s2), "layer-1 inputs fit the layer-2 buffer");
    for (int i = threadIdx.x; i < kDfDX * kDfDY; i += kDfThreads) {
        const int qx = x0 - 4 + i % kDfDX, qy = y0 - 4 + i / kDfDX;
        const uint16_t d = (qx >= 0 && qy >= 0 && qx < W && qy < H) ? depth[qy * W + qx] : (uint16_t)0;
        sd[i] = d;
        sw[i] = dn_w(d);
        sis[i] = dn_inv_sigma(d);
    }
    __syncthreads();
    // Layer 1 on groups of 16 consecutive positions of the flattened
    // layer-1 tile (0 outside the image).
    constexpr int kA1 = kDfA1X * kDfA1Y;
    {
        float a1[5], b1[4];
        load_w1(pk, lane, a1);
        load_bias4(pk->b1, q, b1);
        for (int g = wave; g < (kA1 + 15) / 16; g += kDfThreads / 64) {
            const int i = min(16 * g + n, kA1 - 1);  // the partial last group reads a valid position
            const int lx = i % kDfA1X, ly = i / kDfA1X;
            auto w_at = [&](int t) -> float { return sw[(ly + t / 3) * kDfDX + lx + t % 3]; };
            const v4f acc = l1_group(w_at, sis[(ly + 1) * kDfDX + lx + 1], a1, lane);
            if (16 * g + n >= kA1) continue;
            const int x = x0 - 3 + lx, y = y0 - 3 + ly;
            const uint2 v = (x >= 0 && y >= 0 && x < W && y < H) ? relu_pack4(acc, b1) : make_uint2(0, 0);
            *reinterpret_cast<uint2*>(s1 + i * kDnC + 4 * q) = v;
        }
    }
    v8bf wa[5];
    float bq[4];
    load_wa(pk->w2p, lane, wa);
    load_bias4(pk->b2, q, bq);
    __syncthreads();  // layer 1 done: s2 free for layer 2's outputs
    df_mid_layer(s1, s2, kDfA2X, kDfA2Y, x0 - 2, y0 - 2, W, H, wa, bq);
    load_wa(pk->w3p, lane, wa);
    load_bias4(pk->b3, q, bq);
    __syncthreads();
    df_mid_layer(s2, s1, kDfA3X, kDfA3Y, x0 - 1, y0 - 1, W, H, wa, bq);
    load_wa4(pk->w4p, lane, wa);
    __syncthreads();
    // Layer 4 and the output, 16 pixels of a tile row per group.
    for (int g = wave; g < kDfTY * (kDfTX / 16); g += kDfThreads / 64) {
        const int ly = g / (kDfTX / 16), sx = (g % (kDfTX / 16)) * 16;
        auto in_at = [&](int t, int c0) {
            return *reinterpret_cast<const uint4*>(s1 + ((ly + t / 3) * kDfA3X + sx + n + t % 3) * kDnC + c0);
        };
        const v4f acc = conv_group(in_at, wa, lane);
        const int x = x0 + sx + n, yy = y0 + ly;
        if (lane < 16 && x < W && yy < H)
            dn_store_output(sd[(ly + 4) * kDfDX + sx + n + 4], acc[0] + pk->b4, yy * W + x, ypred, out);
    }
}

// ---------------------------------------------------------------------------
// Training.
// ---------------------------------------------------------------------------

// Loss gradient of the output layer: with y = 1 / (w_p + sigma_w s),
// dz4[p] = dL/ds = -2 (y - t) y^2 sigma_w on pixels with a valid input, a
// valid target and |z - t| < outlier (the clamp of the inverse depth is
// passed straight through); per-block sums of squared error and counts
// (reduced by dn_loss_reduce).
__global__ void __launch_bounds__(256) dn_loss_kernel(const uint16_t* __restrict__ depth,
                                                      const float* __restrict__ ypred,
                                                      const float* __restrict__ target, int P, float outlier,
                                                      float* __restrict__ dz4, double* __restrict__ blk_sse,
                                                      unsigned* __restrict__ blk_cnt) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    double e2 = 0.0;
    unsigned c = 0;
    if (p < P) {
        const uint16_t d = depth[p];
        const float t = target[p];
        float g = 0.0f;
        if (d && t > 0.0f && fabsf(depth_m(d) - t) < outlier) {
            const float yv = ypred[p];
            const float e = yv - t;
            g = -(2.0f * e) * (yv * yv) * noise_sigma_inv(d);
            e2 = (double)e * e;
            c = 1;
        }
        dz4[p] = g;
    }
    for (int off = 32; off > 0; off >>= 1) {
        e2 += __shfl_xor(e2, off);
        c += __shfl_xor(c, off);
    }
    __shared__ double se[4];
    __shared__ unsigned sc[4];
    if ((threadIdx.x & 63) == 0) {
        se[threadIdx.x >> 6] = e2;
        sc[threadIdx.x >> 6] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        blk_sse[blockIdx.x] = se[0] + se[1] + se[2] + se[3];
        blk_cnt[blockIdx.x] = sc[0] + sc[1] + sc[2] + sc[3];
    }
}

// stats[0] = sum of squared error, stats[1] = count (as double), stats[2] =
// gradient scale 1 / count (0 when empty).
__global__ void __launch_bounds__(256) dn_loss_reduce(const double* __restrict__ blk_sse,
                                                      const unsigned* __restrict__ blk_cnt, int nblk,
                                                      double* __restrict__ stats) {
    double s = 0.0, c = 0.0;
    for (int i = threadIdx.x; i < nblk; i += 256) {
        s += blk_sse[i];
        c += (double)blk_cnt[i];
    }
    for (int off = 32; off > 0; off >>= 1) {
        s += __shfl_xor(s, off);
        c += __shfl_xor(c, off);
    }
    __shared__ double ss[4], cc[4];
    if ((threadIdx.x & 63) == 0) {
        ss[threadIdx.x >> 6] = s;
        cc[threadIdx.x >> 6] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const double S = ss[0] + ss[1] + ss[2] + ss[3], Cn = cc[0] + cc[1] + cc[2] + cc[3];
        stats[0] = S;
        stats[1] = Cn;
        stats[2] = Cn > 0.0 ? 1.0 / Cn : 0.0;
    }
}

// Input gradient of layer 4 through the ReLU of layer 3:
// dz3[p][ci] = (a3[p][ci] > 0) * sum_tap dz4[p - off(tap)] w4[ci][tap].
__global__ void __launch_bounds__(256) dn_l4_bwd_input(const float* __restrict__ dz4, const uint16_t* __restrict__ a3,
                                                       const uint16_t* __restrict__ depth, int W, int H,
                                                       const DnPacked* __restrict__ pk, uint16_t* __restrict__ dz3) {
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= W || y >= H) return;
    float g[9];
    for (int ky = 0; ky < 3; ++ky)
        for (int kx = 0; kx < 3; ++kx) {
            // output pixel q = p - (kx-1, ky-1) used input p through tap (ky, kx)
            const int qx = x - (kx - 1), qy = y - (ky - 1);
            float v = 0.0f;
            if (qx >= 0 && qy >= 0 && qx < W && qy < H && depth[qy * W + qx]) v = dz4[qy * W + qx];
            g[ky * 3 + kx] = v;
        }
    const size_t base = (size_t)(y * W + x) * kDnC;
    uint32_t packed[8];
    for (int c = 0; c < 16; c += 2) {
        float r[2];
        for (int h = 0; h < 2; ++h) {
            float s = 0.0f;
            for (int t = 0; t < 9; ++t) s += g[t] * pk->w4[(c + h) * 9 + t];
            r[h] = (a3[base + c + h] & 0x7fffu) != 0 ? s : 0.0f;
        }
        packed[c / 2] = (uint32_t)f2bf(r[0]) | ((uint32_t)f2bf(r[1]) << 16);
    }
    uint4* dst = reinterpret_cast<uint4*>(dz3 + base);
    dst[0] = make_uint4(packed[0], packed[1], packed[2], packed[3]);
    dst[1] = make_uint4(packed[4], packed[5], packed[6], packed[7]);
}

// Hidden-layer input gradient through the previous ReLU (MFMA):
// dz_in[p][ci] = (a_in[p][ci] > 0) * sum_tap sum_co W[tap][co][ci] dz[p - off(tap)][co].
// A[m = ci][k = co] from the transposed pack, B[k = co][n = pixel].
__global__ void __launch_bounds__(256) dn_mid_bwd_input(const uint16_t* __restrict__ dz, const uint16_t* __restrict__ a_in,
                                                        int W, int H, const uint16_t* __restrict__ wt,
                                                        uint16_t* __restrict__ dz_in) {
    const int lane = threadIdx.x & 63;
    const int x0 = (blockIdx.x * 4 + (threadIdx.x >> 6)) * 16;
    const int y = blockIdx.y;
    if (x0 >= W) return;
    const int n = lane & 15, kq = (lane >> 4) * 4;
    v4f acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int tap = 0; tap < 9; ++tap) {
        const int ky = tap / 3, kx = tap % 3;
        const v4s a = wload4(wt + (tap * 16 + n) * 16 + kq);
        const v4s bb = load4(dz, W, H, x0 + n - (kx - 1), y - (ky - 1), kq);
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bb, acc, 0, 0, 0);
    }
    const int px = x0 + n;
    if (px >= W) return;
    const size_t base = (size_t)(y * W + px) * kDnC + kq;
    const uint2 am = *reinterpret_cast<const uint2*>(a_in + base);
    const uint16_t ah[4] = {(uint16_t)(am.x & 0xffffu), (uint16_t)(am.x >> 16), (uint16_t)(am.y & 0xffffu),
                            (uint16_t)(am.y >> 16)};
    uint16_t r[4];
    for (int i = 0; i < 4; ++i) r[i] = (ah[i] & 0x7fffu) != 0 ? f2bf(acc[i]) : (uint16_t)0;
    *reinterpret_cast<uint2*>(dz_in + base) =
        make_uint2((uint32_t)r[0] | ((uint32_t)r[1] << 16), (uint32_t)r[2] | ((uint32_t)r[3] << 16));
}

// Hidden-layer weight gradient (MFMA, K = pixels): per wave, over the
// 16-pixel groups [g0, g1) of row y:
//   dW[tap][co][ci] += sum_p dz[p][co] a_in[p + off(tap)][ci],  db[co] += sum_p dz[p][co].
// A[m = co][k = pixel], B[k = pixel][n = ci].  Writes the wave's partial
// (9*256 weights in [co][ci][tap] order, then 16 biases) to part[wave].
constexpr int kMidGroupsPerWave = 10;
__global__ void __launch_bounds__(64) dn_mid_wgrad(const uint16_t* __restrict__ dz, const uint16_t* __restrict__ a_in,
                                                   int W, int H, float* __restrict__ part) {
    const int lane = threadIdx.x;
    const int groups = (W + 15) / 16;
    const int seg_per_row = (groups + kMidGroupsPerWave - 1) / kMidGroupsPerWave;
    const int wave = blockIdx.x;
    const int y = wave / seg_per_row, seg = wave % seg_per_row;
    float* out = part + (size_t)wave * 2320;
    if (y >= H) return;
    const int g0 = seg * kMidGroupsPerWave, g1 = min(groups, g0 + kMidGroupsPerWave);
    const int m = lane & 15, kq = (lane >> 4) * 4;
    v4f acc[9];
    for (int t = 0; t < 9; ++t) acc[t] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
    float bsum = 0.0f;
    for (int g = g0; g < g1; ++g) {
        const int xb = g * 16;
        // A: dz[pixel xb + kq + j][co = m], j = 0..3
        v4s a;
        for (int j = 0; j < 4; ++j) {
            const int px = xb + kq + j;
            const uint16_t h = px < W ? dz[(size_t)(y * W + px) * kDnC + m] : (uint16_t)0;
            a[j] = (short)h;
            bsum += bf2f(h);
        }
#pragma unroll
        for (int tap = 0; tap < 9; ++tap) {
            const int ky = tap / 3, kx = tap % 3;
            v4s bb;
            for (int j = 0; j < 4; ++j) {
                const int px = xb + kq + j, qx = px + kx - 1, qy = y + ky - 1;
                uint16_t h = 0;
                if (px < W && qx >= 0 && qy >= 0 && qx < W && qy < H) h = a_in[(size_t)(qy * W + qx) * kDnC + m];
                bb[j] = (short)h;
            }
            acc[tap] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bb, acc[tap], 0, 0, 0);
        }
    }
    // D[co = kq + i][ci = m] per tap
    for (int tap = 0; tap < 9; ++tap)
        for (int i = 0; i < 4; ++i) out[((kq + i) * 16 + m) * 9 + tap] = acc[tap][i];
    // bias: lanes with the same m hold partial sums over disjoint pixels
    bsum += __shfl_xor(bsum, 16);
    bsum += __shfl_xor(bsum, 32);
    if (lane < 16) out[2304 + lane] = bsum;
}

// Small-layer weight gradients on the vector ALUs, one wave per row segment:
// layer 4: dW4[ci][tap] = sum_p dz4[p] a3[p + off][ci], db4 = sum dz4 (145);
// layer 1: dW1[co][ci][tap] = sum_p dz1[p][co] x[p + off][ci], db1[co] (304),
// split in two halves of 8 output channels (grid.y).
constexpr int kSmallPixelsPerWave = 320;
__global__ void __launch_bounds__(64) dn_l4_wgrad(const float* __restrict__ dz4, const uint16_t* __restrict__ a3,
                                                  const uint16_t* __restrict__ depth, int W, int H,
                                                  float* __restrict__ part) {
    const int lane = threadIdx.x;
    const int seg_per_row = (W + kSmallPixelsPerWave - 1) / kSmallPixelsPerWave;
    const int y = blockIdx.x / seg_per_row, seg = blockIdx.x % seg_per_row;
    float* out = part + (size_t)blockIdx.x * 145;
    if (y >= H) return;
    float acc[145];
    for (int i = 0; i < 145; ++i) acc[i] = 0.0f;
    for (int x = seg * kSmallPixelsPerWave + lane; x < min(W, (seg + 1) * kSmallPixelsPerWave); x += 64) {
        const int p = y * W + x;
        const float g = depth[p] ? dz4[p] : 0.0f;
        if (g == 0.0f) continue;
        acc[144] += g;
        for (int ky = 0; ky < 3; ++ky)
            for (int kx = 0; kx < 3; ++kx) {
                const int qx = x + kx - 1, qy = y + ky - 1;
                if (qx < 0 || qy < 0 || qx >= W || qy >= H) continue;
                const uint16_t* src = a3 + (size_t)(qy * W + qx) * kDnC;
                for (int c = 0; c < 16; ++c) acc[c * 9 + ky * 3 + kx] += g * bf2f(src[c]);
            }
    }
    for (int i = 0; i < 145; ++i) {
        float v = acc[i];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        if (lane == 0) out[i] = v;
    }
}

__global__ void __launch_bounds__(64) dn_l1_wgrad(const uint16_t* __restrict__ dz1, const uint16_t* __restrict__ depth,
                                                  int W, int H, float* __restrict__ part) {
    const int lane = threadIdx.x;
    const int half = blockIdx.y;  // output channels [8 half, 8 half + 8)
    const int seg_per_row = (W + kSmallPixelsPerWave - 1) / kSmallPixelsPerWave;
    const int y = blockIdx.x / seg_per_row, seg = blockIdx.x % seg_per_row;
    float* out = part + ((size_t)blockIdx.x * 2 + half) * 152;
    if (y >= H) return;
    float acc[152];  // [co 8][ci 2][tap 9] then 8 biases
    for (int i = 0; i < 152; ++i) acc[i] = 0.0f;
    for (int x = seg * kSmallPixelsPerWave + lane; x < min(W, (seg + 1) * kSmallPixelsPerWave); x += 64) {
        const int p = y * W + x;
        const uint16_t dc = depth[p];
        const float wc = dc ? inv_depth(dc) : 0.0f;
        const float inv_sigma = dc ? rcp_rn(noise_sigma_inv(dc)) : 0.0f;
        float g[8];
        bool any = false;
        for (int c = 0; c < 8; ++c) {
            g[c] = bf2f(dz1[(size_t)p * kDnC + 8 * half + c]);
            any |= g[c] != 0.0f;
        }
        if (!any) continue;
        for (int c = 0; c < 8; ++c) acc[144 + c] += g[c];
        for (int ky = 0; ky < 3; ++ky)
            for (int kx = 0; kx < 3; ++kx) {
                const int qx = x + kx - 1, qy = y + ky - 1;
                if (qx < 0 || qy < 0 || qx >= W || qy >= H) continue;
                const uint16_t d = depth[qy * W + qx];
                if (!d) continue;
                const float x0 = dc ? rel_depth(d, wc, inv_sigma) : 0.0f;
                for (int c = 0; c < 8; ++c) {
                    acc[(c * 2 + 0) * 9 + ky * 3 + kx] += g[c] * x0;
                    acc[(c * 2 + 1) * 9 + ky * 3 + kx] += g[c];
                }
            }
    }
    for (int i = 0; i < 152; ++i) {
        float v = acc[i];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        if (lane == 0) out[i] = v;
    }
}

// Input gradient of layer 2 into layer 1's activation: the hidden-layer
// kernel with layer 2's transposed weights (dn_mid_bwd_input).

// Sums the per-wave partials of one layer in wave order (deterministic) and
// accumulates scale * sum into grad (fp32).  Layout maps the partial index to
// the flat parameter index.
__global__ void __launch_bounds__(256) dn_reduce_mid(const float* __restrict__ part, int nwave, int woff, int boff,
                                                     const double* __restrict__ stats, float* __restrict__ grad) {
    const int i = blockIdx.x * 256 + threadIdx.x;  // 0 .. 2319
    if (i >= 2320) return;
    double s = 0.0;
    for (int w = 0; w < nwave; ++w) s += part[(size_t)w * 2320 + i];
    const int dst = i < 2304 ? woff + i : boff + (i - 2304);
    grad[dst] += (float)(s * stats[2]);
}

__global__ void __launch_bounds__(256) dn_reduce_l4(const float* __restrict__ part, int nwave,
                                                    const double* __restrict__ stats, float* __restrict__ grad) {
    const int i = blockIdx.x * 256 + threadIdx.x;  // 0 .. 144
    if (i >= 145) return;
    double s = 0.0;
    for (int w = 0; w < nwave; ++w) s += part[(size_t)w * 145 + i];
    grad[i < 144 ? kOffW4 + i : kOffB4] += (float)(s * stats[2]);
}

__global__ void __launch_bounds__(256) dn_reduce_l1(const float* __restrict__ part, int nwave,
                                                    const double* __restrict__ stats, float* __restrict__ grad) {
    const int i = blockIdx.x * 256 + threadIdx.x;  // 0 .. 303
    if (i >= 304) return;
    // parameter i -> (half, local index)
    int half, local;
    if (i < 288) {
        const int co = i / 18, rem = i % 18;  // [co][ci][tap]
        half = co / 8;
        local = (co % 8) * 18 + rem;
    } else {
        const int co = i - 288;
        half = co / 8;
        local = 144 + co % 8;
    }
    double s = 0.0;
    for (int w = 0; w < nwave; ++w) s += part[((size_t)w * 2 + half) * 152 + local];
    grad[kOffW1 + i] += (float)(s * stats[2]);
}

// Adam (Kingma & Ba, 2015) on the fp32 master weights; grad is the sum of
// per-frame mean gradients over `frames` frames and is cleared afterwards.
__global__ void dn_adam_kernel(float* __restrict__ w, float* __restrict__ grad, float* __restrict__ m,
                               float* __restrict__ v, int n, float lr, float b1, float b2, float eps, float c1,
                               float c2, float inv_frames) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float g = grad[i] * inv_frames;
    const float mi = b1 * m[i] + (1.0f - b1) * g;
    const float vi = b2 * v[i] + (1.0f - b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    grad[i] = 0.0f;
}

}  // namespace kfx

struct kfx_denoiser {
    int W = 0, H = 0;
    float* master = nullptr;      // kDnParams
    float* grad = nullptr;
    float* adam_m = nullptr;
    float* adam_v = nullptr;
    kfx::DnPacked* pk = nullptr;
    uint16_t *a1 = nullptr, *a2 = nullptr, *a3 = nullptr;  // activations
    uint16_t *g1 = nullptr, *g2 = nullptr, *g3 = nullptr;  // gradients of the hidden layers
    float* ypred = nullptr;
    float* dz4 = nullptr;
    double* blk_sse = nullptr;
    unsigned* blk_cnt = nullptr;
    double* stats = nullptr;
    double* host_stats = nullptr;
    float* part = nullptr;
    size_t part_floats = 0;
    int adam_t = 0;
};

using namespace kfx;

namespace {

int mid_waves(int W, int H) {
    const int groups = (W + 15) / 16;
    return H * ((groups + kMidGroupsPerWave - 1) / kMidGroupsPerWave);
}
int small_waves(int W, int H) { return H * ((W + kSmallPixelsPerWave - 1) / kSmallPixelsPerWave); }

int dn_forward(kfx_denoiser* d, const uint16_t* depth, uint16_t* out, float* ypred) {
    const int W = d->W, H = d->H;
    dim3 mid_grid(divup(W, 64), H);
    ProfScope prof("denoise_forward");
    dn_l1_fwd<<<mid_grid, 256, 0, stream()>>>(depth, W, H, d->pk, d->a1);
    KFX_CHECK_LAUNCH();
    dn_mid_fwd<<<mid_grid, 256, 0, stream()>>>(d->a1, d->a2, W, H, d->pk->w2p, d->pk->b2);
    KFX_CHECK_LAUNCH();
    dn_mid_fwd<<<mid_grid, 256, 0, stream()>>>(d->a2, d->a3, W, H, d->pk->w3p, d->pk->b3);
    KFX_CHECK_LAUNCH();
    dn_l4_fwd<<<mid_grid, 256, 0, stream()>>>(depth, d->a3, W, H, d->pk, ypred, out);
    KFX_CHECK_LAUNCH();
    return KFX_OK;
}

// Inference forward in one launch (same output as dn_forward).
int dn_forward_fused(kfx_denoiser* d, const uint16_t* depth, uint16_t* out, float* ypred) {
    ProfScope prof("denoise_forward");
    dim3 grid(divup(d->W, kDfTX), divup(d->H, kDfTY));
    dn_fused_fwd<<<grid, kDfThreads, 0, stream()>>>(depth, d->W, d->H, d->pk, ypred, out);
    KFX_CHECK_LAUNCH();
    return KFX_OK;
}

int dn_pack(kfx_denoiser* d) {
    dn_pack_kernel<<<divup(kDnPackThreads, 256), 256, 0, stream()>>>(d->master, d->pk);
    KFX_CHECK_LAUNCH();
    return KFX_OK;
}

}  // namespace

extern "C" {

int kfx_denoiser_param_count(void) { return kDnParams; }

int kfx_denoiser_create(int cols, int rows, kfx_denoiser** out) {
    KFX_REQUIRE_INIT();
    KFX_REQUIRE(out && cols > 0 && rows > 0, "kfx_denoiser_create: bad arguments");
    kfx_denoiser* d = new kfx_denoiser();
    d->W = cols;
    d->H = rows;
    const size_t P = (size_t)cols * rows;
    const int nblk = (int)divup((unsigned)P, 256u);
    d->part_floats = std::max({(size_t)mid_waves(cols, rows) * 2320, (size_t)small_waves(cols, rows) * 145,
                               (size_t)small_waves(cols, rows) * 2 * 152});
    hipError_t e = hipMalloc((void**)&d->master, sizeof(float) * kDnParams);
    if (e == hipSuccess) e = hipMalloc((void**)&d->grad, sizeof(float) * kDnParams);
    if (e == hipSuccess) e = hipMalloc((void**)&d->adam_m, sizeof(float) * kDnParams);
    if (e == hipSuccess) e = hipMalloc((void**)&d->adam_v, sizeof(float) * kDnParams);
    if (e == hipSuccess) e = hipMalloc((void**)&d->pk, sizeof(DnPacked));
    uint16_t** acts[6] = {&d->a1, &d->a2, &d->a3, &d->g1, &d->g2, &d->g3};
    for (int i = 0; i < 6 && e == hipSuccess; ++i) e = hipMalloc((void**)acts[i], 2 * kDnC * P);
    if (e == hipSuccess) e = hipMalloc((void**)&d->ypred, sizeof(float) * P);
    if (e == hipSuccess) e = hipMalloc((void**)&d->dz4, sizeof(float) * P);
    if (e == hipSuccess) e = hipMalloc((void**)&d->blk_sse, sizeof(double) * nblk);
    if (e == hipSuccess) e = hipMalloc((void**)&d->blk_cnt, sizeof(unsigned) * nblk);
    if (e == hipSuccess) e = hipMalloc((void**)&d->stats, sizeof(double) * 4);
    if (e == hipSuccess) e = hipHostMalloc((void**)&d->host_stats, sizeof(double) * 4, hipHostMallocDefault);
    if (e == hipSuccess) e = hipMalloc((void**)&d->part, sizeof(float) * d->part_floats);
    if (e == hipSuccess) e = hipMemsetAsync(d->master, 0, sizeof(float) * kDnParams, stream());
    if (e == hipSuccess) e = hipMemsetAsync(d->grad, 0, sizeof(float) * kDnParams, stream());
    if (e == hipSuccess) e = hipMemsetAsync(d->adam_m, 0, sizeof(float) * kDnParams, stream());
    if (e == hipSuccess) e = hipMemsetAsync(d->adam_v, 0, sizeof(float) * kDnParams, stream());
    if (e == hipSuccess) e = hipStreamSynchronize(stream());
    if (e != hipSuccess) {
        kfx_denoiser_destroy(d);
        return set_error(KFX_ERR_HIP, "kfx_denoiser_create: %s", hipGetErrorString(e));
    }
    int rc = dn_pack(d);
    if (rc) {
        kfx_denoiser_destroy(d);
        return rc;
    }
    *out = d;
    return KFX_OK;
}

int kfx_denoiser_destroy(kfx_denoiser* d) {
    if (!d) return KFX_OK;
    if (initialised()) (void)hipStreamSynchronize(stream());
    void* ptrs[] = {d->master, d->grad, d->adam_m, d->adam_v, d->pk, d->a1, d->a2, d->a3, d->g1, d->g2, d->g3,
                    d->ypred, d->dz4, d->blk_sse, d->blk_cnt, d->stats, d->part};
    for (void* p : ptrs) (void)hipFree(p);
    if (d->host_stats) (void)hipHostFree(d->host_stats);
    delete d;
    return KFX_OK;
}

int kfx_denoiser_set_params(kfx_denoiser* d, const float* host_params) {
    KFX_REQUIRE_INIT();
    KFX_REQUIRE(d && host_params, "kfx_denoiser_set_params: null pointer");
    KFX_CHECK_HIP(hipMemcpyAsync(d->master, host_params, sizeof(float) * kDnParams, hipMemcpyHostToDevice, stream()));
    KFX_CHECK_HIP(hipStreamSynchronize(stream()));
    d->adam_t = 0;
    KFX_CHECK_HIP(hipMemsetAsync(d->adam_m, 0, sizeof(float) * kDnParams, stream()));
    KFX_CHECK_HIP(hipMemsetAsync(d->adam_v, 0, sizeof(float) * kDnParams, stream()));
    KFX_CHECK_HIP(hipMemsetAsync(d->grad, 0, sizeof(float) * kDnParams, stream()));
    return dn_pack(d);
}

int kfx_denoiser_get_params(kfx_denoiser* d, float* host_params) {
    KFX_REQUIRE_INIT();
    KFX_REQUIRE(d && host_params, "kfx_denoiser_get_params: null pointer");
    KFX_CHECK_HIP(hipMemcpyAsync(host_params, d->master, sizeof(float) * kDnParams, hipMemcpyDeviceToHost, stream()));
    KFX_CHECK_HIP(hipStreamSynchronize(stream()));
    return KFX_OK;
}

int kfx_denoiser_forward(kfx_denoiser* d, const uint16_t* depth, uint16_t* out_depth, float* out_metres) {
    KFX_REQUIRE_INIT();
    KFX_REQUIRE(d && depth, "kfx_denoiser_forward: null pointer");
    return dn_forward_fused(d, depth, out_depth, out_metres);
}

int kfx_denoiser_forward_layers(kfx_denoiser* d, const uint16_t* depth, uint16_t* out_depth, float* out_metres) {
    KFX_REQUIRE_INIT();
    KFX_REQUIRE(d && depth, "kfx_denoiser_forward_layers: null pointer");
    return dn_forward(d, depth, out_depth, out_metres);
}

int kfx_denoiser_accumulate(kfx_denoiser* d, const uint16_t* depth, const float* target_m, float outlier_m,
                            double* loss, uint64_t* count) {
    KFX_REQUIRE_INIT();
    KFX_REQUIRE(d && depth && target_m, "kfx_denoiser_accumulate: null pointer");
    const int W = d->W, H = d->H, P = W * H;
    int rc = dn_forward(d, depth, nullptr, d->ypred);
    if (rc) return rc;
    ProfScope prof("denoise_backward");
    const int nblk = (int)divup((unsigned)P, 256u);
    dn_loss_kernel<<<nblk, 256, 0, stream()>>>(depth, d->ypred, target_m, P, outlier_m, d->dz4, d->blk_sse, d->blk_cnt);
    KFX_CHECK_LAUNCH();
    dn_loss_reduce<<<1, 256, 0, stream()>>>(d->blk_sse, d->blk_cnt, nblk, d->stats);
    KFX_CHECK_LAUNCH();
    dim3 pix_grid(divup(W, 64), divup(H, 4));
    dim3 mid_grid(divup(W, 64), H);
    const int ms = mid_waves(W, H), ss = small_waves(W, H);
    // layer 4
    dn_l4_wgrad<<<ss, 64, 0, stream()>>>(d->dz4, d->a3, depth, W, H, d->part);
    KFX_CHECK_LAUNCH();
    dn_reduce_l4<<<1, 256, 0, stream()>>>(d->part, ss, d->stats, d->grad);
    KFX_CHECK_LAUNCH();
    dn_l4_bwd_input<<<pix_grid, 256, 0, stream()>>>(d->dz4, d->a3, depth, W, H, d->pk, d->g3);
    KFX_CHECK_LAUNCH();
    // layer 3: weights, then input gradient into layer 2's activation
    dn_mid_wgrad<<<ms, 64, 0, stream()>>>(d->g3, d->a2, W, H, d->part);
    KFX_CHECK_LAUNCH();
    dn_reduce_mid<<<divup(2320, 256), 256, 0, stream()>>>(d->part, ms, kOffW3, kOffB3, d->stats, d->grad);
    KFX_CHECK_LAUNCH();
    dn_mid_bwd_input<<<mid_grid, 256, 0, stream()>>>(d->g3, d->a2, W, H, &d->pk->w3t[0][0][0], d->g2);
    KFX_CHECK_LAUNCH();
    // layer 2
    dn_mid_wgrad<<<ms, 64, 0, stream()>>>(d->g2, d->a1, W, H, d->part);
    KFX_CHECK_LAUNCH();
    dn_reduce_mid<<<divup(2320, 256), 256, 0, stream()>>>(d->part, ms, kOffW2, kOffB2, d->stats, d->grad);
    KFX_CHECK_LAUNCH();
    dn_mid_bwd_input<<<mid_grid, 256, 0, stream()>>>(d->g2, d->a1, W, H, &d->pk->w2t[0][0][0], d->g1);
    KFX_CHECK_LAUNCH();
    // layer 1
    dn_l1_wgrad<<<dim3(ss, 2), 64, 0, stream()>>>(d->g1, depth, W, H, d->part);
    KFX_CHECK_LAUNCH();
    dn_reduce_l1<<<divup(304, 256), 256, 0, stream()>>>(d->part, ss, d->stats, d->grad);
    KFX_CHECK_LAUNCH();
    if (loss || count) {
        KFX_CHECK_HIP(hipMemcpyAsync(d->host_stats, d->stats, sizeof(double) * 3, hipMemcpyDeviceToHost, stream()));
        KFX_CHECK_HIP(hipStreamSynchronize(stream()));
        if (loss) *loss = d->host_stats[1] > 0 ? d->host_stats[0] / d->host_stats[1] : 0.0;
        if (count) *count = (uint64_t)d->host_stats[1];
    }
    return KFX_OK;
}

int kfx_denoiser_get_grad(kfx_denoiser* d, float* host_grad) {
    KFX_REQUIRE_INIT();
    KFX_REQUIRE(d && host_grad, "kfx_denoiser_get_grad: null pointer");
    KFX_CHECK_HIP(hipMemcpyAsync(host_grad, d->grad, sizeof(float) * kDnParams, hipMemcpyDeviceToHost, stream()));
    KFX_CHECK_HIP(hipStreamSynchronize(stream()));
    return KFX_OK;
}

int kfx_denoiser_zero_grad(kfx_denoiser* d) {
    KFX_REQUIRE_INIT();
    KFX_REQUIRE(d != nullptr, "kfx_denoiser_zero_grad: null pointer");
    return kfx_fill_u8(d->grad, 0, sizeof(float) * kDnParams);
}

int kfx_denoiser_adam_step(kfx_denoiser* d, float lr, int frames) {
    KFX_REQUIRE_INIT();
    KFX_REQUIRE(d && frames > 0 && lr > 0.0f, "kfx_denoiser_adam_step: bad arguments");
    d->adam_t += 1;
    // The loss is in m^2, so gradients are ~1e-6 and below: eps must sit far
    // under sqrt(v) for Adam to stay scale-free.
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-12f;
    const float c1 = 1.0f - powf(b1, (float)d->adam_t), c2 = 1.0f - powf(b2, (float)d->adam_t);
    dn_adam_kernel<<<divup(kDnParams, 256), 256, 0, stream()>>>(d->master, d->grad, d->adam_m, d->adam_v, kDnParams,
                                                                lr, b1, b2, eps, c1, c2, 1.0f / (float)frames);
    KFX_CHECK_LAUNCH();
    return dn_pack(d);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Model-depth helpers: training targets and the fallback gate.
// ---------------------------------------------------------------------------
namespace kfx {

// z-depth in the camera of each raycast vertex (volume coordinates) at the
// pose (R camera -> volume, t); 0 where the raycast found no surface.
__global__ void __launch_bounds__(256) model_depth_kernel(const kfx_xyz* __restrict__ vmap, int P, kfx_pose_f pose,
                                                          float* __restrict__ out) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= P) return;
    const kfx_xyz v = vmap[p];
    float z = 0.0f;
    if (!isnan(v.x)) {
        const float dx = v.x - pose.t[0], dy = v.y - pose.t[1], dz = v.z - pose.t[2];
        z = pose.R[2] * dx + pose.R[5] * dy + pose.R[8] * dz;  // third row of R^T
        if (!(z > 0.0f)) z = 0.0f;
    }
    out[p] = z;
}

// Consistency of a depth frame with the model: every raycast vertex is
// brought into the camera at the device pose and projected; where the frame
// has depth at that pixel and the two differ by less than max_diff, the
// squared difference is accumulated.  stats: [sum of squares (m^2), count].
// Consistency of a depth frame with the raycast model, per 256-pixel
// workgroup: the model vertices (volume coordinates) projected into the
// camera at the device pose (12 floats: R camera -> volume, then t).  Every
// valid vertex counts towards n_valid; where it lands on a pixel with depth,
// towards n_depth, and when the depth is within max_diff of the vertex depth
// its squared difference and a count are added too.
// partials[g] = {sum e^2, count, n_depth, n_valid}.
__global__ void __launch_bounds__(256) depth_consistency_kernel(const kfx_xyz* __restrict__ vmap,
                                                                const float* __restrict__ dpose,
                                                                const uint16_t* __restrict__ depth,
                                                                kfx_intrinsics in, float max_diff,
                                                                double* __restrict__ partials) {
    __shared__ double red[4][4];
    double v[4];
    gate_terms(vmap, dpose, depth, in, max_diff, blockIdx.x * 256 + threadIdx.x, v);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int k = 0; k < 4; ++k) v[k] = wave_sum_xor(v[k], lane);
    if (lane == 0)
        for (int k = 0; k < 4; ++k) red[wave][k] = v[k];
    __syncthreads();
    if (threadIdx.x < 4) {
        const int k = threadIdx.x;
        partials[blockIdx.x * 4 + k] = ((red[0][k] + red[1][k]) + red[2][k]) + red[3][k];
    }
}

// stats[k] = sum over workgroups of partials[g][k], in a fixed order (one
// workgroup: thread t sums g = t, t + 256, ..., then a tree over the 256).
// With a flag, also the fallback decision (1: fall back) for a gate of
// fallback_m RMS: at least 100 valid model vertices, and an RMS difference
// over the pixels within max_diff above fallback_m or fewer than half of the
// valid vertices within max_diff.
__global__ void __launch_bounds__(256) consistency_reduce_kernel(const double* __restrict__ partials, int nblk,
                                                                 double* __restrict__ stats, double fallback_m,
                                                                 int32_t* __restrict__ flag,
                                                                 int32_t* __restrict__ host_flag) {
    __shared__ double red[4][256];
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    for (int g = threadIdx.x; g < nblk; g += 256)
        for (int k = 0; k < 4; ++k) a[k] += partials[g * 4 + k];
    for (int k = 0; k < 4; ++k) red[k][threadIdx.x] = a[k];
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h)
            for (int k = 0; k < 4; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x < 4) stats[threadIdx.x] = red[threadIdx.x][0];
    if (threadIdx.x == 0 && flag != nullptr) {
        const int32_t f = gate_decision(red[0][0], red[1][0], red[3][0], fallback_m);
        *flag = f;
        if (host_flag != nullptr) *host_flag = f;
    }
}

int depth_consistency(const Workspace& ws, const kfx_xyz* vmap_model, const float* dev_pose, const uint16_t* depth,
                      const kfx_intrinsics* intr, float max_diff, double* dev_stats, double fallback_m,
                      int32_t* dev_flag, int32_t* host_flag) {
    const int P = intr->cols * intr->rows;
    const int nblk = (int)divup((unsigned)P, 256u);
    KFX_REQUIRE(ws.cons != nullptr && (size_t)nblk <= ws.cons_cap,
                "kfx_depth_consistency: workspace holds %zu workgroup partials, %dx%d frames need %d", ws.cons_cap,
                intr->cols, intr->rows, nblk);
    {
        ProfScope prof("consistency");
        depth_consistency_kernel<<<nblk, 256, 0, stream()>>>(vmap_model, dev_pose, depth, *intr, max_diff, ws.cons);
        KFX_CHECK_LAUNCH();
    }
    ProfScope prof("consistency_reduce");
    consistency_reduce_kernel<<<1, 256, 0, stream()>>>(ws.cons, nblk, dev_stats, fallback_m, dev_flag, host_flag);
    KFX_CHECK_LAUNCH();
    return KFX_OK;
}

}  // namespace kfx

extern "C" {

int kfx_model_depth(const kfx_xyz* vmap_model, const kfx_pose* pose_volume, int cols, int rows, float* out_m) {
    KFX_REQUIRE_INIT();
    KFX_REQUIRE(vmap_model && pose_volume && out_m && cols > 0 && rows > 0, "kfx_model_depth: bad arguments");
    kfx_pose_f pf;
    for (int i = 0; i < 9; ++i) pf.R[i] = (float)pose_volume->R[i];
    for (int i = 0; i < 3; ++i) pf.t[i] = (float)pose_volume->t[i];
    const int P = cols * rows;
    model_depth_kernel<<<divup(P, 256), 256, 0, stream()>>>(vmap_model, P, pf, out_m);
    KFX_CHECK_LAUNCH();
    return KFX_OK;
}

int kfx_depth_consistency(const kfx_xyz* vmap_model, const float* dev_pose, const uint16_t* depth,
                          const kfx_intrinsics* intr, float max_diff_m, double* dev_stats) {
    KFX_REQUIRE_INIT();
    KFX_REQUIRE(vmap_model && dev_pose && depth && intr && dev_stats, "kfx_depth_consistency: null pointer");
    KFX_TRY(library_workspace_reserve(intr->cols, intr->rows, false, true));
    return depth_consistency(library_workspace(), vmap_model, dev_pose, depth, intr, max_diff_m, dev_stats, 0.0,
                             nullptr, nullptr);
}

}  // extern "C"
