// Learned warm start for the relaxation loop (no counterpart in the
// reference): a per-segment MLP 64 -> 64 -> 64 -> 16 (51 used inputs, 3 used
// outputs) with BF16 operands and FP32 accumulation on CDNA4 MFMA, trained on
// the GPU from trajectories of the plain relaxer, and a proposal step that
// moves vertices by the predicted displacement. The plain projector always
// runs after a proposal and judges admissibility; the host reverts a
// proposal that raises the residual.
//
// Features of a segment (row of 64 bf16): for its 16 nearest listed
// neighbours by surface gap, the unit direction from its closest point to
// the neighbour's times min((r1 + r2) / distance, 8) (zero when absent;
// above one when overlapping); the unit tangent; and the segment's recent
// motion: with h = (midpoint now - midpoint at the last history mark) /
// diameter and m = |h|, the direction h / m and log10(m) / 8 + 1 (all zero
// when m is below 1e-12 or no mark exists). Target: the midpoint's
// displacement over the next interval in units of m diameters, so the
// network predicts the coming motion relative to the recent one, at every
// stage of the relaxation alike.
//
// Matrix-core tiles: v_mfma_f32_16x16x16_bf16 with, for lane l,
//   A (16x16, MxK): row l%16, k = 4*(l/16) + j       (4 bf16 per lane)
//   B (16x16, KxN): col l%16, k = 4*(l/16) + j
//   D (16x16, MxN): col l%16, row = 4*(l/16) + j     (4 f32 per lane)
// Weights are kept as f32 masters plus two bf16 copies: [out][in] (forward,
// B operand with k = in) and [in][out] (backward data, k = out).

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "learned_device.h"
#include "relax_device.h"

namespace tgh {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float x) {  // round to nearest even
    uint32_t u = __float_as_uint(x);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

__device__ __forceinline__ v4f mfma(v4s a, v4s b, v4f c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ v4s load4(const uint16_t *p) {  // 4 consecutive bf16
    const uint2 u = *(const uint2 *)p;
    v4s r;
    r[0] = (short)(u.x & 0xffffu);
    r[1] = (short)(u.x >> 16);
    r[2] = (short)(u.y & 0xffffu);
    r[3] = (short)(u.y >> 16);
    return r;
}

__device__ __forceinline__ v4s gather4(const uint16_t *p, uint32_t stride) {  // 4 bf16 at a stride
    v4s r;
    r[0] = (short)p[0];
    r[1] = (short)p[stride];
    r[2] = (short)p[2 * stride];
    r[3] = (short)p[3 * stride];
    return r;
}

// ---------------------------------------------------------------------------
// MFMA layout probe: D = A(16x16) * B(16x16), A row-major bf16, B row-major
// bf16 (K x N), D row-major f32. One wave.
__global__ void k_mfma_probe(const uint16_t *A, const uint16_t *B, float *D) {
    const uint32_t l = threadIdx.x;
    const uint32_t r = l % 16, q = l / 16;
    v4s a = load4(A + r * 16 + 4 * q);
    v4s b = gather4(B + (4 * q) * 16 + r, 16);
    v4f acc = {0.f, 0.f, 0.f, 0.f};
    acc = mfma(a, b, acc);
    for (int j = 0; j < 4; ++j) D[(4 * q + j) * 16 + r] = acc[j];
}

// ---------------------------------------------------------------------------
// Features.

__device__ __forceinline__ void closest_offset(const GridParams &g, float4 P1, float4 Q1, float4 P2, float4 Q2,
                                               float &ox, float &oy, float &oz, float &dist) {
    const float eps_len = 1.0e-20f, eps_par = 1.0e-6f;
    float p1x = P1.x, p1y = P1.y, p1z = P1.z, q1x = Q1.x, q1y = Q1.y, q1z = Q1.z;
    float p2x = P2.x, p2y = P2.y, p2z = P2.z, q2x = Q2.x, q2y = Q2.y, q2z = Q2.z;
    float mid[3] = {0.5f * (p2x + q2x - p1x - q1x), 0.5f * (p2y + q2y - p1y - q1y), 0.5f * (p2z + q2z - p1z - q1z)};
    float sh[3] = {0.f, 0.f, 0.f};
    for (int k = 0; k < 3; ++k)
        if (g.periodic[k]) sh[k] = roundf(mid[k] / g.length[k]) * g.length[k];
    p2x -= sh[0]; q2x -= sh[0];
    p2y -= sh[1]; q2y -= sh[1];
    p2z -= sh[2]; q2z -= sh[2];
    float d1x = q1x - p1x, d1y = q1y - p1y, d1z = q1z - p1z;
    float d2x = q2x - p2x, d2y = q2y - p2y, d2z = q2z - p2z;
    float rx = p1x - p2x, ry = p1y - p2y, rz = p1z - p2z;
    float a = d1x * d1x + d1y * d1y + d1z * d1z, e = d2x * d2x + d2y * d2y + d2z * d2z;
    float f = d2x * rx + d2y * ry + d2z * rz;
    float s = 0.f, t = 0.f;
    if (a <= eps_len && e > eps_len) {
        t = fminf(fmaxf(f / e, 0.f), 1.f);
    } else if (a > eps_len) {
        float c = d1x * rx + d1y * ry + d1z * rz;
        if (e <= eps_len) {
            s = fminf(fmaxf(-c / a, 0.f), 1.f);
        } else {
            float b = d1x * d2x + d1y * d2y + d1z * d2z;
            float den = a * e - b * b;
            if (fabsf(den) > eps_par * a * e) s = fminf(fmaxf((b * f - c * e) / den, 0.f), 1.f);
            float proj = (b * s + f) / e;
            if (proj < 0.f) {
                s = fminf(fmaxf(-c / a, 0.f), 1.f);
            } else if (proj > 1.f) {
                t = 1.f;
                s = fminf(fmaxf((b - c) / a, 0.f), 1.f);
            } else {
                t = proj;
            }
        }
    }
    ox = (p2x + d2x * t) - (p1x + d1x * s);
    oy = (p2y + d2y * t) - (p1y + d1y * s);
    oz = (p2z + d2z * t) - (p1z + d1z * s);
    dist = sqrtf(ox * ox + oy * oy + oz * oz);
}

// Lists hold cell-list slots; every segment's geometry is read from the
// compact slot records (first vertex and radius, second vertex and segment)
// and the specimen from slotC, which the refresh before this kernel writes
// from `pos`.
__global__ __launch_bounds__(256) void k_features(const GridParams *grids, const float4 *cP, const float4 *cQ,
                                                  const uint4 *slotC, const uint32_t *ncount, const uint32_t *nbr,
                                                  const uint32_t *nbr_off,
                                                  const uint32_t *ovf, const float4 *history, uint16_t *X,
                                                  float *scale, uint32_t block, uint32_t ovf_cap, uint32_t S) {
    const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= S) return;
    const uint4 me = slotC[slot];
    if (me.x == kNoSegment) return;  // inactive segments keep zero rows
    const uint32_t seg = me.x;
    const GridParams &g = grids[me.w];
    const float4 P1 = cP[slot], Q1 = cQ[slot];
    const float r1 = __uint_as_float(me.z);
    const uint32_t count = ncount[slot];
    const uint32_t *list = count <= block ? nbr + nbr_off[slot]
                                          : (count <= ovf_cap ? ovf + (size_t)slot * ovf_cap : nullptr);
    const uint32_t n = list ? count : 0u;
    float gap[kFeatureNeighbors];
    uint32_t who[kFeatureNeighbors];
#pragma unroll
    for (int k = 0; k < (int)kFeatureNeighbors; ++k) {
        gap[k] = 3.0e38f;
        who[k] = 0xffffffffu;
    }
    for (uint32_t k = 0; k < n; ++k) {
        const uint32_t other = list[k];
        float ox, oy, oz, dist;
        const float4 P2 = cP[other];
        closest_offset(g, P1, Q1, P2, cQ[other], ox, oy, oz, dist);
        float gk = dist - (r1 + P2.w);
        uint32_t wk = other;
        // keep the 16 smallest gaps in ascending order (unrolled insertion)
#pragma unroll
        for (int i = 0; i < (int)kFeatureNeighbors; ++i) {
            if (gk < gap[i]) {
                float tg = gap[i];
                uint32_t tw = who[i];
                gap[i] = gk;
                who[i] = wk;
                gk = tg;
                wk = tw;
            }
        }
    }
    uint16_t *row = X + (size_t)seg * kFeatureWidth;
    for (uint32_t k = 0; k < kFeatureNeighbors; ++k) {
        float fx = 0.f, fy = 0.f, fz = 0.f;
        if (who[k] != 0xffffffffu) {
            float ox, oy, oz, dist;
            const float4 P2 = cP[who[k]];
            closest_offset(g, P1, Q1, P2, cQ[who[k]], ox, oy, oz, dist);
            const float contact = r1 + P2.w;
            const float d = fmaxf(dist, 1.0e-3f * contact);
            const float w = fminf(contact / d, 8.0f) / d;
            fx = ox * w;
            fy = oy * w;
            fz = oz * w;
        }
        row[3 * k] = f2bf(fx);
        row[3 * k + 1] = f2bf(fy);
        row[3 * k + 2] = f2bf(fz);
    }
    float tx = Q1.x - P1.x, ty = Q1.y - P1.y, tz = Q1.z - P1.z;
    float inv = 1.0f / fmaxf(sqrtf(tx * tx + ty * ty + tz * tz), 1.0e-30f);
    row[kFeatTangent] = f2bf(tx * inv);
    row[kFeatTangent + 1] = f2bf(ty * inv);
    row[kFeatTangent + 2] = f2bf(tz * inv);
    float hx = 0.f, hy = 0.f, hz = 0.f, hl = 0.f, m = 0.f;
    if (history) {
        const float4 h = history[seg];
        const float diameter = 2.0f * r1;
        hx = (0.5f * (P1.x + Q1.x) - h.x) / diameter;
        hy = (0.5f * (P1.y + Q1.y) - h.y) / diameter;
        hz = (0.5f * (P1.z + Q1.z) - h.z) / diameter;
        m = sqrtf(hx * hx + hy * hy + hz * hz);
        if (m > kMinMotion) {
            hx /= m;
            hy /= m;
            hz /= m;
            hl = log10f(m) * 0.125f + 1.0f;
        } else {
            hx = hy = hz = 0.f;
            m = 0.f;
        }
    }
    row[kFeatMotion] = f2bf(hx);
    row[kFeatMotion + 1] = f2bf(hy);
    row[kFeatMotion + 2] = f2bf(hz);
    row[kFeatMotionLog] = f2bf(hl);
    for (uint32_t k = kFeatMotionLog + 1; k < kFeatureWidth; ++k) row[k] = 0;
    if (scale) scale[seg] = m * 2.0f * r1;
}

// Segment midpoints; .w = the unit targets are measured in (the diameter
// times the segment's recent motion when `scale` is given).
__global__ void k_midpoints(const float4 *pos, const uint4 *sstat, const float *scale, float4 *mid, uint32_t S) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) return;
    const uint4 me = sstat[s];
    const float4 p = pos[me.x], q = pos[me.x + 1];
    const float unit = scale ? scale[s] : 2.0f * __uint_as_float(me.z);
    mid[s] = make_float4(0.5f * (p.x + q.x), 0.5f * (p.y + q.y), 0.5f * (p.z + q.z), unit);
}

// Target row = (midpoint now - midpoint then) / unit, each component
// clamped to [-kTargetClamp, kTargetClamp] (a segment that barely moved
// before may move far more next: such rows would swamp the loss); zero when
// unit is 0.
constexpr float kTargetClamp = 4.0f;
__global__ void k_targets(const float4 *pos, const uint4 *sstat, const float4 *mid_then, float4 *T, uint32_t S) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) return;
    const uint4 me = sstat[s];
    const float4 p = pos[me.x], q = pos[me.x + 1], m = mid_then[s];
    if (!(m.w > 0.f)) {
        T[s] = make_float4(0.f, 0.f, 0.f, 0.f);
        return;
    }
    const float inv = 1.0f / m.w;
    const float c = kTargetClamp;
    T[s] = make_float4(fminf(fmaxf((0.5f * (p.x + q.x) - m.x) * inv, -c), c),
                       fminf(fmaxf((0.5f * (p.y + q.y) - m.y) * inv, -c), c),
                       fminf(fmaxf((0.5f * (p.z + q.z) - m.z) * inv, -c), c), 0.f);
}

// ---------------------------------------------------------------------------
// MLP on matrix cores. One wave processes 16-row tiles.

struct __align__(16) TileLds {
    uint16_t x[16 * 64];
    uint16_t h1[16 * 64];
    uint16_t h2[16 * 64];
    uint16_t dy[16 * 16];
    uint16_t dz2[16 * 64];
    uint16_t dz1[16 * 64];
};

// One dense layer on a 16-row tile: D[16 x 16*NT] = A[16 x 64] * W^T with W
// stored [out][64] bf16; adds bias, applies ReLU when `relu`, writes the
// result as bf16 into `out_lds` (16 x 64, row-major) and returns the f32
// accumulators (for the last layer).
template <int NT>
__device__ __forceinline__ void dense(const uint16_t *a_lds, const uint16_t *W, const float *bias, bool relu,
                                      uint16_t *out_lds, v4f acc[NT]) {
    const uint32_t l = threadIdx.x & 63u, r = l % 16, q = l / 16;
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) acc[nt] = (v4f){0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kt = 0; kt < 4; ++kt) {
        v4s a = load4(a_lds + r * 64 + 16 * kt + 4 * q);
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            v4s b = load4(W + (16 * nt + r) * 64 + 16 * kt + 4 * q);
            acc[nt] = mfma(a, b, acc[nt]);
        }
    }
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
        const float bb = bias[16 * nt + r];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            float v = acc[nt][j] + bb;
            if (relu) v = fmaxf(v, 0.f);
            acc[nt][j] = v;
            if (out_lds) out_lds[(4 * q + j) * 64 + 16 * nt + r] = f2bf(v);
        }
    }
}

// Inference in the transposed form: each layer computes Z^T = W * H^T on
// 16-row tiles, so a layer's accumulators (lane: column = tile row l%16,
// rows = outputs 4(l/16)+j) are already the next layer's B operand (k =
// outputs 4(l/16)+j, column = tile row), converted to bf16 in registers.
// Every weight sits in registers for the life of the wave (the A operands,
// loaded once), so a tile costs one 2 KiB read of its feature rows, 36
// MFMAs and the 3 outputs of each row; no LDS. Layer 1's inputs are taken
// in the order lane (r, q) reads them — its row's 16 contiguous features
// 16q .. 16q+15 — and its weights in the same order, which leaves the
// product unchanged.
__device__ __forceinline__ v4s pack4(float a, float b, float c, float d) {
    v4s r;
    r[0] = (short)f2bf(a);
    r[1] = (short)f2bf(b);
    r[2] = (short)f2bf(c);
    r[3] = (short)f2bf(d);
    return r;
}

__global__ __launch_bounds__(256) void k_mlp_forward(const MlpDevice m, const uint16_t *X, float4 *Y, uint32_t rows) {
    const uint32_t l = threadIdx.x & 63u, r = l % 16, q = l / 16;
    const uint32_t wave = blockIdx.x * (blockDim.x / 64u) + threadIdx.x / 64u;
    const uint32_t waves = gridDim.x * (blockDim.x / 64u);
    // A operands: W1 with layer 1's input order, W2, W3; biases per output row
    v4s w1[4][4], w2[4][4], w3[4];
    float b1[4][4], b2[4][4], b3[4];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
#pragma unroll
        for (int kt = 0; kt < 4; ++kt) {
            w1[mt][kt] = load4(m.w1 + (16 * mt + r) * 64 + 16 * q + 4 * kt);
            w2[mt][kt] = load4(m.w2 + (16 * mt + r) * 64 + 16 * kt + 4 * q);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            b1[mt][j] = m.b1[16 * mt + 4 * q + j];
            b2[mt][j] = m.b2[16 * mt + 4 * q + j];
        }
    }
#pragma unroll
    for (int kt = 0; kt < 4; ++kt) w3[kt] = load4(m.w3 + r * 64 + 16 * kt + 4 * q);
#pragma unroll
    for (int j = 0; j < 4; ++j) b3[j] = m.b3[4 * q + j];
    const uint32_t tiles = (rows + 15u) / 16u;
    for (uint32_t tile = wave; tile < tiles; tile += waves) {
        const uint32_t row = tile * 16u + r;
        v4s x[4];
        if (row < rows) {
            const uint4 *src = (const uint4 *)(X + (size_t)row * 64 + 16 * q);
            const uint4 u0 = src[0], u1 = src[1];
            const uint32_t w[8] = {u0.x, u0.y, u0.z, u0.w, u1.x, u1.y, u1.z, u1.w};
#pragma unroll
            for (int kt = 0; kt < 4; ++kt) {
                x[kt][0] = (short)(w[2 * kt] & 0xffffu);
                x[kt][1] = (short)(w[2 * kt] >> 16);
                x[kt][2] = (short)(w[2 * kt + 1] & 0xffffu);
                x[kt][3] = (short)(w[2 * kt + 1] >> 16);
            }
        } else {
#pragma unroll
            for (int kt = 0; kt < 4; ++kt) x[kt] = (v4s){0, 0, 0, 0};
        }
        v4s h1[4], h2[4];
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kt = 0; kt < 4; ++kt) acc = mfma(w1[mt][kt], x[kt], acc);
            h1[mt] = pack4(fmaxf(acc[0] + b1[mt][0], 0.f), fmaxf(acc[1] + b1[mt][1], 0.f),
                           fmaxf(acc[2] + b1[mt][2], 0.f), fmaxf(acc[3] + b1[mt][3], 0.f));
        }
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kt = 0; kt < 4; ++kt) acc = mfma(w2[mt][kt], h1[kt], acc);
            h2[mt] = pack4(fmaxf(acc[0] + b2[mt][0], 0.f), fmaxf(acc[1] + b2[mt][1], 0.f),
                           fmaxf(acc[2] + b2[mt][2], 0.f), fmaxf(acc[3] + b2[mt][3], 0.f));
        }
        v4f y = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kt = 0; kt < 4; ++kt) y = mfma(w3[kt], h2[kt], y);
        // lanes 0..15 hold outputs 0..3 of tile row r
        if (q == 0 && row < rows) Y[row] = make_float4(y[0] + b3[0], y[1] + b3[1], y[2] + b3[2], 0.f);
    }
}

// Forward and backward over the rows idx[begin..end) (all rows when idx is
// null) of one minibatch; each workgroup (one wave) accumulates its weight
// gradients over its tiles and writes them, with its loss sum, to its own
// slot of `partial` (no float atomics; reduced in a fixed order afterwards).
// Loss = sum over rows and the 3 outputs of (y - t)^2, scaled by `inv_rows`.
__global__ __launch_bounds__(64) void k_mlp_backward(const MlpDevice m, const uint16_t *X, const float4 *T,
                                                     const uint32_t *idx, uint32_t rows, float inv_rows,
                                                     float *partial) {
    __shared__ TileLds t;
    __shared__ float bias_grad[64 + 64 + 16];
    const uint32_t l = threadIdx.x, r = l % 16, q = l / 16;
    v4f g1[4][4], g2[4][4], g3[4];  // dW1, dW2 as 4x4 16x16 tiles [out tile][in tile]; dW3 [in tile]
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        g3[i] = (v4f){0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int k = 0; k < 4; ++k) g1[i][k] = g2[i][k] = (v4f){0.f, 0.f, 0.f, 0.f};
    }
    for (uint32_t i = l; i < 144; i += 64) bias_grad[i] = 0.f;
    float loss = 0.f;
    for (uint32_t tile = blockIdx.x; tile * 16 < rows; tile += gridDim.x) {
        const uint32_t base = tile * 16;
        for (uint32_t i = l; i < 16 * 64; i += 64) {
            const uint32_t row = base + i / 64;
            uint16_t v = 0;
            if (row < rows) v = X[(size_t)(idx ? idx[row] : row) * 64 + i % 64];
            t.x[i] = v;
        }
        __syncthreads();
        v4f a1[4], a2[4], a3[1];
        dense<4>(t.x, m.w1, m.b1, true, t.h1, a1);
        __syncthreads();
        dense<4>(t.h1, m.w2, m.b2, true, t.h2, a2);
        __syncthreads();
        dense<1>(t.h2, m.w3, m.b3, false, nullptr, a3);
        // dY = 2 (y - t) / rows for the 3 used outputs, rows 4q + j, column r
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint32_t row = base + 4 * q + j;
            float d = 0.f;
            if (r < 3 && row < rows) {
                const float4 tt = T[idx ? idx[row] : row];
                const float target = r == 0 ? tt.x : (r == 1 ? tt.y : tt.z);
                const float e = a3[0][j] - target;
                loss += e * e;
                d = 2.0f * e * inv_rows;
            }
            t.dy[(4 * q + j) * 16 + r] = f2bf(d);
        }
        __syncthreads();
        // dW3[out][in] += dY^T H2: A = dY^T (row = out = r, k = rows 4q..), B = H2 (k = rows, col = in)
        {
            v4s a = gather4(t.dy + (4 * q) * 16 + r, 16);
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) {
                v4s b = gather4(t.h2 + (4 * q) * 64 + 16 * nt + r, 64);
                g3[nt] = mfma(a, b, g3[nt]);
            }
        }
        // db3 and dH2 = dY W3 (W3 as [in][out] bf16: k = out), masked by H2 > 0
        if (l < 16) {
            float s = 0.f;
            for (int row = 0; row < 16; ++row) s += bf2f(t.dy[row * 16 + l]);
            bias_grad[128 + l] += s;
        }
        {
            v4s a = load4(t.dy + r * 16 + 4 * q);  // row r, k = out 4q..4q+3
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) {
                v4s b = load4(m.w3n + (16 * nt + r) * 16 + 4 * q);  // [in][out]
                v4f d = mfma(a, b, (v4f){0.f, 0.f, 0.f, 0.f});
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const uint32_t row = 4 * q + j, col = 16 * nt + r;
                    const float mask = bf2f(t.h2[row * 64 + col]) > 0.f ? 1.f : 0.f;
                    t.dz2[row * 64 + col] = f2bf(d[j] * mask);
                }
            }
        }
        __syncthreads();
        // dW2[out][in] += dZ2^T H1 ; db2 ; dH1 = dZ2 W2 masked by H1 > 0
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            v4s a = gather4(t.dz2 + (4 * q) * 64 + 16 * mt + r, 64);
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) {
                v4s b = gather4(t.h1 + (4 * q) * 64 + 16 * nt + r, 64);
                g2[mt][nt] = mfma(a, b, g2[mt][nt]);
            }
        }
        {
            float s = 0.f;
            for (int row = 0; row < 16; ++row) s += bf2f(t.dz2[row * 64 + l]);
            bias_grad[64 + l] += s;
        }
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            v4f d = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kt = 0; kt < 4; ++kt) {
                v4s a = load4(t.dz2 + r * 64 + 16 * kt + 4 * q);
                v4s b = load4(m.w2n + (16 * nt + r) * 64 + 16 * kt + 4 * q);  // [in][out]
                d = mfma(a, b, d);
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const uint32_t row = 4 * q + j, col = 16 * nt + r;
                const float mask = bf2f(t.h1[row * 64 + col]) > 0.f ? 1.f : 0.f;
                t.dz1[row * 64 + col] = f2bf(d[j] * mask);
            }
        }
        __syncthreads();
        // dW1[out][in] += dZ1^T X ; db1
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            v4s a = gather4(t.dz1 + (4 * q) * 64 + 16 * mt + r, 64);
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) {
                v4s b = gather4(t.x + (4 * q) * 64 + 16 * nt + r, 64);
                g1[mt][nt] = mfma(a, b, g1[mt][nt]);
            }
        }
        {
            float s = 0.f;
            for (int row = 0; row < 16; ++row) s += bf2f(t.dz1[row * 64 + l]);
            bias_grad[l] += s;
        }
        __syncthreads();
    }
    // write this workgroup's partial gradients: [W1 64x64][b1][W2][b2][W3 16x64][b3][loss]
    float *out = partial + (size_t)blockIdx.x * kMlpPartialStride;
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int nt = 0; nt < 4; ++nt)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const uint32_t o = 16 * mt + 4 * q + j, i = 16 * nt + r;
                out[kOffW1 + o * 64 + i] = g1[mt][nt][j];
                out[kOffW2 + o * 64 + i] = g2[mt][nt][j];
            }
#pragma unroll
    for (int nt = 0; nt < 4; ++nt)
#pragma unroll
        for (int j = 0; j < 4; ++j) out[kOffW3 + (4 * q + j) * 64 + 16 * nt + r] = g3[nt][j];
    __syncthreads();
    out[kOffB1 + l] = bias_grad[l];
    out[kOffB2 + l] = bias_grad[64 + l];
    if (l < 16) out[kOffB3 + l] = bias_grad[128 + l];
    // loss: sum over the wave
    for (int offset = 32; offset > 0; offset >>= 1) loss += __shfl_xor(loss, offset, 64);
    if (l == 0) out[kOffLoss] = loss * inv_rows;
}

// Sums the partial slots in block order (deterministic).
__global__ void k_mlp_reduce(const float *partial, uint32_t slots, float *grad) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kMlpPartialStride) return;
    float s = 0.f;
    for (uint32_t k = 0; k < slots; ++k) s += partial[(size_t)k * kMlpPartialStride + i];
    grad[i] = s;
}

// Adam on the f32 masters, then refreshes both bf16 copies of each weight.
__global__ void k_mlp_adam(MlpDevice m, const float *grad, float lr, float beta1, float beta2, float eps,
                           float correction1, float correction2) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kMlpParams) return;
    const float g = grad[i];
    float mo = m.adam_m[i] = beta1 * m.adam_m[i] + (1.f - beta1) * g;
    float ve = m.adam_v[i] = beta2 * m.adam_v[i] + (1.f - beta2) * g * g;
    const float mhat = mo / correction1, vhat = ve / correction2;
    const float w = m.master[i] - lr * mhat / (sqrtf(vhat) + eps);
    m.master[i] = w;
}

// bf16 copies of the masters: [out][in] for the forward, [in][out] for the
// backward data pass; biases are used in f32 directly from the masters.
__global__ void k_mlp_publish(MlpDevice m) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 64 * 64) {
        const uint32_t o = i / 64, k = i % 64;
        m.w1[i] = f2bf(m.master[kOffW1 + i]);
        m.w2[i] = f2bf(m.master[kOffW2 + i]);
        m.w2n[k * 64 + o] = f2bf(m.master[kOffW2 + i]);
    }
    if (i < 16 * 64) {
        const uint32_t o = i / 64, k = i % 64;
        m.w3[i] = f2bf(m.master[kOffW3 + i]);
        m.w3n[k * 16 + o] = f2bf(m.master[kOffW3 + i]);
    }
}

// Proposal: each vertex moves by the mean of its segments' predicted
// displacements (prediction times the segment's unit), scaled, clamped to
// walls.
__global__ void k_apply_proposal(const GridParams *grids, float4 *pos, const float *seg_scale, float4 *backup,
                                 const float4 *Y, const float4 *vA, const float4 *vB, float scale, uint32_t V) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    float4 p = pos[v];
    backup[v] = p;
    const float4 sb = vB[v];
    const uint32_t flags = __float_as_uint(sb.z);
    if (flags & (kPinned | kInactive)) return;
    const GridParams &g = grids[flags >> kSpecShift];
    const uint32_t next_seg = v - __float_as_uint(sb.w);
    float dx = 0.f, dy = 0.f, dz = 0.f, n = 0.f;
    if (flags & kHasPrev) {
        const float4 y = Y[next_seg - 1];
        const float u = seg_scale[next_seg - 1];
        dx += y.x * u; dy += y.y * u; dz += y.z * u; n += 1.f;
    }
    if (flags & kHasNext) {
        const float4 y = Y[next_seg];
        const float u = seg_scale[next_seg];
        dx += y.x * u; dy += y.y * u; dz += y.z * u; n += 1.f;
    }
    if (n == 0.f) return;
    const float k = scale / n;
    float x = p.x + dx * k, y = p.y + dy * k, z = p.z + dz * k;
    const float e = vA[v].w;
    if (!g.periodic[0]) x = fminf(fmaxf(x, g.lower[0] + e), g.upper[0] - e);
    if (!g.periodic[1]) y = fminf(fmaxf(y, g.lower[1] + e), g.upper[1] - e);
    if (!g.periodic[2]) z = fminf(fmaxf(z, g.lower[2] + e), g.upper[2] - e);
    pos[v] = make_float4(x, y, z, p.w);
}

__global__ void k_restore(float4 *pos, const float4 *backup, uint32_t V) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < V) pos[v] = backup[v];
}

// Restores the vertices of the specimens flagged in `revert` ([nspec]).
__global__ void k_restore_specimens(float4 *pos, const float4 *backup, const float4 *vB, const uint32_t *revert,
                                    uint32_t nspec, uint32_t V) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const uint32_t spec = __float_as_uint(vB[v].z) >> kSpecShift;
    if (spec < nspec && revert[spec] != 0u) pos[v] = backup[v];
}

// ---------------------------------------------------------------------------
// launchers

namespace {
inline uint32_t nblocks(uint64_t n, uint32_t t) { return (uint32_t)((n + t - 1) / t); }
}  // namespace

hipError_t launch_mfma_probe(const uint16_t *A, const uint16_t *B, float *D, hipStream_t s) {
    hipLaunchKernelGGL(k_mfma_probe, dim3(1), dim3(64), 0, s, A, B, D);
    return hipGetLastError();
}

hipError_t launch_features(const WorldDev &w, const float4 *history, uint16_t *X, float *scale, hipStream_t s) {
    hipError_t e = hipMemsetAsync(X, 0, (size_t)w.S * kFeatureWidth * sizeof(uint16_t), s);
    if (e != hipSuccess) return e;
    if (scale) e = hipMemsetAsync(scale, 0, (size_t)w.S * sizeof(float), s);
    if (e != hipSuccess) return e;
    e = launch_refresh_slots(w, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_features, dim3(nblocks(w.S, 256)), dim3(256), 0, s, w.grids, w.cP, w.cQ, w.slotC,
                       w.ncount, w.nbr, w.nbr_off, w.ovf, history, X, scale, w.block, w.ovf_cap, w.S);
    return hipGetLastError();
}

hipError_t launch_midpoints(const WorldDev &w, const float *scale, float4 *mid, hipStream_t s) {
    hipLaunchKernelGGL(k_midpoints, dim3(nblocks(w.S, 256)), dim3(256), 0, s, w.pos, w.sstat, scale, mid, w.S);
    return hipGetLastError();
}

hipError_t launch_targets(const WorldDev &w, const float4 *mid_then, float4 *T, hipStream_t s) {
    hipLaunchKernelGGL(k_targets, dim3(nblocks(w.S, 256)), dim3(256), 0, s, w.pos, w.sstat, mid_then, T, w.S);
    return hipGetLastError();
}

hipError_t launch_mlp_forward(const MlpDevice &m, const uint16_t *X, float4 *Y, uint32_t rows, hipStream_t s) {
    // persistent waves: 2 workgroups of 4 waves per CU, each wave looping over tiles
    const uint32_t tiles = nblocks(rows, 16);
    const uint32_t blocks = std::min<uint32_t>(nblocks(tiles, 4), 512u);
    hipLaunchKernelGGL(k_mlp_forward, dim3(blocks), dim3(256), 0, s, m, X, Y, rows);
    return hipGetLastError();
}

hipError_t launch_mlp_gradients(const MlpDevice &m, const uint16_t *X, const float4 *T, const uint32_t *idx,
                                uint32_t rows, uint32_t slots, float *partial, float *grad, hipStream_t s) {
    hipLaunchKernelGGL(k_mlp_backward, dim3(slots), dim3(64), 0, s, m, X, T, idx, rows, 1.0f / (float)rows, partial);
    hipLaunchKernelGGL(k_mlp_reduce, dim3(nblocks(kMlpPartialStride, 256)), dim3(256), 0, s, partial, slots, grad);
    return hipGetLastError();
}

hipError_t launch_mlp_adam(const MlpDevice &m, const float *grad, float lr, uint64_t step, hipStream_t s) {
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
    hipLaunchKernelGGL(k_mlp_adam, dim3(nblocks(kMlpParams, 256)), dim3(256), 0, s, m, grad, lr, b1, b2, 1e-8f, c1,
                       c2);
    hipLaunchKernelGGL(k_mlp_publish, dim3(nblocks(64 * 64, 256)), dim3(256), 0, s, m);
    return hipGetLastError();
}

hipError_t launch_mlp_publish(const MlpDevice &m, hipStream_t s) {
    hipLaunchKernelGGL(k_mlp_publish, dim3(nblocks(64 * 64, 256)), dim3(256), 0, s, m);
    return hipGetLastError();
}

hipError_t launch_apply_proposal(const WorldDev &w, const float4 *Y, const float *seg_scale, float4 *backup,
                                 float scale, hipStream_t s) {
    hipLaunchKernelGGL(k_apply_proposal, dim3(nblocks(w.V, 256)), dim3(256), 0, s, w.grids, w.pos, seg_scale, backup, Y,
                       w.vA, w.vB, scale, w.V);
    return hipGetLastError();
}

hipError_t launch_restore(const WorldDev &w, const float4 *backup, hipStream_t s) {
    hipLaunchKernelGGL(k_restore, dim3(nblocks(w.V, 256)), dim3(256), 0, s, w.pos, backup, w.V);
    return hipGetLastError();
}

hipError_t launch_restore_specimens(const WorldDev &w, const float4 *backup, const uint32_t *revert,
                                    hipStream_t s) {
    hipLaunchKernelGGL(k_restore_specimens, dim3(nblocks(w.V, 256)), dim3(256), 0, s, w.pos, backup, w.vB, revert,
                       w.nspec, w.V);
    return hipGetLastError();
}

}  // namespace tgh
