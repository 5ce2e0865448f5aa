// fe_learn.hip — the learned warm start of the two Jacobi solves (no reference counterpart).
//
// A per-particle MLP, 12 -> 64 -> 64 -> 64 -> 1 with GELU, predicts a particle's final
// pressure from 12 features of its neighbourhood. The prediction replaces the initial
// pressure of the density solve (zero in the reference, density_source_term.comp:73) or of
// the divergence solve (the density solve's pressures in the reference), and plain Jacobi
// then runs to the same error thresholds, so the solver still decides when a result is good
// enough.
//
// Arithmetic: BF16 weights and activations, FP32 accumulation, on the matrix cores
// (v_mfma_f32_16x16x32_bf16, 16 rows x 16 columns x 32 k per instruction). Operand lane
// layouts of that instruction: A[r][k] comes from lane r + 16 (k / 8), element k % 8;
// B[k][c] from lane c + 16 (k / 8), element k % 8; D[r][c] lands in lane c + 16 (r / 4),
// element r % 4. Every product below is a sum of such tiles: the forward products X W1,
// H1 W2, H2 W3, the backward products dZ W^T and the weight gradients H^T dZ (k = the 32
// samples of a wave). The 64 -> 1 output layer is a dot product on the VALU.
//
// Training: one wave takes 32 samples through forward and backward and writes its partial
// gradient; a second kernel sums the partials in a fixed order and applies Adam, so a
// training run is deterministic.
#include "fe_device.h"

namespace fe {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int H = kMlpHidden;   // 64
constexpr int KIN = 32;         // input k padded to one MFMA k-chunk

// Parameter vector (FP32): W1[12][64], b1[64], W2[64][64], b2, W3[64][64], b3, w4[64], b4,
// each W[in][out] row-major.
struct Off {
    static constexpr int W1 = 0, B1 = W1 + kMlpIn * H, W2 = B1 + H, B2 = W2 + H * H, W3 = B2 + H, B3 = W3 + H * H,
                         W4 = B3 + H, B4 = W4 + H, N = B4 + 1;
};
static_assert(Off::N == kMlpParams, "parameter layout");

__device__ __forceinline__ f32x4 mfma(const bf16x8& a, const bf16x8& b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// GELU(z) = z Phi(z), Phi(z) = (1 + erf(z / sqrt 2)) / 2, with erf from Abramowitz & Stegun
// 7.1.26 (Hastings' rational approximation, |error| <= 1.5e-7) on the hardware exponential
// and reciprocal: one exp2 and one rcp per activation. e = exp(-z^2 / 2) is returned for the
// derivative GELU'(z) = Phi(z) + z e / sqrt(2 pi).
__device__ __forceinline__ float phi_of(float z, float& e) {
    const float x = fabsf(z) * 0.70710678118654752f;
    const float t = __builtin_amdgcn_rcpf(fmaf(0.3275911f, x, 1.0f));
    float p = fmaf(1.061405429f, t, -1.453152027f);
    p = fmaf(p, t, 1.421413741f);
    p = fmaf(p, t, -0.284496736f);
    p = fmaf(p, t, 0.254829592f);
    p *= t;
    e = __builtin_amdgcn_exp2f(z * z * -0.72134752044448170f);
    const float half_erfc = 0.5f * p * e;   // (1 - erf(|z| / sqrt 2)) / 2
    return z >= 0.0f ? 1.0f - half_erfc : half_erfc;
}
__device__ __forceinline__ float gelu(float z) {
    float e;
    return z * phi_of(z, e);
}
__device__ __forceinline__ float gelu_grad(float z) {
    float e;
    const float ph = phi_of(z, e);
    return fmaf(z * 0.39894228040143268f, e, ph);
}
__device__ __forceinline__ __bf16 tobf(float x) { return (__bf16)x; }

// Forward weights in LDS: BF16, stored [out][in] (k contiguous, the B operand layout).
struct FwdWeights {
    alignas(16) __bf16 w1t[H][KIN];
    alignas(16) __bf16 w2t[H][H];
    alignas(16) __bf16 w3t[H][H];
    float b1[H], b2[H], b3[H], w4[H];   // w4 holds BF16-rounded values
    float b4;
};
// Backward weights: W2, W3 stored [in][out] (the B operand of dH = dZ W^T).
struct BwdWeights {
    alignas(16) __bf16 w2[H][H];
    alignas(16) __bf16 w3[H][H];
};

__device__ void load_fwd(const float* __restrict__ p, FwdWeights& w) {
    for (int e = threadIdx.x; e < H * KIN; e += blockDim.x) {
        const int o = e / KIN, i = e % KIN;
        w.w1t[o][i] = tobf(i < kMlpIn ? p[Off::W1 + i * H + o] : 0.0f);
    }
    for (int e = threadIdx.x; e < H * H; e += blockDim.x) {
        const int i = e / H, o = e % H;
        w.w2t[o][i] = tobf(p[Off::W2 + e]);
        w.w3t[o][i] = tobf(p[Off::W3 + e]);
    }
    for (int e = threadIdx.x; e < H; e += blockDim.x) {
        w.b1[e] = p[Off::B1 + e];
        w.b2[e] = p[Off::B2 + e];
        w.b3[e] = p[Off::B3 + e];
        w.w4[e] = (float)tobf(p[Off::W4 + e]);
    }
    if (threadIdx.x == 0) w.b4 = p[Off::B4];
}

__device__ void load_bwd(const float* __restrict__ p, BwdWeights& w) {
    for (int e = threadIdx.x; e < H * H; e += blockDim.x) {
        const int i = e / H, o = e % H;
        w.w2[i][o] = tobf(p[Off::W2 + e]);
        w.w3[i][o] = tobf(p[Off::W3 + e]);
    }
}

// A operand (16 rows x 32 k) of a row-major BF16 matrix M[row][ld] at (row0, k0).
__device__ __forceinline__ bf16x8 a_op(const __bf16* m, int ld, int row0, int k0) {
    const int l = threadIdx.x & 63;
    return *reinterpret_cast<const bf16x8*>(m + (row0 + (l & 15)) * ld + k0 + 8 * (l >> 4));
}
// B operand (32 k x 16 columns) of a matrix stored column by column, M[col][ld], at (k0, col0).
__device__ __forceinline__ bf16x8 b_op(const __bf16* m, int ld, int k0, int col0) {
    const int l = threadIdx.x & 63;
    return *reinterpret_cast<const bf16x8*>(m + (col0 + (l & 15)) * ld + k0 + 8 * (l >> 4));
}

// Orders this wave's LDS accesses around it (no cross-wave barrier).
__device__ __forceinline__ void wave_lds_fence() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup"); }

// Pre-activations of a 16-row tile, kept in the accumulator layout: z[layer][column tile].
struct ZTile {
    f32x4 z[3][4];
};

// Forward pass of one 16-row tile by one wave. x: BF16 [16][KIN]; o0, o1, o2: BF16 [16][H]
// activations of the three hidden layers (o2 may alias o0).
template <bool keep_z>
__device__ __forceinline__ void forward16(const FwdWeights& w, const __bf16* x, __bf16* o0, __bf16* o1, __bf16* o2,
                                          ZTile* zt) {
    const int l = threadIdx.x & 63;
    const int col = l & 15, rbase = 4 * (l >> 4);
    const __bf16* in = x;
#pragma unroll
    for (int layer = 0; layer < 3; layer++) {
        const __bf16* wt = layer == 0 ? &w.w1t[0][0] : layer == 1 ? &w.w2t[0][0] : &w.w3t[0][0];
        const float* bias = layer == 0 ? w.b1 : layer == 1 ? w.b2 : w.b3;
        __bf16* out = layer == 0 ? o0 : layer == 1 ? o1 : o2;
        const int kin = layer == 0 ? KIN : H;
        f32x4 acc[4];
#pragma unroll
        for (int ct = 0; ct < 4; ct++) {
            acc[ct] = f32x4{0.f, 0.f, 0.f, 0.f};
            for (int k0 = 0; k0 < kin; k0 += 32) acc[ct] = mfma(a_op(in, kin, 0, k0), b_op(wt, kin, k0, ct * 16), acc[ct]);
        }
        // every read of `in` precedes the writes of `out`, which may alias an older layer
        wave_lds_fence();
#pragma unroll
        for (int ct = 0; ct < 4; ct++) {
            const float b = bias[ct * 16 + col];
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const float z = acc[ct][e] + b;
                if (keep_z) zt->z[layer][ct][e] = z;
                out[(rbase + e) * H + ct * 16 + col] = tobf(gelu(z));
            }
        }
        wave_lds_fence();
        in = out;
    }
}

// ---- inference -------------------------------------------------------------------------
//
// Transposed products, Z^T = W^T H^T: the weights are the A operand (16 output neurons x
// 32 inputs per tile) and stay in registers for the whole launch; the activations of 16
// samples are the B operand (32 inputs x 16 samples). The accumulator then gives lane
// c + 16 q the neurons 4q .. 4q + 3 of sample c, which are contiguous in a [sample][neuron]
// activation row (one 8-byte LDS store per tile), and the layer-1 operand (features 8q ..
// 8q + 7 of sample c) is built in registers from the feature-major input.

// Packed inference copy, in floats: kMlpPackedOps x 64 lanes x 8 BF16 weight operands, then
// b1, b2, b3, the BF16-rounded w4, b4, the feature means and reciprocal standard deviations,
// the target mean and standard deviation.
constexpr int PK_B1 = kMlpPackedOps * 64 * 4, PK_B2 = PK_B1 + H, PK_B3 = PK_B2 + H, PK_W4 = PK_B3 + H,
              PK_B4 = PK_W4 + H, PK_MEAN = PK_B4 + 1, PK_ISTD = PK_MEAN + kMlpIn, PK_TMEAN = PK_ISTD + kMlpIn,
              PK_TSTD = PK_TMEAN + 1, PK_END = PK_TSTD + 1;
static_assert(PK_END <= kMlpPackedFloats, "packed layout");

// GELU of the inference kernel: piecewise linear through GELU at z = -8 + k / 128,
// k = 0 .. 2048, with the end segments extended linearly (GELU(z) is 0 below -8 and z above 8
// to within 1e-14). Inference works in the coordinate U = 1024 z + 8192 = 8 u, u = (z + 8) 128
// the segment coordinate: the packed weights are scaled by 1024 (exact in BF16) and the packed
// biases are 1024 b + 8192, and the bias enters as the products' initial accumulator, so each
// layer's accumulators are U. Entry k holds segment k as (intercept, slope / 8):
// GELU ~ intercept + (slope / 8) U on u in [k, k + 1], and its byte offset in the table is
// floor(U) rounded down to a multiple of 8.
// Interpolation error <= (1/128)^2 / 8 * max|GELU''| < 7e-6, far below a BF16 step of the
// activations it feeds; the table is built from gelu() above.
constexpr int kGeluNodes = 2048;
constexpr float kGeluLo = -8.0f, kGeluScale = 128.0f;
constexpr float kUScale = 8.0f * kGeluScale, kUBias = -8.0f * kGeluLo * kGeluScale;   // U = 1024 z + 8192

__global__ __launch_bounds__(256) void k_gelu_table(float2* __restrict__ table) {
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < kGeluNodes; k += gridDim.x * blockDim.x) {
        const double z0 = (double)kGeluLo + (double)k / kGeluScale, z1 = z0 + 1.0 / kGeluScale;
        const float g0 = gelu((float)z0), g1 = gelu((float)z1);
        const float rise = g1 - g0;
        table[k] = make_float2((float)((double)g0 - (double)rise * k), (float)((double)rise / 8.0));
    }
}

static_assert(PK_END + 2 * kGeluNodes <= kMlpPackedFloats, "packed layout with the GELU table");

// GELU at the table coordinate U (see above): the segment's byte offset is the truncation of
// U clamped to the table and rounded down to 8 bytes (v_cvt_i32_f32: NaN -> 0, saturating).
__device__ __forceinline__ float gelu_lut(const float2* __restrict__ lut, float U) {
    int b = (int)U;
    b = b < 0 ? 0 : (b > 8 * kGeluNodes - 1 ? 8 * kGeluNodes - 1 : b);
    const float2 e = *reinterpret_cast<const float2*>(reinterpret_cast<const char*>(lut) + (b & ~7));
    return fmaf(e.y, U, e.x);
}

// Operand g: layer 1 tiles g = 0..3 (column tile ct = g), layer 2 g = 4 + 2 ct + kc, layer 3
// g = 12 + 2 ct + kc. Lane l holds W[kc 32 + 8 (l / 16) + e][ct 16 + l % 16], e = 0..7.
__global__ __launch_bounds__(256) void k_mlp_pack(const float* __restrict__ params, const float* __restrict__ norm,
                                                  float* __restrict__ packed) {
    __bf16* ops = reinterpret_cast<__bf16*>(packed);
    for (int e = threadIdx.x; e < kMlpPackedOps * 64 * 8; e += blockDim.x) {
        const int g = e / 512, l = (e / 8) % 64, el = e % 8;
        int layer, ct, kc;
        if (g < 4) {
            layer = 0; ct = g; kc = 0;
        } else {
            layer = g < 12 ? 1 : 2;
            ct = ((g - 4) % 8) / 2;
            kc = (g - 4) % 2;
        }
        const int in = kc * 32 + 8 * (l >> 4) + el, out = ct * 16 + (l & 15);
        float v = 0.0f;
        if (layer == 0) v = in < kMlpIn ? params[Off::W1 + in * H + out] : 0.0f;
        else v = params[(layer == 1 ? Off::W2 : Off::W3) + in * H + out];
        ops[e] = tobf(v * kUScale);   // = 1024 bf16(v): a power of two commutes with the rounding
    }
    for (int e = threadIdx.x; e < H; e += blockDim.x) {
        packed[PK_B1 + e] = fmaf(params[Off::B1 + e], kUScale, kUBias);
        packed[PK_B2 + e] = fmaf(params[Off::B2 + e], kUScale, kUBias);
        packed[PK_B3 + e] = fmaf(params[Off::B3 + e], kUScale, kUBias);
        packed[PK_W4 + e] = (float)tobf(params[Off::W4 + e]);
    }
    for (int e = threadIdx.x; e < kMlpIn; e += blockDim.x) {
        packed[PK_MEAN + e] = norm[e];
        packed[PK_ISTD + e] = 1.0f / norm[kMlpIn + e];
    }
    if (threadIdx.x == 0) {
        packed[PK_B4] = params[Off::B4];
        packed[PK_TMEAN] = norm[2 * kMlpIn];
        packed[PK_TSTD] = norm[2 * kMlpIn + 1];
    }
}

constexpr int kInferWaves = 8;

__device__ __forceinline__ bf16x8 as_bf16x8(uint4 u) { return __builtin_bit_cast(bf16x8, u); }

// The 8 inputs lane (c, q) feeds for sample s: features 8q .. 8q + 7 (q < 2), zero otherwise.
__device__ __forceinline__ void load_inputs(const float* __restrict__ feat, uint32_t n, uint32_t s, int q,
                                            float (&x)[8]) {
#pragma unroll
    for (int e = 0; e < 8; e++) {
        const int k = 8 * q + e;
        x[e] = (q < 2 && k < kMlpIn && s < n) ? feat[(size_t)k * n + s] : 0.0f;
    }
}

// The accumulators' initial values: this lane's four neurons' packed biases (U coordinate).
__device__ __forceinline__ f32x4 bias4(const float* sb, int ct, int q) {
    const float4 b = *reinterpret_cast<const float4*>(sb + ct * 16 + 4 * q);
    return f32x4{b.x, b.y, b.z, b.w};
}

// One layer's GELU on the accumulators (U) of the four column tiles, written as BF16
// activation rows act[sample][neuron].
__device__ __forceinline__ void store_act(const float2* lut, const f32x4 (&acc)[4], __bf16* act, int c, int q) {
#pragma unroll
    for (int ct = 0; ct < 4; ct++) {
        const __bf16 h0 = tobf(gelu_lut(lut, acc[ct][0])), h1 = tobf(gelu_lut(lut, acc[ct][1]));
        const __bf16 h2 = tobf(gelu_lut(lut, acc[ct][2])), h3 = tobf(gelu_lut(lut, acc[ct][3]));
        typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
        const bf16x4 v = {h0, h1, h2, h3};
        *reinterpret_cast<bf16x4*>(act + c * H + ct * 16 + 4 * q) = v;
    }
}

// Per sample s < n (s = schedule position): y = MLP(standardised features). With raw != 0,
// y goes to yraw[s] and the inputs are used as given; otherwise p = max(0, y target_std +
// target_mean) goes to pressures[sched[s]], to its mirror pres_m[s], and p / rho^2 to the
// w of the position mirror pos_m[s].
__global__ __launch_bounds__(64 * kInferWaves) void k_mlp_infer(const float* __restrict__ packed,
                                                   const float2* __restrict__ gelu_table,
                                                   const float* __restrict__ feat, uint32_t n, uint32_t raw,
                                                   const uint32_t* __restrict__ sched, const float* __restrict__ dens,
                                                   float* __restrict__ pres, float4* __restrict__ pos_m,
                                                   float* __restrict__ pres_m, float* __restrict__ yraw) {
    __shared__ alignas(16) float sf[PK_END - PK_B1];
    __shared__ alignas(16) __bf16 act[kInferWaves][16 * H];
    __shared__ float2 lut[kGeluNodes];
    // the layer 2 and 3 operands (groups 4 .. 19), one 16-byte row per lane: in LDS rather
    // than in 64 VGPRs per lane, which would allow only two waves per SIMD
    __shared__ uint4 sops[16 * 64];
    for (int e = threadIdx.x; e < PK_END - PK_B1; e += blockDim.x) sf[e] = packed[PK_B1 + e];
    for (int e = threadIdx.x; e < kGeluNodes; e += blockDim.x) lut[e] = gelu_table[e];
    const uint4* ops = reinterpret_cast<const uint4*>(packed);
    for (int e = threadIdx.x; e < 16 * 64; e += blockDim.x) sops[e] = ops[4 * 64 + e];
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int c = l & 15, q = l >> 4;
    bf16x8 A1[4];
#pragma unroll
    for (int g = 0; g < 4; g++) A1[g] = as_bf16x8(ops[g * 64 + l]);
    __syncthreads();
    float mu[8], is[8];
#pragma unroll
    for (int e = 0; e < 8; e++) {
        const int k = 8 * q + e;
        const bool ok = q < 2 && k < kMlpIn;
        mu[e] = ok && !raw ? sf[PK_MEAN - PK_B1 + k] : 0.0f;
        is[e] = ok ? (raw ? 1.0f : sf[PK_ISTD - PK_B1 + k]) : 0.0f;
    }
    const float b4 = sf[PK_B4 - PK_B1], tmean = sf[PK_TMEAN - PK_B1], tstd = sf[PK_TSTD - PK_B1];
    __bf16* a = act[w];
    const uint32_t tiles = (n + 15) / 16, stride = gridDim.x * kInferWaves;
    uint32_t tile = blockIdx.x * kInferWaves + w;
    float nx[8];
    load_inputs(feat, n, tile * 16 + c, q, nx);
    for (; tile < tiles; tile += stride) {
        float x[8];
#pragma unroll
        for (int e = 0; e < 8; e++) x[e] = nx[e];
        if (tile + stride < tiles) load_inputs(feat, n, (tile + stride) * 16 + c, q, nx);   // next tile in flight
        bf16x8 B;
#pragma unroll
        for (int e = 0; e < 8; e++) B[e] = tobf((x[e] - mu[e]) * is[e]);
        f32x4 acc[4];
#pragma unroll
        for (int ct = 0; ct < 4; ct++) acc[ct] = mfma(A1[ct], B, bias4(sf + (PK_B1 - PK_B1), ct, q));
        store_act(lut, acc, a, c, q);
        wave_lds_fence();
#pragma unroll
        for (int layer = 0; layer < 2; layer++) {
            const bf16x8 B0 = *reinterpret_cast<const bf16x8*>(a + c * H + 8 * q);
            const bf16x8 B1 = *reinterpret_cast<const bf16x8*>(a + c * H + 32 + 8 * q);
            wave_lds_fence();   // both operands read before the row is overwritten
#pragma unroll
            for (int ct = 0; ct < 4; ct++) {
                const bf16x8 a0 = as_bf16x8(sops[(8 * layer + 2 * ct) * 64 + l]);
                const bf16x8 a1 = as_bf16x8(sops[(8 * layer + 2 * ct + 1) * 64 + l]);
                acc[ct] = mfma(a1, B1, mfma(a0, B0, bias4(sf + (layer == 0 ? PK_B2 : PK_B3) - PK_B1, ct, q)));
            }
            if (layer == 0) {
                store_act(lut, acc, a, c, q);
                wave_lds_fence();
            }
        }
        // output layer from the layer-3 accumulators: this lane's 16 neurons, then the four
        // lanes of the sample
        float y = 0.0f;
#pragma unroll
        for (int ct = 0; ct < 4; ct++) {
            const float4 w4 = *reinterpret_cast<const float4*>(sf + (PK_W4 - PK_B1) + ct * 16 + 4 * q);
            y += (float)tobf(gelu_lut(lut, acc[ct][0])) * w4.x;
            y += (float)tobf(gelu_lut(lut, acc[ct][1])) * w4.y;
            y += (float)tobf(gelu_lut(lut, acc[ct][2])) * w4.z;
            y += (float)tobf(gelu_lut(lut, acc[ct][3])) * w4.w;
        }
        y += __shfl_xor(y, 16, 64);
        y += __shfl_xor(y, 32, 64);
        y += b4;
        const uint32_t s = tile * 16 + c;
        if (q == 0 && s < n) {
            if (raw) {
                yraw[s] = y;
            } else {
                const uint32_t i = sched[s];
                float p = y * tstd + tmean;
                p = p > 0.0f ? p : 0.0f;
                pres[i] = p;
                pres_m[s] = p;
                pos_m[s].w = prho_of(p, dens[i]);
            }
        }
    }
}

// ---- training ----------------------------------------------------------------------------

struct TrainLds {
    FwdWeights w;
    BwdWeights wb;
    alignas(16) __bf16 x[2][16][KIN];   // inputs, two 16-row tiles (A of layer 1)
    alignas(16) __bf16 xt[KIN][32];     // inputs transposed [k][sample] (A of dW1 = X^T dZ1)
    alignas(16) __bf16 h[3][2][16][H];  // activations of the three hidden layers
    alignas(16) __bf16 ht[H][32];       // activations of the layer below, transposed (A of dW)
    alignas(16) __bf16 dz[2][16][H];    // dZ, row-major (A of dH = dZ W^T)
    alignas(16) __bf16 dzt[H][32];      // dZ transposed [neuron][sample] (B of dW = H^T dZ)
    float dy[32];
};

__device__ __forceinline__ uint32_t sample_row(uint32_t b, uint32_t mod, uint32_t seed) {
    if (mod == 0) return b;
    uint32_t h = b * 0x9E3779B1u + seed * 0x85EBCA77u + 0x165667B1u;
    h ^= h >> 15;
    h *= 0x2C1B3C6Du;
    h ^= h >> 12;
    h *= 0x297A2D39u;
    h ^= h >> 15;
    return h % mod;
}

// Loss L = sum_b (y_b - t_b)^2 / batch. One wave per block takes samples
// [32 blockIdx, 32 blockIdx + 32): forward, output error, backward. Writes the block's partial
// gradient to grads[blockIdx][kMlpParams] and its summed squared error to loss[blockIdx].
__global__ __launch_bounds__(64) void k_mlp_train(const float* __restrict__ params, const float* __restrict__ x,
                                                  const float* __restrict__ tgt, uint32_t batch, uint32_t sample_mod,
                                                  uint32_t seed, float inv_batch, float* __restrict__ grads,
                                                  float* __restrict__ loss) {
    __shared__ TrainLds s;
    load_fwd(params, s.w);
    load_bwd(params, s.wb);
    const int l = threadIdx.x;
    const uint32_t s0 = blockIdx.x * kMlpSamplesPerBlock;
    for (int e = l; e < 32 * KIN; e += 64) {
        const int r = e / KIN, k = e % KIN;
        const uint32_t b = s0 + r;
        float v = 0.0f;
        if (k < kMlpIn && b < batch) v = x[(size_t)sample_row(b, sample_mod, seed) * kMlpIn + k];
        const __bf16 bv = tobf(v);
        s.x[r >> 4][r & 15][k] = bv;
        s.xt[k][r] = bv;
    }
    __syncthreads();
    const int col = l & 15, rbase = 4 * (l >> 4);
    ZTile zt[2];
    forward16<true>(s.w, &s.x[0][0][0], &s.h[0][0][0][0], &s.h[1][0][0][0], &s.h[2][0][0][0], &zt[0]);
    forward16<true>(s.w, &s.x[1][0][0], &s.h[0][1][0][0], &s.h[1][1][0][0], &s.h[2][1][0][0], &zt[1]);
    __syncthreads();
    // output layer and its error: lane b < 32 owns sample b
    float lossp = 0.0f;
    {
        const int b = l & 31;
        const __bf16* h3 = &s.h[2][b >> 4][b & 15][0];
        float y = 0.0f;
        for (int k = 0; k < H; k++) y += (float)h3[k] * s.w.w4[k];
        y += s.w.b4;
        float d = 0.0f;
        if (l < 32 && s0 + l < batch) {
            const float e = y - tgt[sample_row(s0 + l, sample_mod, seed)];
            lossp = e * e;
            d = 2.0f * e * inv_batch;
        }
        if (l < 32) s.dy[l] = d;
    }
    for (int off = 32; off > 0; off >>= 1) lossp += __shfl_xor(lossp, off, 64);
    __syncthreads();
    float* g = grads + (size_t)blockIdx.x * kMlpParams;
    if (l == 0) loss[blockIdx.x] = lossp;
    {
        // dw4 = H3^T dy, db4 = sum dy
        float acc = 0.0f;
        for (int b = 0; b < 32; b++) acc += (float)s.h[2][b >> 4][b & 15][l] * s.dy[b];
        g[Off::W4 + l] = acc;
        if (l == 0) {
            float sb = 0.0f;
            for (int b = 0; b < 32; b++) sb += s.dy[b];
            g[Off::B4] = sb;
        }
    }
    // dZ3 = (dy w4) * gelu'(Z3), in the accumulator layout
    f32x4 dz[2][4];
#pragma unroll
    for (int rt = 0; rt < 2; rt++)
#pragma unroll
        for (int ct = 0; ct < 4; ct++)
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int r = rt * 16 + rbase + e, c = ct * 16 + col;
                dz[rt][ct][e] = s.dy[r] * s.w.w4[c] * gelu_grad(zt[rt].z[2][ct][e]);
            }
#pragma unroll
    for (int layer = 2; layer >= 0; layer--) {
        __syncthreads();
        // dZ (BF16) into LDS in both layouts
#pragma unroll
        for (int rt = 0; rt < 2; rt++)
#pragma unroll
            for (int ct = 0; ct < 4; ct++)
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    const int r = rbase + e, c = ct * 16 + col;
                    const __bf16 v = tobf(dz[rt][ct][e]);
                    s.dz[rt][r][c] = v;
                    s.dzt[c][rt * 16 + r] = v;
                }
        // bias gradient = column sums of dZ (FP32)
        const int boff = layer == 0 ? Off::B1 : layer == 1 ? Off::B2 : Off::B3;
#pragma unroll
        for (int ct = 0; ct < 4; ct++) {
            float sb = 0.0f;
#pragma unroll
            for (int rt = 0; rt < 2; rt++)
#pragma unroll
                for (int e = 0; e < 4; e++) sb += dz[rt][ct][e];
            sb += __shfl_xor(sb, 16, 64);
            sb += __shfl_xor(sb, 32, 64);
            if (l < 16) g[boff + ct * 16 + col] = sb;
        }
        // activations of the layer below, transposed
        if (layer > 0)
            for (int e = l; e < 32 * H; e += 64) {
                const int b = e / H, k = e % H;
                s.ht[k][b] = s.h[layer - 1][b >> 4][b & 15][k];
            }
        __syncthreads();
        // dW[in][out] = A^T dZ over the 32 samples (one MFMA per 16 x 16 tile)
        const int nin = layer == 0 ? kMlpIn : H;
        const int woff = layer == 0 ? Off::W1 : layer == 1 ? Off::W2 : Off::W3;
        const __bf16* at = layer == 0 ? &s.xt[0][0] : &s.ht[0][0];
        for (int it = 0; it * 16 < nin; it++)
            for (int ct = 0; ct < 4; ct++) {
                f32x4 acc = {0.f, 0.f, 0.f, 0.f};
                acc = mfma(a_op(at, 32, it * 16, 0), b_op(&s.dzt[0][0], 32, 0, ct * 16), acc);
                for (int e = 0; e < 4; e++) {
                    const int i = it * 16 + rbase + e, o = ct * 16 + col;
                    if (i < nin) g[woff + i * H + o] = acc[e];
                }
            }
        if (layer == 0) break;
        // dH(below) = dZ W^T, dZ(below) = dH * gelu'(Z(below))
        const __bf16* wback = layer == 1 ? &s.wb.w2[0][0] : &s.wb.w3[0][0];
#pragma unroll
        for (int rt = 0; rt < 2; rt++)
#pragma unroll
            for (int ct = 0; ct < 4; ct++) {
                f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int k0 = 0; k0 < H; k0 += 32)
                    acc = mfma(a_op(&s.dz[rt][0][0], H, 0, k0), b_op(wback, H, k0, ct * 16), acc);
#pragma unroll
                for (int e = 0; e < 4; e++) dz[rt][ct][e] = acc[e] * gelu_grad(zt[rt].z[layer - 1][ct][e]);
            }
    }
}

// Fixed-order sum of the per-block partials; Adam on the FP32 parameters when apply != 0.
__global__ __launch_bounds__(256) void k_mlp_adam(float* __restrict__ params, float* __restrict__ m1,
                                                  float* __restrict__ m2, const float* __restrict__ grads,
                                                  uint32_t parts, const float* __restrict__ loss,
                                                  float* __restrict__ grad_out, float* __restrict__ loss_out, float lr,
                                                  float bc1, float bc2, uint32_t apply) {
    const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < (uint32_t)kMlpParams; p += gridDim.x * blockDim.x) {
        float gsum = 0.0f;
        for (uint32_t k = 0; k < parts; k++) gsum += grads[(size_t)k * kMlpParams + p];
        if (grad_out) grad_out[p] = gsum;
        if (apply) {
            const float a = beta1 * m1[p] + (1.0f - beta1) * gsum;
            const float b = beta2 * m2[p] + (1.0f - beta2) * gsum * gsum;
            m1[p] = a;
            m2[p] = b;
            params[p] -= lr * (a / bc1) / (sqrtf(b / bc2) + eps);
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0 && loss_out) {
        float t = 0.0f;
        for (uint32_t k = 0; k < parts; k++) t += loss[k];
        *loss_out = t;
    }
}

// ---- previous pressures and samples ---------------------------------------------------------

// pprev[t] = plast[vals[sched[t]]]: the last solve's pressure of the particle at schedule
// position t (vals maps the current sort order to the order position_a had, which is the
// order of the previous solve of the same kind).
__global__ __launch_bounds__(kBlock) void k_prev(const uint32_t* __restrict__ sched, const uint32_t* __restrict__ vals,
                                                 const float* __restrict__ plast, uint32_t valid, uint32_t n,
                                                 float* __restrict__ pprev) {
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= n) return;
    const uint32_t v = vals[sched[t]];
    pprev[t] = (valid && v < n) ? plast[v] : 0.0f;   // v >= n: no search has run yet
}

// Append `count` samples (features, pressure) of particles picked by a hash of (salt, k) to
// a ring buffer of `cap` samples starting at `head` (rows of kMlpIn features).
__global__ __launch_bounds__(kBlock) void k_collect(const float* __restrict__ feat, const uint32_t* __restrict__ sched,
                                                    const float* __restrict__ pres, uint32_t n, uint32_t count,
                                                    uint32_t salt, uint32_t head, uint32_t cap, float* __restrict__ dx,
                                                    float* __restrict__ dt) {
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= count) return;
    const uint32_t t = sample_row(k, n, salt);
    const uint32_t slot = (head + k) % cap;
    for (int e = 0; e < kMlpIn; e++) dx[(size_t)slot * kMlpIn + e] = feat[(size_t)e * n + t];
    dt[slot] = pres[sched[t]];
}

inline dim3 grid_for(uint32_t count) { return dim3((count + kBlock - 1) / kBlock); }

// Persistent grid of the inference kernel: 4 waves per block, about 8 waves per CU.
inline uint32_t infer_blocks(uint32_t n) {
    const uint32_t tiles = (n + 15) / 16;
    const uint32_t blocks = (tiles + kInferWaves - 1) / kInferWaves;
    return blocks > 512 ? 512 : (blocks ? blocks : 1);
}

}  // namespace

hipError_t learn_prev(fe_ctx* c, int which) {
    Launch L(c, "ws_prev");
    hipLaunchKernelGGL(k_prev, grid_for(c->n), dim3(kBlock), 0, c->stream, c->sched[0], c->vals[0],
                       c->ws.plast[which], c->ws.plast_valid[which] ? 1u : 0u, c->n, c->ws.pprev);
    return hipGetLastError();
}

hipError_t learn_pack(fe_ctx* c, int which) {
    hipLaunchKernelGGL(k_mlp_pack, dim3(1), dim3(256), 0, c->stream, c->ws.params[which], c->ws.norm[which],
                       c->ws.packed[which]);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_gelu_table, dim3(kGeluNodes / 256), dim3(256), 0, c->stream,
                       reinterpret_cast<float2*>(c->ws.packed[which] + kMlpPackedFloats - 2 * kGeluNodes));
    return hipGetLastError();
}

inline const float2* gelu_table_of(const fe_ctx* c, int which) {
    return reinterpret_cast<const float2*>(c->ws.packed[which] + kMlpPackedFloats - 2 * kGeluNodes);
}

hipError_t learn_infer(fe_ctx* c, int which) {
    Launch L(c, "ws_infer");
    hipLaunchKernelGGL(k_mlp_infer, dim3(infer_blocks(c->n)), dim3(64 * kInferWaves), 0, c->stream,
                       c->ws.packed[which], gelu_table_of(c, which), c->ws.feat, c->n, 0u, c->sched[0], c->dens,
                       c->pres, c->pos_m, c->pres_m, (float*)nullptr);
    return hipGetLastError();
}

hipError_t learn_collect(fe_ctx* c, int which) {
    LearnState& w = c->ws;
    uint32_t count = w.per_solve < w.capacity ? w.per_solve : w.capacity;
    if (count > c->n) count = c->n;
    if (count == 0) return hipSuccess;
    Launch L(c, "ws_collect");
    hipLaunchKernelGGL(k_collect, grid_for(count), dim3(kBlock), 0, c->stream, w.feat, c->sched[0], c->pres, c->n,
                       count, w.salt++, w.head[which], w.capacity, w.data_x[which], w.data_t[which]);
    w.head[which] = (w.head[which] + count) % w.capacity;
    w.size[which] = w.size[which] + count < w.capacity ? w.size[which] + count : w.capacity;
    return hipGetLastError();
}

hipError_t learn_forward_raw(fe_ctx* c, int which, const float* x, uint32_t n, float* y) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_mlp_infer, dim3(infer_blocks(n)), dim3(64 * kInferWaves), 0, c->stream, c->ws.packed[which],
                       gelu_table_of(c, which), x, n, 1u, (const uint32_t*)nullptr, (const float*)nullptr,
                       (float*)nullptr, (float4*)nullptr, (float*)nullptr, y);
    return hipGetLastError();
}

hipError_t learn_step(fe_ctx* c, int which, const float* x, const float* t, uint32_t batch, uint32_t sample_mod,
                      uint32_t seed, float lr, uint32_t apply, float* grad_out, float* loss_out) {
    LearnState& w = c->ws;
    if (batch == 0 || batch > kMlpMaxBatch) return hipErrorInvalidValue;
    const uint32_t parts = (batch + kMlpSamplesPerBlock - 1) / kMlpSamplesPerBlock;
    hipLaunchKernelGGL(k_mlp_train, dim3(parts), dim3(64), 0, c->stream, w.params[which], x, t, batch, sample_mod,
                       seed, 1.0f / (float)batch, w.grads, w.loss);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    float bc1 = 1.0f, bc2 = 1.0f;
    if (apply) {
        w.adam_t[which]++;
        bc1 = 1.0f - powf(0.9f, (float)w.adam_t[which]);
        bc2 = 1.0f - powf(0.999f, (float)w.adam_t[which]);
    }
    hipLaunchKernelGGL(k_mlp_adam, dim3((kMlpParams + 255) / 256), dim3(256), 0, c->stream, w.params[which],
                       w.m1[which], w.m2[which], w.grads, parts, w.loss, grad_out, loss_out, lr, bc1, bc2, apply);
    return hipGetLastError();
}

}  // namespace fe
