// fe_learn.hip — the learned warm start of the two Jacobi solves (no reference counterpart).
//
// A per-particle MLP, 12 -> 64 -> 64 -> 64 -> 1 with GELU, predicts a particle's final
// pressure from 12 features of its neighbourhood. The prediction replaces the initial
// pressure of the density solve (zero in the reference, density_source_term.comp:73) or of
// the divergence solve (the density solve's pressures in the reference), and plain Jacobi
// then runs to the same error thresholds, so the solver stays the judge of the result.
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

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float z) {
    return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
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

// Output layer of a 16-row tile whose last activations are in h (BF16 [16][H]): row
// lane % 16, 16 inputs per lane, summed over the four lanes of the row. Every lane returns
// its row's value.
__device__ __forceinline__ float out_layer16(const FwdWeights& w, const __bf16* h) {
    const int l = threadIdx.x & 63;
    const int r = l & 15, q = l >> 4;
    float y = 0.0f;
#pragma unroll
    for (int k = 0; k < 16; k++) y += (float)h[r * H + q * 16 + k] * w.w4[q * 16 + k];
    y += __shfl_xor(y, 16, 64);
    y += __shfl_xor(y, 32, 64);
    return y + w.b4;
}

// ---- inference -------------------------------------------------------------------------

constexpr int kInferWaves = 4;

struct InferLds {
    FwdWeights w;
    alignas(16) __bf16 x[kInferWaves][16][KIN];
    alignas(16) __bf16 h[kInferWaves][2][16][H];
};

// p = max(0, y * target_std + target_mean) with y = MLP((features - mean) / std); written
// to pressures (hash order) and to the p / rho^2 mirror (schedule order).
__global__ __launch_bounds__(256) void k_mlp_infer(const float* __restrict__ params, const float* __restrict__ norm,
                                                   const float* __restrict__ feat, const uint32_t* __restrict__ sched,
                                                   const float* __restrict__ dens, uint32_t n, float* __restrict__ pres,
                                                   float* __restrict__ prho_m) {
    __shared__ InferLds s;
    load_fwd(params, s.w);
    __syncthreads();
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t tiles = (n + 15) / 16;
    for (uint32_t tile = blockIdx.x * kInferWaves + w; tile < tiles; tile += gridDim.x * kInferWaves) {
        const uint32_t r0 = tile * 16;
        for (int e = l; e < 16 * KIN; e += 64) {
            const int r = e / KIN, k = e % KIN;
            const uint32_t t = r0 + r;
            float v = 0.0f;
            if (k < kMlpIn && t < n) v = (feat[(size_t)t * kMlpIn + k] - norm[k]) / norm[kMlpIn + k];
            s.x[w][r][k] = tobf(v);
        }
        wave_lds_fence();
        forward16<false>(s.w, &s.x[w][0][0], &s.h[w][0][0][0], &s.h[w][1][0][0], &s.h[w][0][0][0], nullptr);
        const float y = out_layer16(s.w, &s.h[w][0][0][0]);
        const uint32_t t = r0 + (l & 15);
        if (l < 16 && t < n) {
            const uint32_t i = sched[t];
            float p = y * norm[2 * kMlpIn + 1] + norm[2 * kMlpIn];
            p = p > 0.0f ? p : 0.0f;
            pres[i] = p;
            prho_m[t] = prho_of(p, dens[i]);
        }
        wave_lds_fence();
    }
}

// Raw network output for already standardised inputs (no clamp, no de-standardisation).
__global__ __launch_bounds__(256) void k_mlp_forward_raw(const float* __restrict__ params,
                                                         const float* __restrict__ x, uint32_t n,
                                                         float* __restrict__ y) {
    __shared__ InferLds s;
    load_fwd(params, s.w);
    __syncthreads();
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t tiles = (n + 15) / 16;
    for (uint32_t tile = blockIdx.x * kInferWaves + w; tile < tiles; tile += gridDim.x * kInferWaves) {
        const uint32_t r0 = tile * 16;
        for (int e = l; e < 16 * KIN; e += 64) {
            const int r = e / KIN, k = e % KIN;
            const uint32_t t = r0 + r;
            s.x[w][r][k] = tobf(k < kMlpIn && t < n ? x[(size_t)t * kMlpIn + k] : 0.0f);
        }
        wave_lds_fence();
        forward16<false>(s.w, &s.x[w][0][0], &s.h[w][0][0][0], &s.h[w][1][0][0], &s.h[w][0][0][0], nullptr);
        const float v = out_layer16(s.w, &s.h[w][0][0][0]);
        if (l < 16 && r0 + l < n) y[r0 + l] = v;
        wave_lds_fence();
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

// ---- features and samples ------------------------------------------------------------------

// pprev[i] = plast[vals[i]]: the last solve's pressure of the particle now at hash index i
// (vals maps the current sort order to the order position_a had, which is the order of the
// previous solve of the same kind).
__global__ __launch_bounds__(kBlock) void k_prev(const uint32_t* __restrict__ vals, const float* __restrict__ plast,
                                                 uint32_t valid, uint32_t n, float* __restrict__ pprev) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t v = vals[i];
    pprev[i] = (valid && v < n) ? plast[v] : 0.0f;   // v >= n: no search has run yet
}

// The 12 features of the particle at schedule position t, over its listed pairs
// (r^2 <= h^2 in the reference's visit order; the first kMaxNbr of them):
//   0 (rho - rho0) / rho0        1 s dt / rho0              2 alpha
//   3 pair count                 4 previous pressure        5 sum W
//   6 sum W p_prev_j             7 |sum r_ij| / h           8 sum |r_ij|^2 / h^2
//   9 sum (v_i - v_j).r_ij / h   10 sum |v_i - v_j|^2       11 -s rho^2 / (dt alpha)
// with v = velocity_a (hash order) and W the reference kernel.
__global__ __launch_bounds__(kBlock) void k_features(Sph s, const uint32_t* __restrict__ sched,
                                                     const float4* __restrict__ pos_m, const uint32_t* __restrict__ nbr,
                                                     const uint32_t* __restrict__ cnt, const float4* __restrict__ vel_a,
                                                     const float* __restrict__ dens, const float* __restrict__ fac,
                                                     const float* __restrict__ src, const float* __restrict__ pprev,
                                                     float rho0, float dt, float* __restrict__ feat) {
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= s.n) return;
    const uint32_t i = sched[t];
    const float4 pi = pos_m[t];
    const float4 vi = vel_a[i];
    const uint32_t count = cnt[t];
    const uint32_t m = count < kMaxNbr ? count : kMaxNbr;
    float sw = 0.f, swp = 0.f, cx = 0.f, cy = 0.f, cz = 0.f, sr2 = 0.f, svr = 0.f, sv2 = 0.f;
    const uint4* tile = list_tile(nbr, t);
    for (uint32_t k = 0; k < m; k++) {
        const uint32_t jt = list_pick(tile[(k >> 2) * 16], k);
        const float4 pj = pos_m[jt];
        const float rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
        const float r2 = rx * rx + ry * ry + rz * rz;
        const float w = exact_w(sqrtf(r2), s);
        const uint32_t j = sched[jt];
        const float4 vj = vel_a[j];
        const float dx = vi.x - vj.x, dy = vi.y - vj.y, dz = vi.z - vj.z;
        sw += w;
        swp += w * pprev[j];
        cx += rx;
        cy += ry;
        cz += rz;
        sr2 += r2;
        svr += dx * rx + dy * ry + dz * rz;
        sv2 += dx * dx + dy * dy + dz * dz;
    }
    const float rho = dens[i], alpha = fac[i], si = src[i];
    float* f = feat + (size_t)t * kMlpIn;
    f[0] = (rho - rho0) / rho0;
    f[1] = si * dt / rho0;
    f[2] = alpha;
    f[3] = (float)count;
    f[4] = pprev[i];
    f[5] = sw;
    f[6] = swp;
    f[7] = sqrtf(cx * cx + cy * cy + cz * cz) / s.h;
    f[8] = sr2 / s.hh;
    f[9] = svr / s.h;
    f[10] = sv2;
    f[11] = (alpha > 1e-6f && rho > 1e-6f && dt > 1e-6f) ? -si * (rho * rho) / (dt * alpha) : 0.0f;
}

// Append `count` samples (features, pressure) of particles picked by a hash of (salt, k) to
// a ring buffer of `cap` samples starting at `head`.
__global__ __launch_bounds__(kBlock) void k_collect(const float* __restrict__ feat, const uint32_t* __restrict__ sched,
                                                    const float* __restrict__ pres, uint32_t n, uint32_t count,
                                                    uint32_t salt, uint32_t head, uint32_t cap, float* __restrict__ dx,
                                                    float* __restrict__ dt) {
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= count) return;
    const uint32_t t = sample_row(k, n, salt);
    const uint32_t slot = (head + k) % cap;
    for (int e = 0; e < kMlpIn; e++) dx[(size_t)slot * kMlpIn + e] = feat[(size_t)t * kMlpIn + e];
    dt[slot] = pres[sched[t]];
}

inline dim3 grid_for(uint32_t count) { return dim3((count + kBlock - 1) / kBlock); }

inline uint32_t infer_blocks(uint32_t n) {
    const uint32_t tiles = (n + 15) / 16;
    const uint32_t blocks = (tiles + kInferWaves - 1) / kInferWaves;
    return blocks > 1024 ? 1024 : (blocks ? blocks : 1);
}

}  // namespace

hipError_t learn_prev(fe_ctx* c, int which) {
    Launch L(c, "ws_prev");
    hipLaunchKernelGGL(k_prev, grid_for(c->n), dim3(kBlock), 0, c->stream, c->vals[0], c->ws.plast[which],
                       c->ws.plast_valid[which] ? 1u : 0u, c->n, c->ws.pprev);
    return hipGetLastError();
}

hipError_t learn_features(fe_ctx* c, int which) {
    hipError_t e = learn_prev(c, which);
    if (e != hipSuccess) return e;
    Launch L(c, "ws_features");
    hipLaunchKernelGGL(k_features, grid_for(c->n), dim3(kBlock), 0, c->stream, sph_of(c), c->sched[0], c->pos_m,
                       c->nbr, c->nbr_count, c->vel_a, c->dens, c->fac, c->src, c->ws.pprev,
                       c->params.target_density, c->params.dt, c->ws.feat);
    return hipGetLastError();
}

hipError_t learn_infer(fe_ctx* c, int which) {
    Launch L(c, "ws_infer");
    hipLaunchKernelGGL(k_mlp_infer, dim3(infer_blocks(c->n)), dim3(64 * kInferWaves), 0, c->stream,
                       c->ws.params[which], c->ws.norm[which], c->ws.feat, c->sched[0], c->dens, c->n, c->pres,
                       c->prho_m);
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
    hipLaunchKernelGGL(k_mlp_forward_raw, dim3(infer_blocks(n)), dim3(64 * kInferWaves), 0, c->stream,
                       c->ws.params[which], x, n, y);
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
