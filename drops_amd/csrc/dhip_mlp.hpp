// dhip_mlp.hpp -- the learned warm start: a per-DOF network 12 -> 64 -> 64 -> 64 -> 2 (tanh
// hidden layers) that predicts the next time level's change of [mu; c] from the last four time
// levels, the step's right-hand side and the preconditioner diagonal.  No reference
// counterpart: GMRES still runs to the same tolerance from the predicted start.
//
//   k_mlp_features  one lane per DOF: 12 features in float64 (differences of the history first),
//                   normalised, stored float32 [S][12].
//   k_mlp_infer     matrix cores: one wave per 16 DOFs, v_mfma_f32_16x16x32_bf16 for every layer
//                   (bf16 weights and layer inputs, float32 accumulation); weights staged in LDS.
//   k_mlp_train     training tile of 32 samples per block: forward with the same rounding points
//                   as k_mlp_infer (float32 FMA on bf16-rounded operands), backward, and the
//                   tile's gradient written to its own slot (no atomics).
//   k_mlp_reduce    gradient = sum of the tile slots in tile order; k_mlp_adam the update.
#pragma once

#include "dhip_common.hpp"

namespace dhip {

constexpr int MLP_NF = 12, MLP_H = 64, MLP_NO = 2;
constexpr int MLP_W1 = 0, MLP_B1 = MLP_W1 + MLP_H * MLP_NF, MLP_W2 = MLP_B1 + MLP_H, MLP_B2 = MLP_W2 + MLP_H * MLP_H,
              MLP_W3 = MLP_B2 + MLP_H, MLP_B3 = MLP_W3 + MLP_H * MLP_H, MLP_W4 = MLP_B3 + MLP_H,
              MLP_B4 = MLP_W4 + MLP_NO * MLP_H, MLP_NPARAM = MLP_B4 + MLP_NO;
static_assert(MLP_NPARAM == 9282, "parameter layout");

// packed bf16 weights for the matrix cores: [out][in] rows, layer 1 inputs padded to 32,
// layer 4 outputs padded to 16
constexpr int MLP_P1 = 0, MLP_P2 = MLP_P1 + MLP_H * 32, MLP_P3 = MLP_P2 + MLP_H * MLP_H,
              MLP_P4 = MLP_P3 + MLP_H * MLP_H, MLP_NPACK = MLP_P4 + 16 * MLP_H;

// float -> bfloat16 bits, round to nearest even
__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bfround(float f) { return __uint_as_float((uint32_t)f2bf(f) << 16); }

struct MlpNorm {
    double in_mean[MLP_NF], in_scale[MLP_NF];   // feature -> (feature - mean) * scale
    double out_mean[MLP_NO], out_scale[MLP_NO]; // correction = y * scale + mean
};

// Step scale: s_mu, s_c = RMS over the DOFs of the last change of mu and c (1 where it is 0).
// Two passes: per-block partial sums of squares, then one block adds them in block order.
constexpr int SCALE_B = 256;

__global__ void __launch_bounds__(SCALE_B) k_mlp_scale_part(const double *x0, const double *x1, int64_t n,
                                                            double *part) {
    double sm = 0.0, sc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * SCALE_B + threadIdx.x; i < n; i += (int64_t)gridDim.x * SCALE_B) {
        const double dm = x0[i] - x1[i], dc = x0[n + i] - x1[n + i];
        sm += dm * dm;
        sc += dc * dc;
    }
    sm = block_sum<SCALE_B>(sm);
    sc = block_sum<SCALE_B>(sc);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = sm;
        part[2 * blockIdx.x + 1] = sc;
    }
}

// one wave: lane-strided partials, then the wave's butterfly sum (the same order in every call)
__global__ void k_mlp_scale_final(const double *part, int nb, int64_t n, double *scale) {
    double sm = 0.0, sc = 0.0;
    for (int b = threadIdx.x & 63; b < nb; b += 64) {
        sm += part[2 * b];
        sc += part[2 * b + 1];
    }
    sm = wave_sum(sm);
    sc = wave_sum(sc);
    if (threadIdx.x != 0) return;
    const double rm = sqrt(sm / (double)n), rc = sqrt(sc / (double)n);
    scale[0] = rm > 0.0 ? rm : 1.0;
    scale[1] = rc > 0.0 ? rc : 1.0;
}

// ---------------------------------------------------------------------------------------------
// Features of DOF i for the step n -> n+1 (h0 = x_n, newest first; s = step scale):
//   c_n, dc_n/s_c, dc_{n-1}/s_c, dc_{n-2}/s_c, mu_n, dmu_n/s_mu, dmu_{n-1}/s_mu, dmu_{n-2}/s_mu,
//   (r_i/d_i)/s_mu, (r_{n+i}/d_{n+i})/s_c, d_i, d_{n+i}
// with dv_k = v_k - v_{k-1}, r = b - K x_n the residual of the new system at the current state
// (the local right-hand side of the correction), d the preconditioner diagonal.
struct FeatArgs {
    const double *h0, *h1, *h2, *h3;
    const double *res, *pcdiag;
    const double *scale;
    int64_t n;
    MlpNorm norm;
    float *X;  // [n][12]
};

__global__ void k_mlp_features(FeatArgs a) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const int64_t n = a.n;
    const double sm = a.scale[0], sc = a.scale[1];
    double f[MLP_NF];
    const double c0 = a.h0[n + i], c1 = a.h1[n + i], c2 = a.h2[n + i], c3 = a.h3[n + i];
    const double m0 = a.h0[i], m1 = a.h1[i], m2 = a.h2[i], m3 = a.h3[i];
    f[0] = c0;
    f[1] = (c0 - c1) / sc;
    f[2] = (c1 - c2) / sc;
    f[3] = (c2 - c3) / sc;
    f[4] = m0;
    f[5] = (m0 - m1) / sm;
    f[6] = (m1 - m2) / sm;
    f[7] = (m2 - m3) / sm;
    f[8] = (a.res[i] / a.pcdiag[i]) / sm;
    f[9] = (a.res[n + i] / a.pcdiag[n + i]) / sc;
    f[10] = a.pcdiag[i];
    f[11] = a.pcdiag[n + i];
#pragma unroll
    for (int k = 0; k < MLP_NF; ++k) a.X[i * MLP_NF + k] = (float)((f[k] - a.norm.in_mean[k]) * a.norm.in_scale[k]);
}

// The prediction is a correction of linear extrapolation in scaled units:
//   x_{n+1} = x_n + s * (dx_n/s + y * out_scale + out_mean).
// (Linear, not quadratic: extrapolating the solver's own tolerance-level errors with weights
// 3, -3, 1 lets them grow from step to step in the slowly converging modes of GMRES.)
__device__ __forceinline__ double extrap_scaled(double v0, double v1, double, double s) { return (v0 - v1) / s; }

// targets: ((x_next - x_n)/s - extrapolation - mean) / out_scale, per DOF [mu, c]
struct TargetArgs {
    const double *xn1, *h0, *h1, *h2;
    const double *scale;
    int64_t n;
    MlpNorm norm;
    float *T;  // [n][2]
};

__global__ void k_mlp_targets(TargetArgs a) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const int64_t n = a.n;
#pragma unroll
    for (int o = 0; o < MLP_NO; ++o) {
        const int64_t j = o == 0 ? i : n + i;  // o = 0: mu, o = 1: c
        const double s = a.scale[o];
        const double corr = (a.xn1[j] - a.h0[j]) / s - extrap_scaled(a.h0[j], a.h1[j], a.h2[j], s);
        a.T[2 * i + o] = (float)((corr - a.norm.out_mean[o]) / a.norm.out_scale[o]);
    }
}

// candidate starts: x0 = x_n + s * (extrapolation + y * out_scale + out_mean) (the network's) and
// xq = x_n + s * extrapolation (linear extrapolation alone)
struct CombineArgs {
    const double *h0, *h1, *h2;
    const float *Y;
    const double *scale;
    int64_t n;
    MlpNorm norm;
    double *x0, *xq;
};

__global__ void k_mlp_combine(CombineArgs a) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const int64_t n = a.n;
#pragma unroll
    for (int o = 0; o < MLP_NO; ++o) {
        const int64_t j = o == 0 ? i : n + i;
        const double s = a.scale[o];
        const double ex = extrap_scaled(a.h0[j], a.h1[j], a.h2[j], s);
        const double corr = (double)a.Y[2 * i + o] * a.norm.out_scale[o] + a.norm.out_mean[o];
        a.x0[j] = a.h0[j] + s * (ex + corr);
        a.xq[j] = a.h0[j] + s * ex;
    }
}

// ---------------------------------------------------------------------------------------------
// packed bf16 weights from the float32 master parameters
__global__ void k_mlp_pack(const float *theta, uint16_t *pk) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= MLP_NPACK) return;
    float v = 0.f;
    if (t < MLP_P2) {
        const int o = t / 32, k = t % 32;
        v = k < MLP_NF ? theta[MLP_W1 + o * MLP_NF + k] : 0.f;
    } else if (t < MLP_P3) {
        v = theta[MLP_W2 + (t - MLP_P2)];
    } else if (t < MLP_P4) {
        v = theta[MLP_W3 + (t - MLP_P3)];
    } else {
        const int o = (t - MLP_P4) / MLP_H, k = (t - MLP_P4) % MLP_H;
        v = o < MLP_NO ? theta[MLP_W4 + o * MLP_H + k] : 0.f;
    }
    pk[t] = f2bf(v);
}

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float floatx4_t __attribute__((ext_vector_type(4)));

constexpr int INFER_WAVES = 4;  // 64 DOFs per block

// Matrix-core inference.  v_mfma_f32_16x16x32_bf16 lane layout: lane l holds A[l%16][8(l/16)+j],
// B[8(l/16)+j][l%16] (j < 8) and D[4(l/16)+i][l%16] (i < 4).  A = layer input (16 DOFs x 32),
// B = W^T (32 inputs x 16 outputs), read as 8 consecutive bf16 of a packed [out][in] row.
__global__ void __launch_bounds__(64 * INFER_WAVES) k_mlp_infer(const float *X, int64_t S, const uint16_t *packed,
                                                               const float *theta, float *Y) {
    __shared__ __align__(16) uint16_t W[MLP_NPACK];
    __shared__ float bias[3 * MLP_H + MLP_NO];
    __shared__ __align__(16) uint16_t act[INFER_WAVES][16][MLP_H];
    for (int t = threadIdx.x; t < MLP_NPACK; t += blockDim.x) W[t] = packed[t];
    for (int t = threadIdx.x; t < MLP_H; t += blockDim.x) {
        bias[t] = theta[MLP_B1 + t];
        bias[MLP_H + t] = theta[MLP_B2 + t];
        bias[2 * MLP_H + t] = theta[MLP_B3 + t];
    }
    if (threadIdx.x < MLP_NO) bias[3 * MLP_H + threadIdx.x] = theta[MLP_B4 + threadIdx.x];
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int row = lane & 15, kq = lane >> 4;
    // each block stages the weights once and walks its share of 64-DOF tiles
    const int64_t ntiles = (S + 16 * INFER_WAVES - 1) / (16 * INFER_WAVES);
    for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int64_t m0 = (tile * INFER_WAVES + wv) * 16;
    // layer 1 input: features 8kq .. 8kq+7 of DOF m0+row (12 real features, zero padding)
    bf16x8_t a1;
    {
        const int64_t s = m0 + row;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int k = 8 * kq + j;
            const float v = (s < S && k < MLP_NF) ? X[s * MLP_NF + k] : 0.f;
            a1[j] = __builtin_bit_cast(__bf16, f2bf(v));
        }
    }
    auto bfrag = [&](int base, int ld, int n0, int k0) {
        return *reinterpret_cast<const bf16x8_t *>(&W[base + (n0 + row) * ld + k0 + 8 * kq]);
    };
    auto store_act = [&](const floatx4_t &d, int n0, const float *b) {
#pragma unroll
        for (int i = 0; i < 4; ++i) act[wv][4 * kq + i][n0 + row] = f2bf(tanhf(d[i] + b[n0 + row]));
    };
    // layer 1 (K = 32)
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        floatx4_t d = {0.f, 0.f, 0.f, 0.f};
        d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, bfrag(MLP_P1, 32, 16 * nt, 0), d, 0, 0, 0);
        store_act(d, 16 * nt, bias);
    }
    __syncthreads();
    // layers 2 and 3 (K = 64)
#pragma unroll
    for (int layer = 0; layer < 2; ++layer) {
        const int base = layer == 0 ? MLP_P2 : MLP_P3;
        bf16x8_t a[2];
#pragma unroll
        for (int kk = 0; kk < 2; ++kk) a[kk] = *reinterpret_cast<const bf16x8_t *>(&act[wv][row][32 * kk + 8 * kq]);
        floatx4_t d[4];
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            d[nt] = floatx4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kk = 0; kk < 2; ++kk)
                d[nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[kk], bfrag(base, MLP_H, 16 * nt, 32 * kk), d[nt], 0, 0, 0);
        }
        __syncthreads();  // every lane has read its layer input
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) store_act(d[nt], 16 * nt, bias + (layer + 1) * MLP_H);
        __syncthreads();
    }
    // output layer (16 padded outputs, 2 real)
    {
        bf16x8_t a[2];
#pragma unroll
        for (int kk = 0; kk < 2; ++kk) a[kk] = *reinterpret_cast<const bf16x8_t *>(&act[wv][row][32 * kk + 8 * kq]);
        floatx4_t d = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kk = 0; kk < 2; ++kk)
            d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[kk], bfrag(MLP_P4, MLP_H, 0, 32 * kk), d, 0, 0, 0);
        if (row < MLP_NO) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int64_t s = m0 + 4 * kq + i;
                if (s < S) Y[s * MLP_NO + row] = d[i] + bias[3 * MLP_H + row];
            }
        }
    }
    __syncthreads();  // the next tile overwrites this wave's activations
    }
}

// ---------------------------------------------------------------------------------------------
// training
constexpr int TS = 32;        // samples per block
constexpr int TRAIN_T = 256;  // threads per block

struct TrainArgs {
    const float *X, *T;       // [S][12], [S][2]
    const int32_t *idx;       // minibatch sample ids
    int64_t B;                // minibatch size
    const float *theta;
    float *part;              // [tiles][MLP_NPARAM + 1] (last: loss)
};

// out[s][n] (+)= sum_k A[s][k] * Wt[n][k], tile TS x 64: thread -> 2 samples x 4 outputs
template <int K>
__device__ inline void fwd_gemm(const float (*A)[MLP_H], const float *Wt, const float *b, float (*Hout)[MLP_H]) {
    const int t = threadIdx.x, s0 = 2 * (t >> 4), n0 = 4 * (t & 15);
    float acc[2][4];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = 0.f;
    for (int k = 0; k < K; ++k) {
        const float a0 = bfround(A[s0][k]), a1 = bfround(A[s0 + 1][k]);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const float w = Wt[(n0 + j) * K + k];
            acc[0][j] = __fmaf_rn(a0, w, acc[0][j]);
            acc[1][j] = __fmaf_rn(a1, w, acc[1][j]);
        }
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) Hout[s0 + i][n0 + j] = tanhf(acc[i][j] + b[n0 + j]);
}

__global__ void __launch_bounds__(TRAIN_T) k_mlp_train(TrainArgs a) {
    __shared__ float Wr[MLP_NPARAM];               // bf16-rounded weights, float32 biases
    __shared__ float A0[TS][MLP_H];                // inputs (first 12 columns)
    __shared__ float H[3][TS][MLP_H];              // tanh outputs (unrounded)
    __shared__ float D0[TS][MLP_H], D1[TS][MLP_H];  // backward buffers
    __shared__ float dY[TS][MLP_NO];
    __shared__ float lsum[TS];
    const int t = threadIdx.x;
    for (int p = t; p < MLP_NPARAM; p += TRAIN_T) {
        const bool isw = (p < MLP_B1) || (p >= MLP_W2 && p < MLP_B2) || (p >= MLP_W3 && p < MLP_B3) ||
                         (p >= MLP_W4 && p < MLP_B4);
        Wr[p] = isw ? bfround(a.theta[p]) : a.theta[p];
    }
    const int64_t base = (int64_t)blockIdx.x * TS;
    for (int e = t; e < TS * MLP_NF; e += TRAIN_T) {
        const int s = e / MLP_NF, k = e % MLP_NF;
        const int64_t g = base + s;
        A0[s][k] = g < a.B ? a.X[(int64_t)a.idx[g] * MLP_NF + k] : 0.f;
    }
    __syncthreads();
    fwd_gemm<MLP_NF>(A0, Wr + MLP_W1, Wr + MLP_B1, H[0]);
    __syncthreads();
    fwd_gemm<MLP_H>(H[0], Wr + MLP_W2, Wr + MLP_B2, H[1]);
    __syncthreads();
    fwd_gemm<MLP_H>(H[1], Wr + MLP_W3, Wr + MLP_B3, H[2]);
    __syncthreads();
    // output, residual, loss (0.5 |y - t|^2 per sample), dY = (y - t) / B
    if (t < TS * MLP_NO) {
        const int s = t / MLP_NO, o = t % MLP_NO;
        const int64_t g = base + s;
        float y = 0.f;
        for (int k = 0; k < MLP_H; ++k) y = __fmaf_rn(bfround(H[2][s][k]), Wr[MLP_W4 + o * MLP_H + k], y);
        y += Wr[MLP_B4 + o];
        const float d = g < a.B ? y - a.T[(int64_t)a.idx[g] * MLP_NO + o] : 0.f;
        dY[s][o] = d / (float)a.B;
        D0[s][o] = 0.5f * d * d;  // scratch for the loss
    }
    __syncthreads();
    if (t < TS) lsum[t] = D0[t][0] + D0[t][1];
    __syncthreads();
    float *G = a.part + (size_t)blockIdx.x * (MLP_NPARAM + 1);
    if (t == 0) {
        float l = 0.f;
        for (int s = 0; s < TS; ++s) l += lsum[s];
        G[MLP_NPARAM] = l;
    }
    // output layer gradients
    for (int p = t; p < MLP_NO * MLP_H; p += TRAIN_T) {
        const int o = p / MLP_H, k = p % MLP_H;
        float g = 0.f;
        for (int s = 0; s < TS; ++s) g = __fmaf_rn(dY[s][o], bfround(H[2][s][k]), g);
        G[MLP_W4 + p] = g;
    }
    if (t < MLP_NO) {
        float g = 0.f;
        for (int s = 0; s < TS; ++s) g += dY[s][t];
        G[MLP_B4 + t] = g;
    }
    // dZ3 = (dY W4) * (1 - h3^2) into D0
    for (int e = t; e < TS * MLP_H; e += TRAIN_T) {
        const int s = e / MLP_H, k = e % MLP_H;
        float v = 0.f;
#pragma unroll
        for (int o = 0; o < MLP_NO; ++o) v = __fmaf_rn(dY[s][o], Wr[MLP_W4 + o * MLP_H + k], v);
        const float h = H[2][s][k];
        D0[s][k] = v * (1.f - h * h);
    }
    __syncthreads();
    // hidden layers 3, 2, 1: weight/bias gradients and the next dZ
    float (*Dz)[MLP_H] = D0;
    float (*Dn)[MLP_H] = D1;
    for (int layer = 2; layer >= 0; --layer) {
        const int wofs = layer == 2 ? MLP_W3 : layer == 1 ? MLP_W2 : MLP_W1;
        const int bofs = layer == 2 ? MLP_B3 : layer == 1 ? MLP_B2 : MLP_B1;
        const int K = layer == 0 ? MLP_NF : MLP_H;
        const float (*In)[MLP_H] = layer == 0 ? A0 : H[layer - 1];
        // dW[n][k] = sum_s dZ[s][n] * bf(In[s][k])
        for (int p = t; p < MLP_H * K; p += TRAIN_T) {
            const int nn = p / K, k = p % K;
            float g = 0.f;
            for (int s = 0; s < TS; ++s) g = __fmaf_rn(Dz[s][nn], bfround(In[s][k]), g);
            G[wofs + p] = g;
        }
        if (t < MLP_H) {
            float g = 0.f;
            for (int s = 0; s < TS; ++s) g += Dz[s][t];
            G[bofs + t] = g;
        }
        if (layer > 0) {
            // dZ_prev[s][k] = (sum_n dZ[s][n] * W[n][k]) * (1 - h^2)
            for (int e = t; e < TS * MLP_H; e += TRAIN_T) {
                const int s = e / MLP_H, k = e % MLP_H;
                float v = 0.f;
                for (int nn = 0; nn < MLP_H; ++nn) v = __fmaf_rn(Dz[s][nn], Wr[wofs + nn * MLP_H + k], v);
                const float h = H[layer - 1][s][k];
                Dn[s][k] = v * (1.f - h * h);
            }
            __syncthreads();
            float (*tmp)[MLP_H] = Dz;
            Dz = Dn;
            Dn = tmp;
        }
    }
}

// grad[p] = sum over tiles (tile order) of part[tile][p]; p = MLP_NPARAM is the loss
__global__ void k_mlp_reduce(const float *part, int ntiles, double *grad) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > MLP_NPARAM) return;
    double s = 0.0;
    for (int b = 0; b < ntiles; ++b) s += (double)part[(size_t)b * (MLP_NPARAM + 1) + p];
    grad[p] = s;
}

struct AdamArgs {
    float *theta, *m, *v;
    const double *grad;
    float lr, beta1, beta2, eps;
    int step;  // 1-based
};

__global__ void k_mlp_adam(AdamArgs a) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= MLP_NPARAM) return;
    const float g = (float)a.grad[p];
    const float m = a.beta1 * a.m[p] + (1.f - a.beta1) * g;
    const float v = a.beta2 * a.v[p] + (1.f - a.beta2) * g * g;
    a.m[p] = m;
    a.v[p] = v;
    const float mh = m / (1.f - powf(a.beta1, (float)a.step));
    const float vh = v / (1.f - powf(a.beta2, (float)a.step));
    a.theta[p] -= a.lr * mh / (sqrtf(vh) + a.eps);
}

}  // namespace dhip
