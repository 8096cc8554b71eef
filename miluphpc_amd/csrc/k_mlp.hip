// k_mlp.hip -- learned smoothing length: a 7 -> 64 -> 64 -> 1 tanh MLP with BF16 weights
// and activations on the matrix cores (v_mfma_f32_16x16x16_bf16, FP32 accumulation),
// FP32 master weights and Adam.  It predicts log(h_converged / h_previous) from a
// particle's state; the plain h-iteration starts from the prediction and has the last word.
// No counterpart in the reference.
//
// Work split, training: one wavefront owns 64 samples.  Per layer it computes a 64 x 64 tile
// product as 4 x 4 output blocks of 16 x 16 with K in steps of 16; activations are staged in
// LDS as BF16 row-major [sample][feature] (the backward pass reads them).  Training writes
// per-wave gradient partials that one block folds in wave order (deterministic), then Adam
// updates the weights.  Inference keeps the activations in registers (mlp_forward_16).
#include <math.h>

#include "mhpc_internal.h"

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short f2bf(float f) {
    unsigned u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);   // round to nearest even (inputs are finite)
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short b) { return __uint_as_float(((unsigned)b) << 16); }

// D(16x16) += A(16x16) B(16x16).  Fragments: lane l holds A[l%16][4(l/16)..+3],
// B[4(l/16)..+3][l%16] and D[4(l/16)+r][l%16], r = 0..3.
__device__ __forceinline__ f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

#define MLP_IN 8      // 7 features + 1 zero pad
#define MLP_H 64
#define MLP_KPAD 16   // layer-1 K padded to one MFMA step
// LDS row strides (BF16 elements): four more than the row length, so that the 16 rows one MFMA
// operand read touches fall in different banks (with the bare length every row starts in the
// same bank: 16-way conflicts)
#define MLP_KLD (MLP_KPAD + 4)
#define MLP_HLD (MLP_H + 4)
#define MLP_NPARAM (MLP_H * MLP_IN + MLP_H + MLP_H * MLP_H + MLP_H + MLP_H + 1)   // 4801

// parameter offsets in the flat FP32 vector
#define P_W1 0
#define P_B1 (P_W1 + MLP_H * MLP_IN)
#define P_W2 (P_B1 + MLP_H)
#define P_B2 (P_W2 + MLP_H * MLP_H)
#define P_W3 (P_B2 + MLP_H)
#define P_B3 (P_W3 + MLP_H)

static_assert(MLP_NPARAM == MHPC_MLP_NPARAM, "parameter layout of include/mhpc.h");

int mlp_num_params() { return MLP_NPARAM; }

// weights as the kernels stage them: W1 [64][16] bf16 (cols >= 8 zero), W2 [64][64] bf16
struct alignas(16) MlpLds {
    unsigned short w1[MLP_H][MLP_KLD];
    unsigned short w2[MLP_H][MLP_HLD];
    float b1[MLP_H], b2[MLP_H], w3[MLP_H];
    float b3;
};

__device__ void load_weights(MlpLds &W, const float *__restrict__ p) {
    for (int t = threadIdx.x; t < MLP_H * MLP_KPAD; t += blockDim.x) {
        const int o = t / MLP_KPAD, k = t % MLP_KPAD;
        W.w1[o][k] = k < MLP_IN ? f2bf(p[P_W1 + o * MLP_IN + k]) : 0;
    }
    for (int t = threadIdx.x; t < MLP_H * MLP_H; t += blockDim.x) W.w2[t / MLP_H][t % MLP_H] = f2bf(p[P_W2 + t]);
    for (int t = threadIdx.x; t < MLP_H; t += blockDim.x) {
        W.b1[t] = p[P_B1 + t];
        W.b2[t] = p[P_B2 + t];
        W.w3[t] = p[P_W3 + t];
    }
    if (threadIdx.x == 0) W.b3 = p[P_B3];
}

// per-wave activations
struct alignas(16) MlpWave {
    unsigned short x[64][MLP_KLD];
    unsigned short a1[64][MLP_HLD];
    unsigned short a2[64][MLP_HLD];
};

// tanh(x) = 1 - 2 / (exp(2x) + 1) from the hardware exp2 and reciprocal estimates: absolute
// error ~1e-7, far below the BF16 rounding the activations are stored with; saturates to
// +-1 for large |x|
__device__ __forceinline__ float tanh_fast(float x) {
    const float e = __builtin_amdgcn_exp2f(x * 2.8853900817779268f);   // exp(2x) = 2^(2x log2 e)
    return 1.f - 2.f * __builtin_amdgcn_rcpf(e + 1.f);
}

// Z(64 x 64) = act(Ain(64 x K) Wt(K x 64) + b); Ain row-major [sample][k], W row-major
// [out][k] (so B[k][n] = W[n][k]); writes tanh(Z) as BF16 into Aout.
template <int K, int LDA>
__device__ __forceinline__ void dense_tanh(const unsigned short (*Ain)[LDA],
                                           const unsigned short (*Wt)[LDA], const float *bias,
                                           unsigned short (*Aout)[MLP_HLD]) {
    const int lane = wave_lane(), r16 = lane & 15, q = lane >> 4;
    for (int mt = 0; mt < 4; mt++) {
        for (int nt = 0; nt < 4; nt++) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kt = 0; kt < K / 16; kt++) {
                // four consecutive k of one row: one 8-byte LDS read per operand
                const int k0 = kt * 16 + 4 * q;
                const bf16x4 a = *reinterpret_cast<const bf16x4 *>(&Ain[mt * 16 + r16][k0]);
                const bf16x4 b = *reinterpret_cast<const bf16x4 *>(&Wt[nt * 16 + r16][k0]);
                acc = mfma16(a, b, acc);
            }
            const int n = nt * 16 + r16;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int m = mt * 16 + 4 * q + r;
                Aout[m][n] = f2bf(tanh_fast(acc[r] + bias[n]));
            }
        }
    }
}

__device__ __forceinline__ float mlp_forward_wave(const MlpLds &W, MlpWave &S) {
    dense_tanh<MLP_KPAD, MLP_KLD>(S.x, W.w1, W.b1, S.a1);
    wave_sync();
    dense_tanh<MLP_H, MLP_HLD>(S.a1, W.w2, W.b2, S.a2);
    wave_sync();
    const int lane = wave_lane();
    float y = W.b3;
#pragma unroll 8
    for (int k = 0; k < MLP_H; k++) y += bf2f(S.a2[lane][k]) * W.w3[k];
    return y;
}

// ------------------------------------------------------------------------------------
// features, caller order: 0 log10 h, 1 neighbour count / target, 2 log10 rho, 3 div v h/c
// (div v = -drho/dt / rho), 4 |curl v| h / c, 5 |grad rho| h / rho, 6 dt c / h;
// c = cs + (tiny) so cold gas gives finite values.  Normalised with the stored mean / std.
__global__ __launch_bounds__(256) void k_mlp_features(int64_t n, const double *__restrict__ h,
                                                      const int32_t *__restrict__ noi,
                                                      const double *__restrict__ rho,
                                                      const double *__restrict__ drhodt,
                                                      const double *__restrict__ curl,
                                                      const double *__restrict__ gradrho,
                                                      const double *__restrict__ cs, double dt,
                                                      int n_target, const float *__restrict__ norm,
                                                      float clip, float *__restrict__ X,
                                                      double *__restrict__ hprev) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double hi = h[i];
    const double r = rho[i] > 0.0 ? rho[i] : 1e-300;
    const double c = cs[i] + 1e-12;
    float f[7];
    f[0] = (float)log10(hi);
    f[1] = (float)((double)noi[i] / (double)n_target);
    f[2] = (float)log10(r);
    f[3] = (float)(-drhodt[i] / r * hi / c);
    f[4] = (float)(curl[i] * hi / c);
    f[5] = (float)(gradrho[i] * hi / r);
    f[6] = (float)(dt * c / hi);
#pragma unroll
    for (int k = 0; k < 7; k++) {
        float v = (f[k] - norm[k]) / norm[8 + k];
        v = fminf(fmaxf(v, -clip), clip);
        X[i * MLP_IN + k] = v;
    }
    X[i * MLP_IN + 7] = 0.f;
    hprev[i] = hi;
}

// normalised features are clipped to +-MLP_CLIP standard deviations
#define MLP_CLIP 8.f

// normalise raw features in place (after the statistics have been taken from them)
__global__ __launch_bounds__(256) void k_mlp_renorm(int64_t n, const float *__restrict__ norm, float *__restrict__ X) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= n * MLP_IN) return;
    const int k = (int)(t % MLP_IN);
    if (k >= 7) return;
    const float v = (X[t] - norm[k]) / norm[8 + k];
    X[t] = fminf(fmaxf(v, -MLP_CLIP), MLP_CLIP);
}

void launch_mlp_renorm(mhpc_ctx *c) {
    if (c->n <= 0) return;
    hipLaunchKernelGGL(k_mlp_renorm, dim3((unsigned)ceil_div(c->n * MLP_IN, 256)), dim3(256), 0, c->stream, c->n,
                       c->mlp_norm, c->mlp_x);
}

// Inference with the activations kept in registers.  The layers are computed transposed --
// hidden units as rows, samples as columns -- so the fragment a lane holds after one layer,
// D[4q + r][c] (q = lane / 16, c = lane % 16, r = 0..3), is exactly its B operand of the next
// layer, B[k = 4q + r][c]: no LDS round trip and no layout change between the layers.  The
// weights are the A operands and stay in registers for the whole kernel.  One step takes 16
// samples (one column block): layer 1 is 4 MFMAs (K = 16, features 8..15 zero), layer 2 is 16
// (4 row blocks x 4 K steps), and the output a per-lane partial over the lane's 16 hidden
// units plus two cross-lane sums.  Same BF16 roundings as the training kernel's forward
// (inputs, weights and both layers' activations), FP32 accumulation.
struct MlpRegs {
    bf16x4 w1[4];                         // W1[mt*16 + c][4q .. 4q+3]
    bf16x4 w2[4][4];                      // W2[mt*16 + c][kt*16 + 4q .. +3]
    float b1[4][4], b2[4][4], w3[4][4];   // [mt][r]: hidden unit mt*16 + 4q + r
    float b3;
};

__device__ __forceinline__ void load_regs(MlpRegs &R, const float *__restrict__ p) {
    const int lane = wave_lane(), c = lane & 15, q = lane >> 4;
#pragma unroll
    for (int mt = 0; mt < 4; mt++) {
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int k = 4 * q + j;
            R.w1[mt][j] = (short)(k < MLP_IN ? f2bf(p[P_W1 + (mt * 16 + c) * MLP_IN + k]) : 0);
        }
#pragma unroll
        for (int kt = 0; kt < 4; kt++)
#pragma unroll
            for (int j = 0; j < 4; j++)
                R.w2[mt][kt][j] = (short)f2bf(p[P_W2 + (mt * 16 + c) * MLP_H + kt * 16 + 4 * q + j]);
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int hu = mt * 16 + 4 * q + r;
            R.b1[mt][r] = p[P_B1 + hu];
            R.b2[mt][r] = p[P_B2 + hu];
            R.w3[mt][r] = p[P_W3 + hu];
        }
    }
    R.b3 = p[P_B3];
}

// the network's output for samples s0 .. s0 + 15 (lane c and its copies c + 16, c + 32, c + 48
// return the value of sample s0 + c; samples at or beyond n read zero features)
__device__ __forceinline__ float mlp_forward_16(const MlpRegs &R, const float *__restrict__ X, int64_t s0,
                                                int64_t n) {
    const int lane = wave_lane(), c = lane & 15, q = lane >> 4;
    bf16x4 xb = {0, 0, 0, 0};
    const int64_t s = s0 + c;
    if (q < 2 && s < n) {
        const float4 v = *reinterpret_cast<const float4 *>(X + s * MLP_IN + 4 * q);
        xb[0] = (short)f2bf(v.x);
        xb[1] = (short)f2bf(v.y);
        xb[2] = (short)f2bf(v.z);
        xb[3] = (short)f2bf(v.w);
    }
    bf16x4 a1[4];
#pragma unroll
    for (int mt = 0; mt < 4; mt++) {
        const f32x4 d = mfma16(R.w1[mt], xb, f32x4{0.f, 0.f, 0.f, 0.f});
#pragma unroll
        for (int r = 0; r < 4; r++) a1[mt][r] = (short)f2bf(tanh_fast(d[r] + R.b1[mt][r]));
    }
    float part = 0.f;
#pragma unroll
    for (int mt = 0; mt < 4; mt++) {
        f32x4 d = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kt = 0; kt < 4; kt++) d = mfma16(R.w2[mt][kt], a1[kt], d);
#pragma unroll
        for (int r = 0; r < 4; r++) part += bf2f(f2bf(tanh_fast(d[r] + R.b2[mt][r]))) * R.w3[mt][r];
    }
    part += __shfl_xor(part, 16);
    part += __shfl_xor(part, 32);
    return part + R.b3;
}

// inference: h = h_prev * exp(y * sigma_t + mu_t), clamped to [1/4, 4] x h_prev.  A bounded
// grid of wavefronts looping over 16-sample steps; each wave converts its weight fragments once.
#define MLP_INF_WAVES 4
#define MLP_INF_BLOCKS 1024
__global__ __launch_bounds__(64 * MLP_INF_WAVES) void k_mlp_predict(int64_t n, const float *__restrict__ X,
                                                                   const float *__restrict__ params,
                                                                   const float *__restrict__ norm,
                                                                   const double *__restrict__ hprev,
                                                                   double *__restrict__ h_out) {
    MlpRegs R;
    load_regs(R, params);
    const int w = threadIdx.x / 64, lane = wave_lane();
    const double mu = norm[16], sd = norm[16 + 1];
    for (int64_t s0 = ((int64_t)blockIdx.x * MLP_INF_WAVES + w) * 16; s0 < n;
         s0 += (int64_t)gridDim.x * MLP_INF_WAVES * 16) {
        const float y = mlp_forward_16(R, X, s0, n);
        const int64_t i = s0 + lane;
        if (lane < 16 && i < n) {
            double t = (double)y * sd + mu;
            t = fmin(fmax(t, -1.3862943611198906), 1.3862943611198906);   // ln 4
            h_out[i] = hprev[i] * exp(t);
        }
    }
}

// forward only on given inputs (test hook): Y[i] = y
__global__ __launch_bounds__(64 * MLP_INF_WAVES) void k_mlp_forward_raw(int64_t n, const float *__restrict__ X,
                                                                       const float *__restrict__ params,
                                                                       float *__restrict__ Y) {
    MlpRegs R;
    load_regs(R, params);
    const int w = threadIdx.x / 64, lane = wave_lane();
    for (int64_t s0 = ((int64_t)blockIdx.x * MLP_INF_WAVES + w) * 16; s0 < n;
         s0 += (int64_t)gridDim.x * MLP_INF_WAVES * 16) {
        const float y = mlp_forward_16(R, X, s0, n);
        const int64_t i = s0 + lane;
        if (lane < 16 && i < n) Y[i] = y;
    }
}

// ------------------------------------------------------------------------------------
// training: per wave 64 samples -> forward, loss, backward; gradient partials per wave
#define MLP_TR_WAVES 2
struct MlpTrainWave {
    MlpWave f;
    unsigned short dz2[64][MLP_HLD];
    unsigned short dz1[64][MLP_HLD];
    float dy[64];
};

__global__ __launch_bounds__(64 * MLP_TR_WAVES) void k_mlp_train_grad(int64_t nb, const float *__restrict__ X,
                                                                     const float *__restrict__ T,
                                                                     const int32_t *__restrict__ idx,
                                                                     const float *__restrict__ params,
                                                                     float *__restrict__ partial,
                                                                     float *__restrict__ loss_partial) {
    __shared__ MlpLds W;
    __shared__ MlpTrainWave Sw[MLP_TR_WAVES];
    load_weights(W, params);
    __syncthreads();
    const int w = threadIdx.x / 64, lane = wave_lane(), r16 = lane & 15, q = lane >> 4;
    MlpTrainWave &S = Sw[w];
    const int64_t wid = (int64_t)blockIdx.x * MLP_TR_WAVES + w;
    const int64_t base = wid * 64;
    float *g = partial + wid * MLP_NPARAM;
    if (base >= nb) return;
    const int64_t s = base + lane;
    const bool valid = s < nb;
    const int64_t src = valid ? (idx ? (int64_t)idx[s] : s) : 0;
#pragma unroll
    for (int k = 0; k < MLP_KPAD; k++) S.f.x[lane][k] = (valid && k < MLP_IN) ? f2bf(X[src * MLP_IN + k]) : 0;
    wave_sync();
    const float y = mlp_forward_wave(W, S.f);
    const float e = valid ? y - T[src] : 0.f;
    const float dy = e / (float)nb;
    S.dy[lane] = dy;
    loss_partial[wid] = wave_sum(0.5f * e * e);
    wave_sync();
    // layer 3: gW3[k] = sum_i dy_i a2[i][k], gb3 = sum dy; dz2 = dy w3 (1 - a2^2)
    {
        float gw3 = 0.f;
        for (int i = 0; i < 64; i++) gw3 += S.dy[i] * bf2f(S.f.a2[i][lane]);
        g[P_W3 + lane] = gw3;
        const float gb3 = wave_sum(dy);
        if (lane == 0) g[P_B3] = gb3;
        for (int k = 0; k < MLP_H; k++) {
            const float a = bf2f(S.f.a2[lane][k]);
            S.dz2[lane][k] = f2bf(dy * W.w3[k] * (1.f - a * a));
        }
    }
    wave_sync();
    // gb2[j] = sum_i dz2[i][j]; gW2[j][k] = sum_i dz2[i][j] a1[i][k] (M = j, K = i, N = k)
    {
        float gb2 = 0.f;
        for (int i = 0; i < 64; i++) gb2 += bf2f(S.dz2[i][lane]);
        g[P_B2 + lane] = gb2;
        for (int mt = 0; mt < 4; mt++)
            for (int nt = 0; nt < 4; nt++) {
                f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int kt = 0; kt < 4; kt++) {
                    bf16x4 a, b;
                    const int i0 = kt * 16 + 4 * q;
#pragma unroll
                    for (int u = 0; u < 4; u++) {
                        a[u] = (short)S.dz2[i0 + u][mt * 16 + r16];
                        b[u] = (short)S.f.a1[i0 + u][nt * 16 + r16];
                    }
                    acc = mfma16(a, b, acc);
                }
#pragma unroll
                for (int r = 0; r < 4; r++) g[P_W2 + (mt * 16 + 4 * q + r) * MLP_H + nt * 16 + r16] = acc[r];
            }
    }
    // dA1[i][k] = sum_j dz2[i][j] W2[j][k] (M = i, K = j, N = k); dz1 = dA1 (1 - a1^2)
    for (int mt = 0; mt < 4; mt++)
        for (int nt = 0; nt < 4; nt++) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kt = 0; kt < 4; kt++) {
                bf16x4 a, b;
                const int j0 = kt * 16 + 4 * q;
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    a[u] = (short)S.dz2[mt * 16 + r16][j0 + u];
                    b[u] = (short)W.w2[j0 + u][nt * 16 + r16];
                }
                acc = mfma16(a, b, acc);
            }
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int i = mt * 16 + 4 * q + r, k = nt * 16 + r16;
                const float a1 = bf2f(S.f.a1[i][k]);
                S.dz1[i][k] = f2bf(acc[r] * (1.f - a1 * a1));
            }
        }
    wave_sync();
    // gb1[j] = sum_i dz1[i][j]; gW1[j][f] = sum_i dz1[i][j] x[i][f] (M = j, K = i, N = f)
    {
        float gb1 = 0.f;
        for (int i = 0; i < 64; i++) gb1 += bf2f(S.dz1[i][lane]);
        g[P_B1 + lane] = gb1;
        for (int mt = 0; mt < 4; mt++) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kt = 0; kt < 4; kt++) {
                bf16x4 a, b;
                const int i0 = kt * 16 + 4 * q;
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    a[u] = (short)S.dz1[i0 + u][mt * 16 + r16];
                    b[u] = (short)S.f.x[i0 + u][r16];
                }
                acc = mfma16(a, b, acc);
            }
#pragma unroll
            for (int r = 0; r < 4; r++)
                if (r16 < MLP_IN) g[P_W1 + (mt * 16 + 4 * q + r) * MLP_IN + r16] = acc[r];
        }
    }
}

// fold per-wave partials in a fixed order (deterministic), in two levels: chunks of
// MLP_RED_CHUNK waves in parallel (stage 1), then the chunk sums (stage 2);
// loss = mean of 0.5 (y - t)^2
#define MLP_RED_CHUNK 32
__global__ __launch_bounds__(256) void k_mlp_reduce1(int nw, const float *__restrict__ partial,
                                                     const float *__restrict__ loss_partial,
                                                     float *__restrict__ part2, float *__restrict__ loss2) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int w0 = blockIdx.y * MLP_RED_CHUNK;
    const int w1 = w0 + MLP_RED_CHUNK < nw ? w0 + MLP_RED_CHUNK : nw;
    if (p < MLP_NPARAM) {
        float s = 0.f;
        for (int w = w0; w < w1; w++) s += partial[(int64_t)w * MLP_NPARAM + p];
        part2[(int64_t)blockIdx.y * MLP_NPARAM + p] = s;
    }
    if (p == 0) {
        float s = 0.f;
        for (int w = w0; w < w1; w++) s += loss_partial[w];
        loss2[blockIdx.y] = s;
    }
}

__global__ __launch_bounds__(256) void k_mlp_reduce2(int nchunks, int64_t nb, const float *__restrict__ part2,
                                                     const float *__restrict__ loss2, float *__restrict__ grad,
                                                     float *__restrict__ loss) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < MLP_NPARAM) {
        float s = 0.f;
        for (int c = 0; c < nchunks; c++) s += part2[(int64_t)c * MLP_NPARAM + p];
        grad[p] = s;
    }
    if (p == 0) {
        float s = 0.f;
        for (int c = 0; c < nchunks; c++) s += loss2[c];
        *loss = s / (float)nb;
    }
}

// Adam (beta1 0.9, beta2 0.999, eps 1e-8) on the FP32 master weights
__global__ __launch_bounds__(256) void k_mlp_adam(float *__restrict__ params, float *__restrict__ m1,
                                                  float *__restrict__ m2, const float *__restrict__ grad,
                                                  float lr, float bc1, float bc2) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= MLP_NPARAM) return;
    const float gg = grad[p];
    const float a = 0.9f * m1[p] + 0.1f * gg;
    const float b = 0.999f * m2[p] + 0.001f * gg * gg;
    m1[p] = a;
    m2[p] = b;
    params[p] -= lr * (a / bc1) / (sqrtf(b / bc2) + 1e-8f);
}

// ------------------------------------------------------------------------------------
void launch_mlp_features(mhpc_ctx *c, double dt) {
    if (c->n <= 0) return;
    double **f = c->f;
    hipLaunchKernelGGL(k_mlp_features, dim3((unsigned)ceil_div(c->n, 256)), dim3(256), 0, c->stream, c->n,
                       f[MHPC_F_SML], c->noi, f[MHPC_F_RHO], f[MHPC_F_DRHODT], f[MHPC_F_CURLV],
                       f[MHPC_F_GRADRHO], f[MHPC_F_CS], dt, c->p.n_target, c->mlp_norm,
                       c->mlp_norm_set ? MLP_CLIP : 3.0e38f, c->mlp_x, c->mlp_hprev);
}

void launch_mlp_predict(mhpc_ctx *c) {
    if (c->n <= 0) return;
    const unsigned g = (unsigned)std::min<int64_t>(ceil_div(c->n, 16 * MLP_INF_WAVES), MLP_INF_BLOCKS);
    hipLaunchKernelGGL(k_mlp_predict, dim3(g), dim3(64 * MLP_INF_WAVES), 0, c->stream, c->n, c->mlp_x,
                       c->mlp_params, c->mlp_norm, c->mlp_hprev, c->f[MHPC_F_SML]);
}

void launch_mlp_forward_raw(mhpc_ctx *c, int64_t n, const float *X, float *Y) {
    const unsigned g = (unsigned)std::min<int64_t>(ceil_div(n, 16 * MLP_INF_WAVES), MLP_INF_BLOCKS);
    hipLaunchKernelGGL(k_mlp_forward_raw, dim3(g), dim3(64 * MLP_INF_WAVES), 0, c->stream, n, X,
                       c->mlp_params, Y);
}

// one gradient evaluation over nb samples (X, T or X[idx], T[idx]); grad and loss on the device
void launch_mlp_gradient(mhpc_ctx *c, int64_t nb, const float *X, const float *T, const int32_t *idx,
                         float *grad, float *loss) {
    const int64_t nw = ceil_div(nb, 64);
    hipLaunchKernelGGL(k_mlp_train_grad, dim3((unsigned)ceil_div(nw, MLP_TR_WAVES)), dim3(64 * MLP_TR_WAVES), 0,
                       c->stream, nb, X, T, idx, c->mlp_params, c->mlp_partial, c->mlp_loss_partial);
    // stage-1 sums go behind the per-wave partials in the same buffer (sized for the
    // maximum wave count plus its chunk sums)
    const int nchunks = (int)ceil_div(nw, MLP_RED_CHUNK);
    float *part2 = c->mlp_partial + (int64_t)c->mlp_max_waves * MLP_NPARAM;
    float *loss2 = c->mlp_loss_partial + c->mlp_max_waves;
    hipLaunchKernelGGL(k_mlp_reduce1, dim3((unsigned)ceil_div(MLP_NPARAM, 256), (unsigned)nchunks), dim3(256), 0,
                       c->stream, (int)nw, c->mlp_partial, c->mlp_loss_partial, part2, loss2);
    hipLaunchKernelGGL(k_mlp_reduce2, dim3((unsigned)ceil_div(MLP_NPARAM, 256)), dim3(256), 0, c->stream, nchunks,
                       nb, part2, loss2, grad, loss);
}

void launch_mlp_adam(mhpc_ctx *c, float lr, int step) {
    const float bc1 = 1.f - powf(0.9f, (float)step), bc2 = 1.f - powf(0.999f, (float)step);
    hipLaunchKernelGGL(k_mlp_adam, dim3((unsigned)ceil_div(MLP_NPARAM, 256)), dim3(256), 0, c->stream,
                       c->mlp_params, c->mlp_m1, c->mlp_m2, c->mlp_grad, lr, bc1, bc2);
}

// training targets: random sample indices (hash of seed and slot) and
// T = (log(sml / h_previous) - mu_t) / sigma_t for every particle
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ __launch_bounds__(256) void k_mlp_targets(int64_t n, const double *__restrict__ h,
                                                     const double *__restrict__ hprev,
                                                     const float *__restrict__ norm, float *__restrict__ T) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) T[i] = (float)((log(h[i] / hprev[i]) - norm[16]) / norm[17]);
}

__global__ __launch_bounds__(256) void k_mlp_sample(int64_t n, int64_t ns, uint64_t seed, int32_t *__restrict__ idx) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < ns) idx[i] = (int32_t)(mix64(seed * 0x9E3779B97F4A7C15ull + (uint64_t)i) % (uint64_t)n);
}

void launch_mlp_targets(mhpc_ctx *c) {
    if (c->n <= 0) return;
    hipLaunchKernelGGL(k_mlp_targets, dim3((unsigned)ceil_div(c->n, 256)), dim3(256), 0, c->stream, c->n,
                       c->f[MHPC_F_SML], c->mlp_hprev, c->mlp_norm, c->mlp_t);
}

void launch_mlp_sample(mhpc_ctx *c, int64_t nsamples, uint64_t seed) {
    hipLaunchKernelGGL(k_mlp_sample, dim3((unsigned)ceil_div(nsamples, 256)), dim3(256), 0, c->stream, c->n,
                       nsamples, seed, c->mlp_idx);
}
