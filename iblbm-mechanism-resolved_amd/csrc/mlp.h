// mlp.h -- the learned MDF force predictor: a 3-layer MLP (8 -> 64 -> 64 -> 2, tanh).
//
// Per marker it maps (change of the forcing slip q = R (ud - U) before forcing since the
// previous step / sig, the direction of q, marker direction and sub-cell offset) to the change
// of the marker's converged MDF force from the previous step / sig; sig, the largest change of
// q over the wavefront's markers, makes the prediction exact in scale.  Training is FP32 on the
// device (forward, backward and Adam in one workgroup, summation orders fixed); inference runs
// in the window workgroup on the MFMA units (BF16 operands, FP32 accumulation).
//
// Parameter vector (FP32, MLP_NP entries): W1[MLP_IN][MLP_H], b1[MLP_H], W2[MLP_H][MLP_H],
// b2[MLP_H], W3[MLP_H][MLP_OUT], b3[MLP_OUT]; a layer computes y_n = b_n + sum_k x_k W[k][n].
// Normalisation (MLP_NORM floats): feature mean[8], feature scale[8], target scale[2].
#pragma once
#include <hip/hip_runtime.h>

namespace iblbm {

constexpr int MLP_IN = 8, MLP_H = 64, MLP_OUT = 2;
constexpr int MLP_W1 = 0, MLP_B1 = MLP_W1 + MLP_IN * MLP_H, MLP_W2 = MLP_B1 + MLP_H,
              MLP_B2 = MLP_W2 + MLP_H * MLP_H, MLP_W3 = MLP_B2 + MLP_H,
              MLP_B3 = MLP_W3 + MLP_H * MLP_OUT, MLP_NP = MLP_B3 + MLP_OUT;
constexpr int MLP_NORM = 2 * MLP_IN + MLP_OUT;
constexpr int MLP_ROW = MLP_IN + MLP_OUT;  // one training sample: features, target
// BF16 weight blocks in MFMA B-operand order (64 lanes x 4 values each): layer 1 (K padded
// to 16) 4 blocks, layer 2 16 blocks, layer 3 (N padded to 16) 4 blocks
constexpr int MLP_BLK1 = 0, MLP_BLK2 = 4, MLP_BLK3 = 20, MLP_NBLK = 24;

typedef short mlp_short4 __attribute__((ext_vector_type(4)));
typedef float mlp_float4 __attribute__((ext_vector_type(4)));
typedef __bf16 mlp_bf16x4 __attribute__((ext_vector_type(4)));

// Four floats to BF16 (round to nearest even) with gfx950's v_cvt_pk_bf16_f32.
__device__ inline mlp_short4 mlp_pack_bf16(float a, float b, float c, float d) {
    const mlp_bf16x4 v = {(__bf16)a, (__bf16)b, (__bf16)c, (__bf16)d};
    return __builtin_bit_cast(mlp_short4, v);
}

__host__ __device__ inline unsigned short f32_to_bf16(float f) {  // round to nearest even
    union {
        float f;
        unsigned u;
    } v;
    v.f = f;
    if ((v.u & 0x7F800000u) == 0x7F800000u && (v.u & 0x007FFFFFu))  // NaN stays NaN
        return (unsigned short)((v.u >> 16) | 0x40u);
    const unsigned r = v.u + 0x7FFFu + ((v.u >> 16) & 1u);
    return (unsigned short)(r >> 16);
}
__host__ __device__ inline float bf16_to_f32(unsigned short h) {
    union {
        unsigned u;
        float f;
    } v;
    v.u = (unsigned)h << 16;
    return v.f;
}

// B-operand block (kb, nb) of a K x N row-major FP32 matrix: lane l holds
// B[16 kb + 4 (l / 16) + j][16 nb + l % 16], j = 0..3 (zero outside K x N).
__host__ __device__ inline void mlp_pack_block(const float *W, int K, int N, int kb, int nb,
                                               unsigned short *blk /* 256 */) {
    for (int l = 0; l < 64; ++l)
        for (int j = 0; j < 4; ++j) {
            const int k = 16 * kb + 4 * (l / 16) + j, n = 16 * nb + l % 16;
            blk[4 * l + j] = (k < K && n < N) ? f32_to_bf16(W[k * N + n]) : (unsigned short)0;
        }
}

// tanh for activations that are rounded to BF16 next (8 significant bits): 1 - 2 / (e^{2x} + 1)
// with the hardware exp2 and reciprocal, relative error about 1e-6 where |tanh| > 1e-3 and
// absolute error below 1e-7 elsewhere; saturates to +-1 for large |x|.
__device__ inline float mlp_tanh_bf16(float x) {
    const float e = __builtin_amdgcn_exp2f(2.8853900817779268f * x);  // e^{2x}
    return 1.f - 2.f * __builtin_amdgcn_rcpf(e + 1.f);
}

// LDS writes of a wavefront made visible to its other lanes.
__device__ inline void mlp_wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Inference of one 16-row tile on one wavefront, as transposed products H^T = W^T X^T.
// With v_mfma_f32_16x16x16bf16_1k lane l = 16 g + r holds A[r][4 g + j], B[4 g + j][r] and
// D[4 g + j][r] (j = 0..3).  Taking the weights as the A operand (a packed block holds
// W[16 kb + 4 g + j][16 nb + r], which is W^T[16 nb + r][16 kb + 4 g + j]) and the activations
// as the B operand, a layer's output block (hidden 4 g + j of block ib, tile row r) is in
// exactly the lane and slot the next layer's B operand needs: the hidden activations never
// leave registers.  x16: 16 rows x 16 bf16 (features, then zeros) in LDS; y16: 16 rows x
// MLP_OUT floats.  Every lane of the wavefront must call it.
__device__ inline void mlp_tile16(const mlp_short4 *__restrict__ wblk,
                                  const float *__restrict__ b1, const float *__restrict__ b2,
                                  const float *__restrict__ b3, const unsigned short *x16,
                                  float *y16) {
    const int l = threadIdx.x & 63, r = l % 16, g = l / 16;
    const mlp_short4 xb = *reinterpret_cast<const mlp_short4 *>(x16 + r * 16 + 4 * g);
    mlp_short4 h1[MLP_H / 16], h2[MLP_H / 16];
#pragma unroll
    for (int ib = 0; ib < MLP_H / 16; ++ib) {  // layer 1
        mlp_float4 c = {0.f, 0.f, 0.f, 0.f};
        c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(wblk[(MLP_BLK1 + ib) * 64 + l], xb, c, 0,
                                                       0, 0);
        const mlp_float4 bb = *reinterpret_cast<const mlp_float4 *>(b1 + 16 * ib + 4 * g);
        h1[ib] = mlp_pack_bf16(mlp_tanh_bf16(c[0] + bb[0]), mlp_tanh_bf16(c[1] + bb[1]),
                               mlp_tanh_bf16(c[2] + bb[2]), mlp_tanh_bf16(c[3] + bb[3]));
    }
#pragma unroll
    for (int ib = 0; ib < MLP_H / 16; ++ib) {  // layer 2
        mlp_float4 c = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kb = 0; kb < MLP_H / 16; ++kb)
            c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(wblk[(MLP_BLK2 + kb * 4 + ib) * 64 + l],
                                                           h1[kb], c, 0, 0, 0);
        const mlp_float4 bb = *reinterpret_cast<const mlp_float4 *>(b2 + 16 * ib + 4 * g);
        h2[ib] = mlp_pack_bf16(mlp_tanh_bf16(c[0] + bb[0]), mlp_tanh_bf16(c[1] + bb[1]),
                               mlp_tanh_bf16(c[2] + bb[2]), mlp_tanh_bf16(c[3] + bb[3]));
    }
    mlp_float4 c = {0.f, 0.f, 0.f, 0.f};  // layer 3: outputs 0, 1 in lanes g = 0
#pragma unroll
    for (int kb = 0; kb < MLP_H / 16; ++kb)
        c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(wblk[(MLP_BLK3 + kb) * 64 + l], h2[kb], c,
                                                       0, 0, 0);
    if (g == 0) {
#pragma unroll
        for (int o = 0; o < MLP_OUT; ++o) y16[r * MLP_OUT + o] = c[o] + b3[o];
    }
}

// bf16 feature row of the tiles: MLP_IN normalised features, then zeros up to 16.
__device__ inline void mlp_put_row(unsigned short *row16, const float *xn) {
    static_assert(MLP_IN == 8, "two packed quads of features");
    mlp_short4 *r = reinterpret_cast<mlp_short4 *>(row16);
    r[0] = mlp_pack_bf16(xn[0], xn[1], xn[2], xn[3]);
    r[1] = mlp_pack_bf16(xn[4], xn[5], xn[6], xn[7]);
    r[2] = mlp_short4{0, 0, 0, 0};
    r[3] = mlp_short4{0, 0, 0, 0};
}

// LDS copy of the inference weights: the BF16 blocks and the three biases.
struct MlpLds {
    alignas(16) mlp_short4 blk[MLP_NBLK * 64];
    alignas(16) float bias[2 * MLP_H + 4];  // b1, b2, b3 (b3 lands with 2 trailing floats)
};

// ------------------------------------------------------------------ FP32 training

constexpr int MLP_TT = 1024;  // threads of the training workgroup
constexpr int MLP_B = 128;    // minibatch rows

// LDS of the training workgroup
struct MlpTrainShared {
    float p[MLP_NP];               // parameters
    float x[MLP_B * MLP_IN];       // normalised features of the batch
    float y[MLP_B * MLP_OUT];      // normalised targets of the batch
    float h1[MLP_B * MLP_H];       // layer-1 activations
    float h2[MLP_B * MLP_H];       // layer-2 activations
    float d2[MLP_B * MLP_H];       // dL/d(layer-2 pre-activation)
    float d3[MLP_B * MLP_OUT];     // dL/d(output)
    float red[MLP_TT / 64];
};

__device__ inline float mlp_wave_sum(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Forward + backward of one batch held in S.x / S.y (nb rows).  g: gradient (MLP_NP floats,
// global).  Returns the loss mean((y_hat - y)^2) over rows and outputs (valid in all threads).
// Every sum runs in a fixed order (thread-private loops over the batch or the layer width,
// then a fixed shuffle tree), so the result does not depend on timing.
__device__ inline float mlp_forward_backward(MlpTrainShared &S, int nb, float *g) {
    const int t = threadIdx.x;
    const float *p = S.p;
    // layer 1 and 2 forward
    for (int e = t; e < nb * MLP_H; e += MLP_TT) {
        const int b = e / MLP_H, n = e % MLP_H;
        float s = p[MLP_B1 + n];
        for (int k = 0; k < MLP_IN; ++k) s += S.x[b * MLP_IN + k] * p[MLP_W1 + k * MLP_H + n];
        S.h1[e] = tanhf(s);
    }
    __syncthreads();
    for (int e = t; e < nb * MLP_H; e += MLP_TT) {
        const int b = e / MLP_H, n = e % MLP_H;
        float s = p[MLP_B2 + n];
        for (int k = 0; k < MLP_H; ++k) s += S.h1[b * MLP_H + k] * p[MLP_W2 + k * MLP_H + n];
        S.h2[e] = tanhf(s);
    }
    __syncthreads();
    // output, loss and dL/dy_hat = 2 (y_hat - y) / (nb * OUT)
    float lsum = 0.f;
    const float inv = 1.f / (float)(nb * MLP_OUT);
    for (int e = t; e < nb * MLP_OUT; e += MLP_TT) {
        const int b = e / MLP_OUT, o = e % MLP_OUT;
        float s = p[MLP_B3 + o];
        for (int k = 0; k < MLP_H; ++k) s += S.h2[b * MLP_H + k] * p[MLP_W3 + k * MLP_OUT + o];
        const float r = s - S.y[e];
        lsum += r * r;
        S.d3[e] = 2.f * r * inv;
    }
    lsum = mlp_wave_sum(lsum);
    if ((t & 63) == 0) S.red[t >> 6] = lsum;
    __syncthreads();
    float loss = 0.f;
    for (int w = 0; w < MLP_TT / 64; ++w) loss += S.red[w];
    loss *= inv;
    // dL/d(layer-2 pre-activation) = (d3 W3^T) (1 - h2^2)
    for (int e = t; e < nb * MLP_H; e += MLP_TT) {
        const int b = e / MLP_H, k = e % MLP_H;
        float s = 0.f;
        for (int o = 0; o < MLP_OUT; ++o) s += S.d3[b * MLP_OUT + o] * p[MLP_W3 + k * MLP_OUT + o];
        const float h = S.h2[e];
        S.d2[e] = s * (1.f - h * h);
    }
    __syncthreads();
    // gradients of W3, b3, W2, b2: one thread per parameter, sum over the batch in order
    for (int q = t; q < MLP_H * MLP_OUT + MLP_OUT; q += MLP_TT) {
        float s = 0.f;
        if (q < MLP_H * MLP_OUT) {
            const int k = q / MLP_OUT, o = q % MLP_OUT;
            for (int b = 0; b < nb; ++b) s += S.h2[b * MLP_H + k] * S.d3[b * MLP_OUT + o];
            g[MLP_W3 + q] = s;
        } else {
            const int o = q - MLP_H * MLP_OUT;
            for (int b = 0; b < nb; ++b) s += S.d3[b * MLP_OUT + o];
            g[MLP_B3 + o] = s;
        }
    }
    for (int q = t; q < MLP_H * MLP_H + MLP_H; q += MLP_TT) {
        float s = 0.f;
        if (q < MLP_H * MLP_H) {
            const int k = q / MLP_H, n = q % MLP_H;
            for (int b = 0; b < nb; ++b) s += S.h1[b * MLP_H + k] * S.d2[b * MLP_H + n];
            g[MLP_W2 + q] = s;
        } else {
            const int n = q - MLP_H * MLP_H;
            for (int b = 0; b < nb; ++b) s += S.d2[b * MLP_H + n];
            g[MLP_B2 + n] = s;
        }
    }
    __syncthreads();
    // dL/d(layer-1 pre-activation) = (d2 W2^T) (1 - h1^2), kept in h1's place
    for (int e = t; e < nb * MLP_H; e += MLP_TT) {
        const int b = e / MLP_H, k = e % MLP_H;
        float s = 0.f;
        for (int n = 0; n < MLP_H; ++n) s += S.d2[b * MLP_H + n] * p[MLP_W2 + k * MLP_H + n];
        const float h = S.h1[e];
        S.h2[e] = s * (1.f - h * h);  // h2 is free now: d1 goes there
    }
    __syncthreads();
    for (int q = t; q < MLP_IN * MLP_H + MLP_H; q += MLP_TT) {
        float s = 0.f;
        if (q < MLP_IN * MLP_H) {
            const int k = q / MLP_H, n = q % MLP_H;
            for (int b = 0; b < nb; ++b) s += S.x[b * MLP_IN + k] * S.h2[b * MLP_H + n];
            g[MLP_W1 + q] = s;
        } else {
            const int n = q - MLP_IN * MLP_H;
            for (int b = 0; b < nb; ++b) s += S.h2[b * MLP_H + n];
            g[MLP_B1 + n] = s;
        }
    }
    __threadfence_block();
    __syncthreads();
    return loss;
}

// ------------------------------------------------------------------ kernels

// Normalisation from n samples: feature mean and standard deviation (1 where a feature is
// constant), target root-mean-square (1 where zero).  FP64 sums in a fixed order.
static __global__ __launch_bounds__(MLP_TT) void k_mlp_norm(const float *__restrict__ data,
                                                    long long n, float *__restrict__ norm) {
    __shared__ double red[MLP_TT / 64][MLP_ROW * 2];
    const int t = threadIdx.x;
    double s1[MLP_ROW], s2[MLP_ROW];
    for (int q = 0; q < MLP_ROW; ++q) s1[q] = s2[q] = 0.0;
    for (long long r = t; r < n; r += MLP_TT)
        for (int q = 0; q < MLP_ROW; ++q) {
            const double v = data[r * MLP_ROW + q];
            s1[q] += v;
            s2[q] += v * v;
        }
    for (int q = 0; q < MLP_ROW; ++q) {
        double a = s1[q], b = s2[q];
        for (int o = 32; o > 0; o >>= 1) {
            a += __shfl_xor(a, o, 64);
            b += __shfl_xor(b, o, 64);
        }
        if ((t & 63) == 0) {
            red[t >> 6][q] = a;
            red[t >> 6][MLP_ROW + q] = b;
        }
    }
    __syncthreads();
    if (t < MLP_ROW) {
        double a = 0.0, b = 0.0;
        for (int w = 0; w < MLP_TT / 64; ++w) {
            a += red[w][t];
            b += red[w][MLP_ROW + t];
        }
        const double nn = n > 0 ? (double)n : 1.0;
        const double mean = a / nn, var = b / nn - mean * mean;
        if (t < MLP_IN) {
            const double sd = var > 1e-24 ? sqrt(var) : 1.0;
            norm[t] = (float)mean;
            norm[MLP_IN + t] = (float)sd;
        } else {
            const double rms = sqrt(b / nn);
            norm[2 * MLP_IN + (t - MLP_IN)] = rms > 1e-30 ? (float)rms : 1.f;
        }
    }
}

// A reproducible row index for minibatch row b of training step s.
__device__ inline long long mlp_pick(unsigned seed, int s, int b, long long n) {
    unsigned long long z = ((unsigned long long)seed << 32) ^ ((unsigned long long)s * 1024u + b);
    z += 0x9E3779B97F4A7C15ull;  // splitmix64
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (long long)(z % (unsigned long long)n);
}

// `steps` Adam steps (beta1 0.9, beta2 0.999, eps 1e-8, bias-corrected; t0 steps done before)
// on minibatches of MLP_B samples drawn from the n stored ones.  One workgroup; losses[s]
// receives the minibatch loss of step s.
static __global__ __launch_bounds__(MLP_TT) void k_mlp_train(float *__restrict__ p, float *__restrict__ m,
                                                     float *__restrict__ v, float *__restrict__ g,
                                                     const float *__restrict__ norm,
                                                     const float *__restrict__ data, long long n,
                                                     int steps, int t0, float lr, unsigned seed,
                                                     float *__restrict__ losses) {
    __shared__ MlpTrainShared S;
    const int t = threadIdx.x;
    for (int q = t; q < MLP_NP; q += MLP_TT) S.p[q] = p[q];
    for (int s = 0; s < steps; ++s) {
        for (int e = t; e < MLP_B * MLP_ROW; e += MLP_TT) {
            const int b = e / MLP_ROW, q = e % MLP_ROW;
            const long long r = mlp_pick(seed, t0 + s, b, n);
            const float val = data[r * MLP_ROW + q];
            if (q < MLP_IN)
                S.x[b * MLP_IN + q] = (val - norm[q]) / norm[MLP_IN + q];
            else
                S.y[b * MLP_OUT + (q - MLP_IN)] = val / norm[2 * MLP_IN + (q - MLP_IN)];
        }
        __syncthreads();
        const float loss = mlp_forward_backward(S, MLP_B, g);
        const int tt = t0 + s + 1;
        const float c1 = 1.f - powf(0.9f, (float)tt), c2 = 1.f - powf(0.999f, (float)tt);
        for (int q = t; q < MLP_NP; q += MLP_TT) {
            const float gq = g[q];
            const float mq = 0.9f * m[q] + 0.1f * gq;
            const float vq = 0.999f * v[q] + 0.001f * gq * gq;
            m[q] = mq;
            v[q] = vq;
            S.p[q] = S.p[q] - lr * (mq / c1) / (sqrtf(vq / c2) + 1e-8f);
        }
        if (t == 0) losses[s] = loss;
        __threadfence_block();
        __syncthreads();
    }
    for (int q = t; q < MLP_NP; q += MLP_TT) p[q] = S.p[q];
}

// Loss and gradient of one batch (nb <= MLP_B rows of normalised x and y).
static __global__ __launch_bounds__(MLP_TT) void k_mlp_grad(const float *__restrict__ p,
                                                    const float *__restrict__ x,
                                                    const float *__restrict__ y, int nb,
                                                    float *__restrict__ g, float *__restrict__ loss) {
    __shared__ MlpTrainShared S;
    const int t = threadIdx.x;
    for (int q = t; q < MLP_NP; q += MLP_TT) S.p[q] = p[q];
    for (int e = t; e < nb * MLP_IN; e += MLP_TT) S.x[e] = x[e];
    for (int e = t; e < nb * MLP_OUT; e += MLP_TT) S.y[e] = y[e];
    __syncthreads();
    const float l = mlp_forward_backward(S, nb, g);
    if (t == 0) *loss = l;
}

// BF16 weight blocks of a parameter vector.
static __global__ void k_mlp_pack(const float *__restrict__ p, unsigned short *__restrict__ wb) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= MLP_NBLK) return;
    if (b < MLP_BLK2)
        mlp_pack_block(p + MLP_W1, MLP_IN, MLP_H, 0, b - MLP_BLK1, wb + 256 * b);
    else if (b < MLP_BLK3)
        mlp_pack_block(p + MLP_W2, MLP_H, MLP_H, (b - MLP_BLK2) / 4, (b - MLP_BLK2) % 4,
                       wb + 256 * b);
    else
        mlp_pack_block(p + MLP_W3, MLP_H, MLP_OUT, b - MLP_BLK3, 0, wb + 256 * b);
}

// BF16 MFMA inference of n rows of normalised features (64 rows per workgroup, four tiles).
static __global__ __launch_bounds__(64) void k_mlp_infer(const unsigned short *__restrict__ wb,
                                                 const float *__restrict__ p,
                                                 const float *__restrict__ x, int n,
                                                 float *__restrict__ out) {
    __shared__ alignas(16) unsigned short xs[64 * 16];
    __shared__ float ys[64 * MLP_OUT];
    const int row = blockIdx.x * 64 + threadIdx.x;
    float xi[MLP_IN];
#pragma unroll
    for (int q = 0; q < MLP_IN; ++q) xi[q] = row < n ? x[row * MLP_IN + q] : 0.f;
    mlp_put_row(xs + threadIdx.x * 16, xi);
    mlp_wave_sync();
    for (int tile = 0; tile < 4; ++tile)
        mlp_tile16(reinterpret_cast<const mlp_short4 *>(wb), p + MLP_B1, p + MLP_B2, p + MLP_B3,
                   xs + 256 * tile, ys + 16 * MLP_OUT * tile);
    mlp_wave_sync();
    if (row < n) {
#pragma unroll
        for (int o = 0; o < MLP_OUT; ++o) out[row * MLP_OUT + o] = ys[threadIdx.x * MLP_OUT + o];
    }
}

}  // namespace iblbm
