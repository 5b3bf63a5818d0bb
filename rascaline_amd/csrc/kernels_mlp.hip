// Learned radial-integral surrogate: training (FP32), output-layer solve (FP64)
// and inference (FP16 hi + lo operands on MFMA with FP32 accumulation).
//
// The network is 1 -> 64 -> 64 -> n_out with SiLU activations,
// silu(z) = z * sigmoid(z). Its inputs are pair distances mapped to [-1, 1];
// its outputs are R_nl(r) and dR_nl/dr for every (l, n), each divided by its
// largest magnitude over [0, cutoff).
#include "kernels_mlp.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "device.hpp"
#include "hypers.hpp"

namespace rascaline_amd {

namespace {

constexpr int H = MLP_HIDDEN;
constexpr int BLOCK = 256;

// splitmix64 finaliser (Steele, Lea and Flood, OOPSLA 2014)
__host__ __device__ inline unsigned long long mix64(unsigned long long z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__device__ inline float sigmoidf(float z) { return 1.0f / (1.0f + expf(-z)); }

__device__ inline double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

__device__ inline double wave_max(double v) {
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
    return v;
}

// ------------------------------------------------------------------- training

__global__ void __launch_bounds__(BLOCK)
mlp_train_kernel(int n_out, int ts, long long batch, long long n_data, const float* __restrict__ x,
                 const float* __restrict__ targets, unsigned long long seed, long long step,
                 const float* __restrict__ params, const float* __restrict__ w2t, const float* __restrict__ w3t,
                 float* __restrict__ partial, double* __restrict__ loss_partial) {
    extern __shared__ float smem[];
    float* xs = smem;          // [64]
    float* a1 = xs + 64;       // [ts][64] silu(z1)
    float* d1 = a1 + ts * H;   // [ts][64] silu'(z1), later dL/dz1
    float* a2 = d1 + ts * H;   // [ts][64] silu(z2)
    float* d2 = a2 + ts * H;   // [ts][64] silu'(z2), later dL/dz2
    float* dy = d2 + ts * H;   // [ts][n_out] dL/dy
    __shared__ long long src[64];
    __shared__ double red[BLOCK / 64];

    const int t = threadIdx.x;
    const long long P = mlp_n_params(n_out);
    const float* W1 = params;
    const float* b1 = W1 + H;
    const float* W2 = b1 + H;
    const float* b2 = W2 + H * H;
    const float* W3 = b2 + H;
    const float* b3 = W3 + (long long)n_out * H;
    float* gW1 = partial + (long long)blockIdx.x * P;
    float* gb1 = gW1 + H;
    float* gW2 = gb1 + H;
    float* gb2 = gW2 + H * H;
    float* gW3 = gb2 + H;
    float* gb3 = gW3 + (long long)n_out * H;

    if (t < ts) {
        const long long b = (long long)blockIdx.x * ts + t;
        long long s = -1;
        if (b < batch) {
            s = seed ? (long long)(mix64(seed ^ mix64((unsigned long long)(step * batch + b))) %
                                   (unsigned long long)n_data)
                     : b;
        }
        src[t] = s;
        xs[t] = s >= 0 ? x[s] : 0.0f;
    }
    __syncthreads();

    // layer 1
    for (int i = t; i < ts * H; i += BLOCK) {
        const int s = i >> 6, u = i & 63;
        const float z = W1[u] * xs[s] + b1[u];
        const float sg = sigmoidf(z);
        a1[i] = z * sg;
        d1[i] = sg * (1.0f + z * (1.0f - sg));
    }
    __syncthreads();
    // layer 2
    for (int i = t; i < ts * H; i += BLOCK) {
        const int s = i >> 6, u = i & 63;
        float z = b2[u];
        for (int k = 0; k < H; k++) z += w2t[k * H + u] * a1[s * H + k];
        const float sg = sigmoidf(z);
        a2[i] = z * sg;
        d2[i] = sg * (1.0f + z * (1.0f - sg));
    }
    __syncthreads();
    // layer 3 and the residuals
    const float scale = 2.0f / (float)((double)batch * n_out);
    double lsum = 0.0;
    for (int i = t; i < ts * n_out; i += BLOCK) {
        const int s = i / n_out, o = i - s * n_out;
        float y = b3[o];
        for (int k = 0; k < H; k++) y += w3t[(long long)k * n_out + o] * a2[s * H + k];
        float res = 0.0f;
        if (src[s] >= 0) {
            res = y - targets[src[s] * n_out + o];
            lsum += (double)res * res;
        }
        dy[i] = scale * res;
    }
    lsum = wave_sum(lsum);
    if ((t & 63) == 0) red[t >> 6] = lsum;
    __syncthreads();
    if (t == 0) {
        double total = 0.0;
        for (int w = 0; w < BLOCK / 64; w++) total += red[w];
        loss_partial[blockIdx.x] = total;
    }

    // output layer gradients
    for (int i = t; i < n_out * H; i += BLOCK) {
        const int o = i >> 6, k = i & 63;
        float acc = 0.0f;
        for (int s = 0; s < ts; s++) acc += dy[s * n_out + o] * a2[s * H + k];
        gW3[i] = acc;
    }
    for (int o = t; o < n_out; o += BLOCK) {
        float acc = 0.0f;
        for (int s = 0; s < ts; s++) acc += dy[s * n_out + o];
        gb3[o] = acc;
    }
    // dL/dz2 = (W3^T dy) * silu'(z2), in place of d2
    for (int i = t; i < ts * H; i += BLOCK) {
        const int s = i >> 6, k = i & 63;
        float acc = 0.0f;
        for (int o = 0; o < n_out; o++) acc += W3[(long long)o * H + k] * dy[s * n_out + o];
        d2[i] *= acc;
    }
    __syncthreads();
    // hidden layer gradients
    for (int i = t; i < H * H; i += BLOCK) {
        const int u = i >> 6, k = i & 63;
        float acc = 0.0f;
        for (int s = 0; s < ts; s++) acc += d2[s * H + u] * a1[s * H + k];
        gW2[i] = acc;
    }
    if (t < H) {
        float acc = 0.0f;
        for (int s = 0; s < ts; s++) acc += d2[s * H + t];
        gb2[t] = acc;
    }
    // dL/dz1 = (W2^T dz2) * silu'(z1), in place of d1
    for (int i = t; i < ts * H; i += BLOCK) {
        const int s = i >> 6, k = i & 63;
        float acc = 0.0f;
        for (int u = 0; u < H; u++) acc += W2[u * H + k] * d2[s * H + u];
        d1[i] *= acc;
    }
    __syncthreads();
    if (t < H) {
        float acc = 0.0f;
        for (int s = 0; s < ts; s++) acc += d1[s * H + t] * xs[s];
        gW1[t] = acc;
    } else if (t < 2 * H) {
        float acc = 0.0f;
        for (int s = 0; s < ts; s++) acc += d1[s * H + t - H];
        gb1[t - H] = acc;
    }
}

__device__ inline void write_transposed(int n_out, long long p, float w, float* w2t, float* w3t) {
    const long long w2 = 2 * H, w3 = 3 * H + H * H;
    if (p >= w2 && p < w2 + H * H) {
        const long long q = p - w2;
        w2t[(q % H) * H + q / H] = w;
    } else if (p >= w3 && p < w3 + (long long)n_out * H) {
        const long long q = p - w3;
        w3t[(q % H) * n_out + q / H] = w;
    }
}

__global__ void __launch_bounds__(BLOCK)
mlp_update_kernel(int n_out, int n_wg, const float* __restrict__ partial, const double* __restrict__ loss_partial,
                  long long batch, int adam, float step_size, float bias2, float* __restrict__ params,
                  float* __restrict__ m, float* __restrict__ v, float* __restrict__ w2t, float* __restrict__ w3t,
                  float* __restrict__ gradient, double* __restrict__ loss_out) {
    const long long P = mlp_n_params(n_out);
    const long long p = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (blockIdx.x == 0 && threadIdx.x == 0 && loss_out) {
        double l = 0.0;
        for (int w = 0; w < n_wg; w++) l += loss_partial[w];
        loss_out[0] = l / ((double)batch * n_out);
    }
    if (p >= P) return;
    float g = 0.0f;
    for (int w = 0; w < n_wg; w++) g += partial[(long long)w * P + p];
    if (!adam) {
        gradient[p] = g;
        return;
    }
    const float mm = 0.9f * m[p] + 0.1f * g;
    const float vv = 0.999f * v[p] + 0.001f * g * g;
    m[p] = mm;
    v[p] = vv;
    const float denom = sqrtf(vv / bias2) + 1e-8f;
    const float w = params[p] - step_size * mm / denom;
    params[p] = w;
    write_transposed(n_out, p, w, w2t, w3t);
}

__global__ void mlp_transpose_kernel(int n_out, const float* __restrict__ params, float* __restrict__ w2t,
                                     float* __restrict__ w3t) {
    const long long P = mlp_n_params(n_out);
    const long long p = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (p < P) write_transposed(n_out, p, params[p], w2t, w3t);
}

// ------------------------------------------------------------- output layer

__global__ void __launch_bounds__(BLOCK)
mlp_features_kernel(long long n, const float* __restrict__ x, const float* __restrict__ params,
                    const float* __restrict__ w2t, float* __restrict__ features) {
    __shared__ float xs[64];
    __shared__ float a1[64 * H];
    const int t = threadIdx.x;
    const long long s0 = (long long)blockIdx.x * 64;
    const float* W1 = params;
    const float* b1 = W1 + H;
    const float* b2 = b1 + H + H * H;
    if (t < 64) xs[t] = s0 + t < n ? x[s0 + t] : 0.0f;
    __syncthreads();
    for (int i = t; i < 64 * H; i += BLOCK) {
        const int s = i >> 6, u = i & 63;
        const float z = W1[u] * xs[s] + b1[u];
        a1[i] = z * sigmoidf(z);
    }
    __syncthreads();
    for (int i = t; i < 64 * H; i += BLOCK) {
        const int s = i >> 6, u = i & 63;
        if (s0 + s >= n) continue;
        float z = b2[u];
        for (int k = 0; k < H; k++) z += w2t[k * H + u] * a1[s * H + k];
        features[(s0 + s) * H + u] = z * sigmoidf(z);
    }
}

constexpr int GRAM_ROWS = H + 1;

__global__ void __launch_bounds__(BLOCK)
mlp_gram_kernel(long long n, int n_out, long long per_chunk, const float* __restrict__ features,
                const float* __restrict__ targets, double* __restrict__ partial) {
    __shared__ double As[64][GRAM_ROWS];
    const int t = threadIdx.x;
    const int C = GRAM_ROWS + n_out;
    const int j = blockIdx.x * BLOCK + t;
    const long long begin = (long long)blockIdx.y * per_chunk;
    const long long end = begin + per_chunk < n ? begin + per_chunk : n;
    double acc[GRAM_ROWS];
#pragma unroll
    for (int i = 0; i < GRAM_ROWS; i++) acc[i] = 0.0;
    for (long long base = begin; base < end; base += 64) {
        for (int idx = t; idx < 64 * GRAM_ROWS; idx += BLOCK) {
            const int s = idx / GRAM_ROWS, i = idx - s * GRAM_ROWS;
            const long long sg = base + s;
            double v = 0.0;
            if (sg < end) v = i < H ? (double)features[sg * H + i] : 1.0;
            As[s][i] = v;
        }
        __syncthreads();
        if (j < C) {
            for (int s = 0; s < 64; s++) {
                const long long sg = base + s;
                double cj = 0.0;
                if (sg < end) cj = j < GRAM_ROWS ? As[s][j] : (double)targets[sg * n_out + (j - GRAM_ROWS)];
#pragma unroll
                for (int i = 0; i < GRAM_ROWS; i++) acc[i] += As[s][i] * cj;
            }
        }
        __syncthreads();
    }
    if (j < C) {
        double* out = partial + (long long)blockIdx.y * GRAM_ROWS * C;
#pragma unroll
        for (int i = 0; i < GRAM_ROWS; i++) out[(long long)i * C + j] = acc[i];
    }
}

__global__ void mlp_gram_reduce_kernel(int n_out, int chunks, const double* __restrict__ partial,
                                       double* __restrict__ gram) {
    const long long E = (long long)GRAM_ROWS * (GRAM_ROWS + n_out);
    const long long e = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (e >= E) return;
    double s = 0.0;
    for (int c = 0; c < chunks; c++) s += partial[c * E + e];
    gram[e] = s;
}

__global__ void __launch_bounds__(BLOCK)
mlp_solve_kernel(int n_out, double ridge, const double* __restrict__ gram, float* __restrict__ params,
                 float* __restrict__ w3t, int* __restrict__ status) {
    __shared__ double G[GRAM_ROWS * GRAM_ROWS];
    __shared__ int failed;
    const int t = threadIdx.x;
    const int C = GRAM_ROWS + n_out;
    const int R = GRAM_ROWS;
    for (int idx = t; idx < R * R; idx += BLOCK) G[idx] = gram[(idx / R) * C + idx % R];
    if (t == 0) failed = 0;
    __syncthreads();
    if (t == 0) {
        double trace = 0.0;
        for (int i = 0; i < R; i++) trace += G[i * R + i];
        for (int i = 0; i < R; i++) G[i * R + i] += ridge * trace / R;
    }
    __syncthreads();
    // right-looking Cholesky, lower triangle in place
    for (int k = 0; k < R; k++) {
        if (t == 0) {
            const double d = G[k * R + k];
            if (!(d > 0.0)) failed = 1;
            G[k * R + k] = d > 0.0 ? sqrt(d) : 1.0;
        }
        __syncthreads();
        if (failed) break;
        for (int i = k + 1 + t; i < R; i += BLOCK) G[i * R + k] /= G[k * R + k];
        __syncthreads();
        const int m = R - k - 1;
        for (int idx = t; idx < m * m; idx += BLOCK) {
            const int i = k + 1 + idx / m, j = k + 1 + idx % m;
            if (j <= i) G[i * R + j] -= G[i * R + k] * G[j * R + k];
        }
        __syncthreads();
    }
    if (failed) {
        if (t == 0) status[0] = 1;
        return;
    }
    if (t == 0) status[0] = 0;
    float* W3 = params + 3 * H + H * H;
    float* b3 = W3 + (long long)n_out * H;
    for (int o = t; o < n_out; o += BLOCK) {
        double w[GRAM_ROWS];
        for (int i = 0; i < R; i++) {
            double s = gram[(long long)i * C + R + o];
            for (int k = 0; k < i; k++) s -= G[i * R + k] * w[k];
            w[i] = s / G[i * R + i];
        }
        for (int i = R - 1; i >= 0; i--) {
            double s = w[i];
            for (int k = i + 1; k < R; k++) s -= G[k * R + i] * w[k];
            w[i] = s / G[i * R + i];
        }
        for (int k = 0; k < H; k++) {
            W3[(long long)o * H + k] = (float)w[k];
            w3t[(long long)k * n_out + o] = (float)w[k];
        }
        b3[o] = (float)w[H];
    }
}

// ------------------------------------------------------------------ inference

typedef _Float16 half4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

constexpr float LO_SCALE = 4096.0f;  // lo parts are stored times 2^12
constexpr float LO_UNSCALE = 1.0f / 4096.0f;

__device__ __forceinline__ void split_f16(float x, _Float16& hi, _Float16& lo) {
    hi = (_Float16)x;
    lo = (_Float16)((x - (float)hi) * LO_SCALE);
}

// One wave evaluates 16 distances at a time. Every layer runs transposed
// (units x distances), so the accumulator of v_mfma_f32_16x16x16f16 (lane
// p + 16q holds rows 4q..4q+3 of column p) is already the B operand of the
// next layer (lane p + 16q holds rows 4q..4q+3 of column p).
template <typename T, bool VEC4>
__global__ void __launch_bounds__(BLOCK)
mlp_inference_kernel(int n_out, int n_pad, int split, double x_scale, double x_shift,
                     const uint16_t* __restrict__ frag2, const uint16_t* __restrict__ frag3,
                     const float* __restrict__ small, long long n, const double* __restrict__ r, T* __restrict__ out0,
                     T* __restrict__ out1) {
    extern __shared__ __align__(16) unsigned char raw[];
    const int NOB = n_pad / 16;
    half4_t* f2 = reinterpret_cast<half4_t*>(raw);                // [2][16][64]
    half4_t* f3 = f2 + 2 * 16 * 64;                               // [2][NOB * 4][64]
    float* sm = reinterpret_cast<float*>(f3 + 2 * NOB * 4 * 64);  // see MlpInference::small
    const int t = threadIdx.x;
    {
        const uint4* g2 = reinterpret_cast<const uint4*>(frag2);
        uint4* l2 = reinterpret_cast<uint4*>(f2);
        for (int i = t; i < 2 * 16 * 64 / 2; i += BLOCK) l2[i] = g2[i];
        const uint4* g3 = reinterpret_cast<const uint4*>(frag3);
        uint4* l3 = reinterpret_cast<uint4*>(f3);
        for (int i = t; i < 2 * NOB * 4 * 64 / 2; i += BLOCK) l3[i] = g3[i];
        for (int i = t; i < 3 * H + 2 * n_pad + 4; i += BLOCK) sm[i] = small[i];
    }
    __syncthreads();
    const float* W1 = sm;
    const float* b1 = sm + H;
    const float* b2 = sm + 2 * H;
    const float* b3 = sm + 3 * H;
    const float* oscale = b3 + n_pad;
    const float sa1 = oscale[n_pad], inv2 = oscale[n_pad + 1], sa2 = oscale[n_pad + 2], inv3 = oscale[n_pad + 3];

    const int lane = t & 63, wave = t >> 6;
    const int p = lane & 15, q = lane >> 4;
    const int n1 = n_out - split;
    const long long n_tiles = (n + 15) / 16;
    for (long long tile = (long long)blockIdx.x * (BLOCK / 64) + wave; tile < n_tiles;
         tile += (long long)gridDim.x * (BLOCK / 64)) {
        const long long e = tile * 16 + p;
        const bool valid = e < n;
        const float xin = valid ? (float)(r[e] * x_scale + x_shift) : 0.0f;

        half4_t bh[4], bl[4];
#pragma unroll
        for (int kb = 0; kb < 4; kb++) {
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int k = kb * 16 + 4 * q + j;
                const float z = W1[k] * xin + b1[k];
                _Float16 h, l;
                split_f16(z * sigmoidf(z) * sa1, h, l);
                bh[kb][j] = h;
                bl[kb][j] = l;
            }
        }
        half4_t ch[4], cl[4];
#pragma unroll
        for (int ib = 0; ib < 4; ib++) {
            float4_t hh = {0.0f, 0.0f, 0.0f, 0.0f}, corr = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int kb = 0; kb < 4; kb++) {
                const half4_t ah = f2[(ib * 4 + kb) * 64 + lane];
                const half4_t al = f2[(16 + ib * 4 + kb) * 64 + lane];
                hh = __builtin_amdgcn_mfma_f32_16x16x16f16(ah, bh[kb], hh, 0, 0, 0);
                corr = __builtin_amdgcn_mfma_f32_16x16x16f16(ah, bl[kb], corr, 0, 0, 0);
                corr = __builtin_amdgcn_mfma_f32_16x16x16f16(al, bh[kb], corr, 0, 0, 0);
            }
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const float z = (hh[j] + corr[j] * LO_UNSCALE) * inv2 + b2[ib * 16 + 4 * q + j];
                _Float16 h, l;
                split_f16(z * sigmoidf(z) * sa2, h, l);
                ch[ib][j] = h;
                cl[ib][j] = l;
            }
        }
        for (int ob = 0; ob < NOB; ob++) {
            float4_t hh = {0.0f, 0.0f, 0.0f, 0.0f}, corr = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int kb = 0; kb < 4; kb++) {
                const half4_t ah = f3[(ob * 4 + kb) * 64 + lane];
                const half4_t al = f3[(NOB * 4 + ob * 4 + kb) * 64 + lane];
                hh = __builtin_amdgcn_mfma_f32_16x16x16f16(ah, ch[kb], hh, 0, 0, 0);
                corr = __builtin_amdgcn_mfma_f32_16x16x16f16(ah, cl[kb], corr, 0, 0, 0);
                corr = __builtin_amdgcn_mfma_f32_16x16x16f16(al, ch[kb], corr, 0, 0, 0);
            }
            const int o0 = ob * 16 + 4 * q;
            if (!valid || o0 >= n_out) continue;
            float y[4];
#pragma unroll
            for (int j = 0; j < 4; j++) y[j] = ((hh[j] + corr[j] * LO_UNSCALE) * inv3 + b3[o0 + j]) * oscale[o0 + j];
            if constexpr (VEC4) {
                // split and n_out are multiples of 4: the four outputs share a destination
                float4_t v = {y[0], y[1], y[2], y[3]};
                if (o0 < split) {
                    *reinterpret_cast<float4_t*>(out0 + e * split + o0) = v;
                } else {
                    *reinterpret_cast<float4_t*>(out1 + e * n1 + (o0 - split)) = v;
                }
            } else {
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const int o = o0 + j;
                    if (o >= n_out) break;
                    if (o < split) {
                        out0[e * split + o] = (T)y[j];
                    } else {
                        out1[e * n1 + (o - split)] = (T)y[j];
                    }
                }
            }
        }
    }
}

// ---------------------------------------------------------------------- judge

__global__ void __launch_bounds__(BLOCK)
column_error_kernel(long long n, int cols, const float* __restrict__ a, const float* __restrict__ b,
                    double* __restrict__ out) {
    __shared__ double red[2][BLOCK / 64];
    const int c = blockIdx.x;
    double diff = 0.0, ref = 0.0;
    for (long long i = threadIdx.x; i < n; i += BLOCK) {
        const double va = a[i * cols + c], vb = b[i * cols + c];
        diff = fmax(diff, fabs(va - vb));
        ref = fmax(ref, fabs(vb));
    }
    diff = wave_max(diff);
    ref = wave_max(ref);
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = diff;
        red[1][threadIdx.x >> 6] = ref;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < BLOCK / 64; w++) {
            diff = fmax(diff, red[0][w]);
            ref = fmax(ref, red[1][w]);
        }
        out[c] = diff;
        out[cols + c] = ref;
    }
}

__global__ void __launch_bounds__(BLOCK)
column_maxabs_kernel(long long n, int cols, const double* __restrict__ values, double* __restrict__ out) {
    __shared__ double red[BLOCK / 64];
    const int c = blockIdx.x;
    double m = 0.0;
    for (long long i = threadIdx.x; i < n; i += BLOCK) m = fmax(m, fabs(values[i * cols + c]));
    m = wave_max(m);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < BLOCK / 64; w++) m = fmax(m, red[w]);
        out[c] = m;
    }
}

__global__ void normalise_targets_kernel(long long n, int cols, const double* __restrict__ values,
                                         const double* __restrict__ derivs, const double* __restrict__ scale,
                                         float* __restrict__ targets) {
    const long long idx = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (idx >= n * 2 * cols) return;
    const long long i = idx / (2 * cols);
    const int o = (int)(idx - i * 2 * cols);
    const double v = o < cols ? values[i * cols + o] : derivs[i * cols + o - cols];
    const double s = scale[o];
    targets[idx] = (float)(s > 0.0 ? v / s : 0.0);
}

__global__ void uniform_samples_kernel(long long n, double r_max, unsigned long long seed, double* __restrict__ r) {
    const long long i = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    const double u = (double)(mix64(seed ^ mix64((unsigned long long)i)) >> 11) * 0x1.0p-53;
    r[i] = r_max * ((double)i + u) / (double)n;
}

__global__ void network_inputs_kernel(long long n, const double* __restrict__ r, double x_scale, double x_shift,
                                      float* __restrict__ x) {
    const long long i = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) x[i] = (float)(r[i] * x_scale + x_shift);
}

unsigned blocks_for(long long n) { return (unsigned)((n + BLOCK - 1) / BLOCK); }

}  // namespace

// --------------------------------------------------------------------- launchers

int mlp_train_tile(int n_out) {
    for (int ts : {32, 16, 8}) {
        const size_t bytes = (64 + (size_t)ts * (4 * H + n_out)) * sizeof(float);
        if (bytes <= 60 * 1024) return ts;
    }
    return 4;
}

void launch_mlp_train(hipStream_t stream, int n_out, long long batch, long long n_data, const float* x,
                      const float* targets, unsigned long long seed, long long step, const float* params,
                      const float* w2t, const float* w3t, float* partial, double* loss_partial) {
    if (batch <= 0) return;
    const int ts = mlp_train_tile(n_out);
    const long long n_wg = (batch + ts - 1) / ts;
    const size_t shmem = (64 + (size_t)ts * (4 * H + n_out)) * sizeof(float);
    mlp_train_kernel<<<(unsigned)n_wg, BLOCK, shmem, stream>>>(n_out, ts, batch, n_data, x, targets, seed, step,
                                                               params, w2t, w3t, partial, loss_partial);
}

void launch_mlp_update(hipStream_t stream, int n_out, int n_wg, const float* partial, const double* loss_partial,
                       long long batch, int adam, float step_size, float bias2, float* params, float* m, float* v,
                       float* w2t, float* w3t, float* gradient, double* loss_out) {
    mlp_update_kernel<<<blocks_for(mlp_n_params(n_out)), BLOCK, 0, stream>>>(
        n_out, n_wg, partial, loss_partial, batch, adam, step_size, bias2, params, m, v, w2t, w3t, gradient, loss_out);
}

void launch_mlp_transpose(hipStream_t stream, int n_out, const float* params, float* w2t, float* w3t) {
    mlp_transpose_kernel<<<blocks_for(mlp_n_params(n_out)), BLOCK, 0, stream>>>(n_out, params, w2t, w3t);
}

void launch_mlp_features(hipStream_t stream, long long n, const float* x, const float* params, const float* w2t,
                         float* features) {
    if (n <= 0) return;
    mlp_features_kernel<<<(unsigned)((n + 63) / 64), BLOCK, 0, stream>>>(n, x, params, w2t, features);
}

int mlp_gram_chunks(long long n) {
    const long long tiles = (n + 63) / 64;
    return (int)(tiles < 128 ? (tiles > 0 ? tiles : 1) : 128);
}

void launch_mlp_gram(hipStream_t stream, long long n, int n_out, const float* features, const float* targets,
                     double* partial) {
    const int chunks = mlp_gram_chunks(n);
    const long long tiles = (n + 63) / 64;
    const long long per_chunk = ((tiles + chunks - 1) / chunks) * 64;
    const int C = GRAM_ROWS + n_out;
    dim3 grid((C + BLOCK - 1) / BLOCK, chunks);
    mlp_gram_kernel<<<grid, BLOCK, 0, stream>>>(n, n_out, per_chunk, features, targets, partial);
}

void launch_mlp_gram_reduce(hipStream_t stream, int n_out, int chunks, const double* partial, double* gram) {
    const long long E = (long long)GRAM_ROWS * (GRAM_ROWS + n_out);
    mlp_gram_reduce_kernel<<<blocks_for(E), BLOCK, 0, stream>>>(n_out, chunks, partial, gram);
}

void launch_mlp_solve(hipStream_t stream, int n_out, double ridge, double* gram, float* params, float* w3t,
                      int* status) {
    mlp_solve_kernel<<<1, BLOCK, 0, stream>>>(n_out, ridge, gram, params, w3t, status);
}

namespace {

// power of two p with max_abs * p in (2^12, 2^13]; 1 for an all-zero operand
float operand_scale(double max_abs) {
    if (!(max_abs > 0.0) || !std::isfinite(max_abs)) return 1.0f;
    return std::ldexp(1.0f, 13 - (int)std::ceil(std::log2(max_abs)));
}

double silu_host(double z) { return z / (1.0 + std::exp(-z)); }

}  // namespace

void mlp_pack_inference(int n_out, const float* params, const float* scale, double x_min, double x_max,
                        uint16_t* frag2, uint16_t* frag3, float* small) {
    const int n_pad = mlp_padded_outputs(n_out), NOB = n_pad / 16;
    const float* W1 = params;
    const float* b1 = W1 + H;
    const float* W2 = b1 + H;
    const float* b2 = W2 + H * H;
    const float* W3 = b2 + H;
    const float* b3 = W3 + (size_t)n_out * H;

    // largest activations over the input range (a dense grid; the network is a
    // smooth function of one variable), with a factor 2 of headroom
    double max_a1 = 0.0, max_a2 = 0.0, max_w2 = 0.0, max_w3 = 0.0;
    const int G = 4097;
    std::vector<double> a1(H);
    for (int g = 0; g < G; g++) {
        const double x = x_min + (x_max - x_min) * g / (G - 1);
        for (int u = 0; u < H; u++) {
            a1[u] = silu_host((double)W1[u] * x + b1[u]);
            max_a1 = std::max(max_a1, std::fabs(a1[u]));
        }
        for (int u = 0; u < H; u++) {
            double z = b2[u];
            for (int k = 0; k < H; k++) z += (double)W2[u * H + k] * a1[k];
            max_a2 = std::max(max_a2, std::fabs(silu_host(z)));
        }
    }
    for (int i = 0; i < H * H; i++) max_w2 = std::max(max_w2, std::fabs((double)W2[i]));
    for (size_t i = 0; i < (size_t)n_out * H; i++) max_w3 = std::max(max_w3, std::fabs((double)W3[i]));
    const float sa1 = operand_scale(2.0 * max_a1), sa2 = operand_scale(2.0 * max_a2);
    const float sw2 = operand_scale(max_w2), sw3 = operand_scale(max_w3);

    auto bits = [](_Float16 h) {
        uint16_t u;
        std::memcpy(&u, &h, sizeof(u));
        return u;
    };
    // fragment (ib, kb), lane l, element e holds A[ib*16 + l%16][kb*16 + 4*(l/16) + e]
    auto pack = [&](const float* W, float sw, int rows, int blocks, uint16_t* out) {
        const size_t half = (size_t)blocks * 4 * 64 * 4;
        for (int ib = 0; ib < blocks; ib++)
            for (int kb = 0; kb < 4; kb++)
                for (int l = 0; l < 64; l++)
                    for (int e = 0; e < 4; e++) {
                        const int row = ib * 16 + l % 16, col = kb * 16 + 4 * (l / 16) + e;
                        const float w = row < rows ? W[(size_t)row * H + col] * sw : 0.0f;
                        const _Float16 hi = (_Float16)w;
                        const _Float16 lo = (_Float16)((w - (float)hi) * LO_SCALE);
                        const size_t idx = (((size_t)ib * 4 + kb) * 64 + l) * 4 + e;
                        out[idx] = bits(hi);
                        out[half + idx] = bits(lo);
                    }
    };
    pack(W2, sw2, H, 4, frag2);
    pack(W3, sw3, n_out, NOB, frag3);
    for (int k = 0; k < H; k++) {
        small[k] = W1[k];
        small[H + k] = b1[k];
        small[2 * H + k] = b2[k];
    }
    for (int o = 0; o < n_pad; o++) {
        small[3 * H + o] = o < n_out ? b3[o] : 0.0f;
        small[3 * H + n_pad + o] = o < n_out ? scale[o] : 0.0f;
    }
    float* extra = small + 3 * H + 2 * n_pad;
    extra[0] = sa1;
    extra[1] = 1.0f / (sa1 * sw2);
    extra[2] = sa2;
    extra[3] = 1.0f / (sa2 * sw3);
}

size_t mlp_inference_shmem(int n_out) {
    const int n_pad = mlp_padded_outputs(n_out), NOB = n_pad / 16;
    return (size_t)2 * 16 * 64 * 8 + (size_t)2 * NOB * 4 * 64 * 8 + mlp_small_size(n_out) * sizeof(float);
}

template <typename T>
void launch_mlp_inference(hipStream_t stream, const MlpInference& net, long long n, const double* r, T* out0,
                          T* out1) {
    if (n <= 0) return;
    const int n_pad = mlp_padded_outputs(net.n_out);
    const long long tiles = (n + 15) / 16;
    long long grid = (tiles + 3) / 4;
    if (grid > 1024) grid = 1024;
    const size_t shmem = mlp_inference_shmem(net.n_out);
    const bool vec4 = net.split % 4 == 0 && net.n_out % 4 == 0;
    if constexpr (sizeof(T) == 4) {
        if (vec4) {
            mlp_inference_kernel<T, true><<<(unsigned)grid, BLOCK, shmem, stream>>>(
                net.n_out, n_pad, net.split, net.x_scale, net.x_shift, net.frag2, net.frag3, net.small, n, r, out0,
                out1);
            return;
        }
    }
    {
        mlp_inference_kernel<T, false><<<(unsigned)grid, BLOCK, shmem, stream>>>(
            net.n_out, n_pad, net.split, net.x_scale, net.x_shift, net.frag2, net.frag3, net.small, n, r, out0, out1);
    }
}
template void launch_mlp_inference<float>(hipStream_t, const MlpInference&, long long, const double*, float*, float*);
template void launch_mlp_inference<double>(hipStream_t, const MlpInference&, long long, const double*, double*,
                                           double*);

void launch_column_error(hipStream_t stream, long long n, int cols, const float* a, const float* b, double* out) {
    if (cols <= 0) return;
    column_error_kernel<<<cols, BLOCK, 0, stream>>>(n, cols, a, b, out);
}

void launch_column_maxabs(hipStream_t stream, long long n, int cols, const double* values, double* out) {
    if (cols <= 0) return;
    column_maxabs_kernel<<<cols, BLOCK, 0, stream>>>(n, cols, values, out);
}

void launch_normalise_targets(hipStream_t stream, long long n, int cols, const double* values, const double* derivs,
                              const double* scale, float* targets) {
    if (n <= 0) return;
    normalise_targets_kernel<<<blocks_for(n * 2 * cols), BLOCK, 0, stream>>>(n, cols, values, derivs, scale, targets);
}

void launch_uniform_samples(hipStream_t stream, long long n, double r_max, unsigned long long seed, double* r) {
    if (n <= 0) return;
    uniform_samples_kernel<<<blocks_for(n), BLOCK, 0, stream>>>(n, r_max, seed, r);
}

void launch_network_inputs(hipStream_t stream, long long n, const double* r, double x_scale, double x_shift, float* x) {
    if (n <= 0) return;
    network_inputs_kernel<<<blocks_for(n), BLOCK, 0, stream>>>(n, r, x_scale, x_shift, x);
}

// ---- allocation-free training step (featomic_amd_mlp_loss_gradient_noalloc)
//
// The scratch of one loss-and-gradient pass (transposed weights, one partial
// gradient and one partial loss per tile of samples) lives in this file-scope
// device array instead of device memory allocated per call. Every GPU the library
// runs on holds its own copy until the process ends; calls that overlap on one GPU
// share it and race.
constexpr size_t MLP_SCRATCH_BYTES = size_t(32) << 20;
static __device__ __align__(256) unsigned char scratch[MLP_SCRATCH_BYTES];

constexpr int MLP_SCRATCH_MAX_DEVICES = 64;

// The device address of `scratch` on the current device, looked up once per device
// (the array's name in host code is a stand-in, not a device address).
static unsigned char* scratch_address() {
    static std::mutex mutex;
    static void* address[MLP_SCRATCH_MAX_DEVICES] = {};
    int device = 0;
    RA_HIP_CHECK(hipGetDevice(&device));
    if (device < 0 || device >= MLP_SCRATCH_MAX_DEVICES) throw Error::internal("unexpected device index");
    std::lock_guard<std::mutex> lock(mutex);
    if (!address[device]) {
        void* ptr = nullptr;
        const hipError_t status = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (status != hipSuccess || !ptr) {
            throw Error::internal(std::string("could not find the scratch array on the device: ") + hipGetErrorString(status));
        }
        address[device] = ptr;
    }
    return static_cast<unsigned char*>(address[device]);
}

size_t mlp_noalloc_scratch_bytes() { return MLP_SCRATCH_BYTES; }

// scratch bytes of one pass: w2t [H][H], w3t [H][n_out], partial [tiles][P] floats, then
// the partial losses [tiles] doubles (8-byte aligned)
static size_t noalloc_layout(int n_out, long long n_samples, size_t* partial_offset, size_t* loss_offset) {
    const long long P = mlp_n_params(n_out);
    const int ts = mlp_train_tile(n_out);
    const long long tiles = (n_samples + ts - 1) / ts;
    const size_t floats = (size_t)H * H + (size_t)H * n_out;
    *partial_offset = floats * sizeof(float);
    const size_t after = *partial_offset + (size_t)tiles * P * sizeof(float);
    *loss_offset = (after + 7) / 8 * 8;
    return *loss_offset + (size_t)tiles * sizeof(double);
}

long long mlp_noalloc_max_samples(int n_out) {
    long long lo = 0, hi = 1;
    size_t a, b;
    while (noalloc_layout(n_out, hi, &a, &b) <= MLP_SCRATCH_BYTES) hi *= 2;
    while (hi - lo > 1) {
        const long long mid = (lo + hi) / 2;
        if (noalloc_layout(n_out, mid, &a, &b) <= MLP_SCRATCH_BYTES) {
            lo = mid;
        } else {
            hi = mid;
        }
    }
    return lo;
}

void mlp_loss_gradient_noalloc(int n_out, long long n_samples, const float* params, const float* inputs,
                               const float* targets, double* loss, float* gradient, hipStream_t stream) {
    if (n_out < 1 || n_out > MLP_MAX_OUTPUTS) throw Error::invalid_parameter("n_out must be in [1, 512]");
    if (n_samples < 1) throw Error::invalid_parameter("n_samples must be positive");
    size_t partial_offset = 0, loss_offset = 0;
    const size_t bytes = noalloc_layout(n_out, n_samples, &partial_offset, &loss_offset);
    if (bytes > MLP_SCRATCH_BYTES) {
        throw Error::invalid_parameter("featomic_amd_mlp_loss_gradient_noalloc takes at most " +
                                       std::to_string(mlp_noalloc_max_samples(n_out)) + " samples for " +
                                       std::to_string(n_out) + " outputs (its scratch is " +
                                       std::to_string(MLP_SCRATCH_BYTES >> 20) + " MiB), got " +
                                       std::to_string(n_samples));
    }
    unsigned char* base = scratch_address();
    float* w2t = reinterpret_cast<float*>(base);
    float* w3t = w2t + H * H;
    float* partial = reinterpret_cast<float*>(base + partial_offset);
    double* loss_partial = reinterpret_cast<double*>(base + loss_offset);
    const int ts = mlp_train_tile(n_out);
    const int tiles = static_cast<int>((n_samples + ts - 1) / ts);
    // every word the passes read is written first: the transposed weights by the
    // transpose kernel, each tile's partial gradient and loss by its training workgroup
    launch_mlp_transpose(stream, n_out, params, w2t, w3t);
    launch_mlp_train(stream, n_out, n_samples, n_samples, inputs, targets, 0, 0, params, w2t, w3t, partial,
                     loss_partial);
    launch_mlp_update(stream, n_out, tiles, partial, loss_partial, n_samples, 0, 0.0f, 1.0f, nullptr, nullptr, nullptr,
                      nullptr, nullptr, gradient, loss);
    RA_HIP_CHECK(hipGetLastError());
}

}  // namespace rascaline_amd
