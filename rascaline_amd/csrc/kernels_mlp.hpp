// Kernels of the learned radial-integral surrogate: a 1 -> 64 -> 64 -> n_out
// SiLU network that maps a pair distance to every R_nl(r) and dR_nl/dr at once.
//
// Flat FP32 parameter layout (shared with the training, inference and the
// host): W1[64], b1[64], W2[64][64] (row = output unit), b2[64],
// W3[n_out][64], b3[n_out].
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace rascaline_amd {

constexpr int MLP_HIDDEN = 64;
constexpr int MLP_MAX_OUTPUTS = 512;

__host__ __device__ inline long long mlp_n_params(int n_out) {
    return 3LL * MLP_HIDDEN + (long long)MLP_HIDDEN * MLP_HIDDEN + (long long)n_out * MLP_HIDDEN + n_out;
}
inline int mlp_padded_outputs(int n_out) { return (n_out + 15) / 16 * 16; }

// ---- training (FP32)

// Samples per workgroup of the training kernel for this output count.
int mlp_train_tile(int n_out);

// One forward + backward pass over a batch. Sample b of the batch reads
// data[index(b)]: index(b) = b when `seed` is 0, else a hash of (seed, step, b)
// modulo n_data. Each workgroup writes the gradient of
// sum((y - t)^2) / (batch * n_out) over its samples to partial[wg][n_params]
// and its sum of squared residuals to loss_partial[wg].
// params: flat layout; w2t [64][64] and w3t [64][n_out] are the transposed
// copies of W2 and W3 (kept by the update kernel).
void launch_mlp_train(hipStream_t stream, int n_out, long long batch, long long n_data, const float* x,
                      const float* targets, unsigned long long seed, long long step, const float* params,
                      const float* w2t, const float* w3t, float* partial, double* loss_partial);

// Sums the partial gradients of `n_wg` workgroups. With adam = 0 the sum is
// written to `gradient`; otherwise an Adam step (beta1 0.9, beta2 0.999, eps
// 1e-8, bias-corrected with step_size = lr / (1 - beta1^t) and
// denom = sqrt(v / (1 - beta2^t)) + eps) updates params, m, v and the transposed
// copies. loss_out[0] receives sum(loss_partial) / (batch * n_out).
void launch_mlp_update(hipStream_t stream, int n_out, int n_wg, const float* partial, const double* loss_partial,
                       long long batch, int adam, float step_size, float bias2, float* params, float* m, float* v,
                       float* w2t, float* w3t, float* gradient, double* loss_out);

// Transposed copies of W2 and W3 from the flat parameters.
void launch_mlp_transpose(hipStream_t stream, int n_out, const float* params, float* w2t, float* w3t);

// One loss-and-gradient pass of the training kernels on device arrays with no device
// allocation: the scratch is a file-scope device array of mlp_noalloc_scratch_bytes()
// per GPU, shared by every call (overlapping calls on one GPU race). Queued on
// `stream`; loss (one double) and gradient ([n_params] floats) are device pointers.
void mlp_loss_gradient_noalloc(int n_out, long long n_samples, const float* params, const float* inputs,
                               const float* targets, double* loss, float* gradient, hipStream_t stream);
size_t mlp_noalloc_scratch_bytes();
// the largest n_samples whose scratch fits
long long mlp_noalloc_max_samples(int n_out);

// ---- output-layer solve (FP64)

// Hidden features a2 (FP32) of n samples: features[s][64].
void launch_mlp_features(hipStream_t stream, long long n, const float* x, const float* params, const float* w2t,
                         float* features);
// Number of sample chunks of the Gram kernel.
int mlp_gram_chunks(long long n);
// partial[chunk][65][65 + n_out] (FP64) of [A 1]^T [A 1 Y] over each chunk's
// samples, A = features, Y = targets [n][n_out].
void launch_mlp_gram(hipStream_t stream, long long n, int n_out, const float* features, const float* targets,
                     double* partial);
// Sums the chunks into gram[65][65 + n_out].
void launch_mlp_gram_reduce(hipStream_t stream, int n_out, int chunks, const double* partial, double* gram);
// Solves (G + ridge * trace(G) / 65 * I) w = R in FP64 by Cholesky for every
// output and writes W3 / b3 (FP32) into params and w3t. status[0] = 0 on
// success, 1 when the matrix is not positive definite.
void launch_mlp_solve(hipStream_t stream, int n_out, double ridge, double* gram, float* params, float* w3t,
                      int* status);

// ---- inference (FP16 hi + lo operands on MFMA, FP32 accumulation)

// Inference weights in MFMA operand order, built on the host by
// mlp_pack_inference: W2 / W3 split into FP16 hi and lo fragments, plus the
// FP32 first layer, biases and output scales.
//
// FP16 keeps 11 significant bits between 6.1e-5 and 65504; below that range
// values lose precision or are flushed to zero. Each operand v of the two
// products is therefore scaled by a power of two chosen per layer so that its
// largest magnitude (measured on the host for the activations) sits near 2^13,
// split as hi = fp16(v) and lo = fp16((v - hi) * 2^12), and the hi * hi
// products and the hi * lo + lo * hi corrections accumulate separately in FP32:
// z = (acc_hh + acc_corr * 2^-12) / (scale_a * scale_w).
struct MlpInference {
    int n_out = 0;
    int split = 0;         // outputs [0, split) go to out0, [split, n_out) to out1
    double x_scale = 1.0;  // network input = r * x_scale + x_shift
    double x_shift = 0.0;
    const uint16_t* frag2 = nullptr;  // [2][16 fragments][64 lanes][4]  (hi, lo)
    const uint16_t* frag3 = nullptr;  // [2][n_pad/16 * 4 fragments][64 lanes][4]
    // W1[64], b1[64], b2[64], b3[n_pad], scale[n_pad], then the operand scales
    // scale_a1, 1 / (scale_a1 scale_w2), scale_a2, 1 / (scale_a2 scale_w3)
    const float* small = nullptr;
};

inline size_t mlp_frag2_size() { return 2 * 16 * 256; }
inline size_t mlp_frag3_size(int n_out) { return 2 * (size_t)(mlp_padded_outputs(n_out) / 16) * 4 * 256; }
inline size_t mlp_small_size(int n_out) { return 3 * MLP_HIDDEN + 2 * (size_t)mlp_padded_outputs(n_out) + 4; }

// Host-side packing of the flat FP32 parameters and per-output scales into the
// three device arrays above, for network inputs in [x_min, x_max] (the
// activation maxima that set the operand scales are measured there).
void mlp_pack_inference(int n_out, const float* params, const float* scale, double x_min, double x_max,
                        uint16_t* frag2, uint16_t* frag3, float* small);
size_t mlp_inference_shmem(int n_out);

// y = network(r * x_scale + x_shift) * scale for every r; out0[i][split],
// out1[i][n_out - split].
template <typename T>
void launch_mlp_inference(hipStream_t stream, const MlpInference& net, long long n, const double* r, T* out0,
                          T* out1);

// ---- judge

// Per-column maximum of |a - b| and of |b| over n rows of `cols` values
// (FP32 arrays); out[0][cols] = max|a - b|, out[1][cols] = max|b|.
void launch_column_error(hipStream_t stream, long long n, int cols, const float* a, const float* b, double* out);

// Per-column max |values| (FP64 input [n][cols]), for output scaling.
void launch_column_maxabs(hipStream_t stream, long long n, int cols, const double* values, double* out);

// targets[i][o] = values[i][o] / scale[o] (o < cols) then derivs[i][o - cols] /
// scale[o] (cols <= o < 2 cols), as FP32.
void launch_normalise_targets(hipStream_t stream, long long n, int cols, const double* values, const double* derivs,
                              const double* scale, float* targets);

// r[i] = r_max * (i + u_i) / n with u_i in [0, 1) from a hash of (seed, i):
// stratified uniform samples over [0, r_max).
void launch_uniform_samples(hipStream_t stream, long long n, double r_max, unsigned long long seed, double* r);

// x[i] = (float)(r[i] * x_scale + x_shift)
void launch_network_inputs(hipStream_t stream, long long n, const double* r, double x_scale, double x_shift, float* x);

}  // namespace rascaline_amd
