// Training, judging and evaluation of the learned radial-integral surrogate.
#include "surrogate.hpp"

#include <algorithm>
#include <cmath>
#include <random>

#include "kernels.hpp"

namespace rascaline_amd {

namespace {

constexpr double PI = 3.141592653589793;

template <typename T>
struct DeviceArray {
    T* ptr = nullptr;
    explicit DeviceArray(size_t count) { RA_HIP_CHECK(hipMalloc(&ptr, std::max<size_t>(count, 1) * sizeof(T))); }
    ~DeviceArray() {
        if (ptr) (void)hipFree(ptr);
    }
    DeviceArray(const DeviceArray&) = delete;
    DeviceArray& operator=(const DeviceArray&) = delete;
};

// Waits for the stream when leaving a scope, before the arrays declared earlier
// in that scope are freed (also when an exception unwinds it).
struct StreamSync {
    hipStream_t stream;
    ~StreamSync() { (void)hipStreamSynchronize(stream); }
};

template <typename T>
void to_device(T* dst, const T* src, size_t count, hipStream_t stream) {
    if (count) RA_HIP_CHECK(hipMemcpyAsync(dst, src, count * sizeof(T), hipMemcpyHostToDevice, stream));
}

template <typename T>
void to_host(T* dst, const T* src, size_t count, hipStream_t stream) {
    if (count) RA_HIP_CHECK(hipMemcpyAsync(dst, src, count * sizeof(T), hipMemcpyDeviceToHost, stream));
}

// Hidden transitions spread over the input range: unit u of the first layer
// switches at x = -1.05 + 2.1 u / 63 with slope +-8 (alternating), so the first
// layer starts as a smooth piecewise-linear basis of [-1, 1]. The other layers
// start uniform in +-1/sqrt(64).
std::vector<float> initial_parameters(int n_out, unsigned long long seed) {
    const int H = MLP_HIDDEN;
    std::vector<float> p(static_cast<size_t>(mlp_n_params(n_out)));
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> uni(-0.125, 0.125);
    for (int u = 0; u < H; u++) {
        const double slope = u % 2 == 0 ? 8.0 : -8.0;
        const double knot = -1.05 + 2.1 * u / (H - 1);
        p[u] = static_cast<float>(slope);
        p[H + u] = static_cast<float>(-slope * knot);
    }
    for (size_t k = 2 * H; k < p.size(); k++) p[k] = static_cast<float>(uni(rng));
    return p;
}

}  // namespace

RadialSurrogate::RadialSurrogate(int max_angular, int n_radial, double cutoff)
    : L_(max_angular + 1), n_(n_radial), n_out_(2 * (max_angular + 1) * n_radial), cutoff_(cutoff) {
    if (n_out_ > MLP_MAX_OUTPUTS) {
        throw Error::invalid_parameter("the radial surrogate supports at most " + std::to_string(MLP_MAX_OUTPUTS / 2) +
                                       " (l, n) channels, (max_angular + 1) * n_radial is " +
                                       std::to_string(L_ * n_));
    }
}

RadialSurrogate::~RadialSurrogate() {
    if (d_frag2_) (void)hipFree(d_frag2_);
    if (d_frag3_) (void)hipFree(d_frag3_);
    if (d_small_) (void)hipFree(d_small_);
}

MlpInference RadialSurrogate::inference() const {
    MlpInference net;
    net.n_out = n_out_;
    net.split = n_out_ / 2;
    net.x_scale = 2.0 / cutoff_;
    net.x_shift = -1.0;
    net.frag2 = d_frag2_;
    net.frag3 = d_frag3_;
    net.small = d_small_;
    return net;
}

void RadialSurrogate::upload_inference() {
    std::vector<uint16_t> frag2(mlp_frag2_size()), frag3(mlp_frag3_size(n_out_));
    std::vector<float> small(mlp_small_size(n_out_));
    // network inputs are distances mapped to [-1, 1]
    mlp_pack_inference(n_out_, params_.data(), scale_.data(), -1.0, 1.0, frag2.data(), frag3.data(), small.data());
    if (!d_frag2_) {
        RA_HIP_CHECK(hipMalloc(&d_frag2_, frag2.size() * sizeof(uint16_t)));
        RA_HIP_CHECK(hipMalloc(&d_frag3_, frag3.size() * sizeof(uint16_t)));
        RA_HIP_CHECK(hipMalloc(&d_small_, small.size() * sizeof(float)));
    }
    RA_HIP_CHECK(hipMemcpy(d_frag2_, frag2.data(), frag2.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
    RA_HIP_CHECK(hipMemcpy(d_frag3_, frag3.data(), frag3.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
    RA_HIP_CHECK(hipMemcpy(d_small_, small.data(), small.size() * sizeof(float), hipMemcpyHostToDevice));
}

namespace {

size_t device_lds_bytes() {
    int dev = 0;
    RA_HIP_CHECK(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    RA_HIP_CHECK(hipGetDeviceProperties(&prop, dev));
    return prop.sharedMemPerBlock > 0 ? static_cast<size_t>(prop.sharedMemPerBlock) : 65536;
}

void check_inference_fits(int n_out) {
    const size_t need = mlp_inference_shmem(n_out);
    const size_t have = device_lds_bytes();
    if (need > have) {
        throw Error::invalid_parameter("the radial surrogate's inference weights need " + std::to_string(need) +
                                       " bytes of local memory, the GPU has " + std::to_string(have));
    }
}

}  // namespace

template <typename T>
void RadialSurrogate::evaluate(hipStream_t stream, long long count, const double* r, T* values, T* derivatives) const {
    if (!d_frag2_) throw Error::internal("the radial surrogate has not been trained");
    launch_mlp_inference<T>(stream, inference(), count, r, values, derivatives);
    RA_HIP_CHECK(hipGetLastError());
}
template void RadialSurrogate::evaluate<float>(hipStream_t, long long, const double*, float*, float*) const;
template void RadialSurrogate::evaluate<double>(hipStream_t, long long, const double*, double*, double*) const;

void check_surrogate_options(const SurrogateOptions& options) {
    if (options.n_samples < 1 || options.steps < 0 || options.batch < 1 || options.n_validation < 2 ||
        !(options.learning_rate > 0.0) || !(options.ridge >= 0.0) || !(options.tolerance > 0.0)) {
        throw Error::invalid_parameter(
            "invalid radial surrogate options: n_samples >= 1, steps >= 0, batch >= 1, n_validation >= 2, "
            "learning_rate > 0, ridge >= 0 and tolerance > 0 are required");
    }
}

SurrogateReport RadialSurrogate::train(const RadialTables& tables, const double* d_gto, const double* d_ortho,
                                       const DeviceRadial<float>* judge, hipStream_t stream,
                                       const SurrogateOptions& options) {
    check_surrogate_options(options);
    check_inference_fits(n_out_);
    accepted_ = false;
    const long long N = options.n_samples;
    const int cols = L_ * n_;
    const long long P = mlp_n_params(n_out_);
    const int ts = mlp_train_tile(n_out_);
    const long long batch = std::min<long long>(options.batch, N);
    const int n_wg = static_cast<int>((batch + ts - 1) / ts);

    hipEvent_t start, stop;
    RA_HIP_CHECK(hipEventCreate(&start));
    RA_HIP_CHECK(hipEventCreate(&stop));
    RA_HIP_CHECK(hipEventRecord(start, stream));

    DeviceArray<double> r(N);
    DeviceArray<float> x(N);
    DeviceArray<float> targets(static_cast<size_t>(N) * n_out_);
    DeviceArray<double> scale(n_out_);
    DeviceArray<float> params(P), m(P), v(P), w2t(MLP_HIDDEN * MLP_HIDDEN), w3t(static_cast<size_t>(MLP_HIDDEN) * n_out_);
    DeviceArray<float> partial(static_cast<size_t>(n_wg) * P);
    DeviceArray<double> loss_partial(n_wg);
    DeviceArray<double> loss(std::max<long long>(options.steps, 1));
    StreamSync sync{stream};

    // 1. samples and exact targets (FP64), normalised per column
    launch_uniform_samples(stream, N, cutoff_, options.seed, r.ptr);
    {
        DeviceArray<double> val(static_cast<size_t>(N) * cols), der(static_cast<size_t>(N) * cols);
        StreamSync inner{stream};
        launch_radial_exact<double>(stream, N, L_, n_, r.ptr, tables.sigma2_atomic, d_gto, d_ortho, val.ptr, der.ptr);
        launch_column_maxabs(stream, N, cols, val.ptr, scale.ptr);
        launch_column_maxabs(stream, N, cols, der.ptr, scale.ptr + cols);
        launch_normalise_targets(stream, N, cols, val.ptr, der.ptr, scale.ptr, targets.ptr);
        RA_HIP_CHECK(hipGetLastError());
    }
    launch_network_inputs(stream, N, r.ptr, 2.0 / cutoff_, -1.0, x.ptr);

    // 2. Adam over all parameters
    params_ = initial_parameters(n_out_, options.seed);
    to_device(params.ptr, params_.data(), params_.size(), stream);
    RA_HIP_CHECK(hipMemsetAsync(m.ptr, 0, P * sizeof(float), stream));
    RA_HIP_CHECK(hipMemsetAsync(v.ptr, 0, P * sizeof(float), stream));
    launch_mlp_transpose(stream, n_out_, params.ptr, w2t.ptr, w3t.ptr);
    const double lr_min = options.learning_rate * 1e-3;
    const unsigned long long batch_seed = options.seed * 0x2545F4914F6CDD1Dull + 1;
    for (long long step = 0; step < options.steps; step++) {
        const double lr = lr_min + 0.5 * (options.learning_rate - lr_min) *
                                       (1.0 + std::cos(PI * static_cast<double>(step) / options.steps));
        const double t = static_cast<double>(step + 1);
        const double bias1 = 1.0 - std::pow(0.9, t), bias2 = 1.0 - std::pow(0.999, t);
        launch_mlp_train(stream, n_out_, batch, N, x.ptr, targets.ptr, batch_seed, step, params.ptr, w2t.ptr, w3t.ptr,
                         partial.ptr, loss_partial.ptr);
        launch_mlp_update(stream, n_out_, n_wg, partial.ptr, loss_partial.ptr, batch, 1,
                          static_cast<float>(lr / bias1), static_cast<float>(bias2), params.ptr, m.ptr, v.ptr, w2t.ptr,
                          w3t.ptr, nullptr, loss.ptr + step);
    }
    RA_HIP_CHECK(hipGetLastError());

    // 3. output layer by ridge least squares in FP64 over every sample
    {
        const int chunks = mlp_gram_chunks(N);
        const size_t E = static_cast<size_t>(MLP_HIDDEN + 1) * (MLP_HIDDEN + 1 + n_out_);
        DeviceArray<float> features(static_cast<size_t>(N) * MLP_HIDDEN);
        DeviceArray<double> gram_partial(static_cast<size_t>(chunks) * E), gram(E);
        DeviceArray<int> status(1);
        StreamSync inner{stream};
        launch_mlp_features(stream, N, x.ptr, params.ptr, w2t.ptr, features.ptr);
        launch_mlp_gram(stream, N, n_out_, features.ptr, targets.ptr, gram_partial.ptr);
        launch_mlp_gram_reduce(stream, n_out_, chunks, gram_partial.ptr, gram.ptr);
        launch_mlp_solve(stream, n_out_, options.ridge, gram.ptr, params.ptr, w3t.ptr, status.ptr);
        RA_HIP_CHECK(hipGetLastError());
        int st = 0;
        to_host(&st, status.ptr, 1, stream);
        RA_HIP_CHECK(hipStreamSynchronize(stream));
        if (st != 0) {
            throw Error::internal("radial surrogate: the output-layer normal equations are not positive definite; "
                                  "increase the ridge");
        }
    }

    SurrogateReport report;
    to_host(params_.data(), params.ptr, params_.size(), stream);
    std::vector<double> scale_host(n_out_);
    to_host(scale_host.data(), scale.ptr, scale_host.size(), stream);
    if (options.steps > 0) to_host(&report.adam_loss, loss.ptr + options.steps - 1, 1, stream);
    RA_HIP_CHECK(hipStreamSynchronize(stream));
    scale_.assign(scale_host.begin(), scale_host.end());
    upload_inference();

    // 4. judge on a uniform grid of [0, cutoff)
    {
        const long long nv = options.n_validation;
        std::vector<double> grid(nv);
        for (long long k = 0; k < nv; k++) grid[k] = cutoff_ * static_cast<double>(k) / static_cast<double>(nv);
        DeviceArray<double> rv(nv);
        DeviceArray<float> ref_val(static_cast<size_t>(nv) * cols), ref_der(static_cast<size_t>(nv) * cols);
        DeviceArray<float> net_val(static_cast<size_t>(nv) * cols), net_der(static_cast<size_t>(nv) * cols);
        DeviceArray<double> err(4 * static_cast<size_t>(cols));
        StreamSync inner{stream};
        to_device(rv.ptr, grid.data(), grid.size(), stream);
        if (judge) {
            launch_radial_eval<float>(stream, nv, *judge, rv.ptr, ref_val.ptr, ref_der.ptr);
        } else {
            launch_radial_exact<float>(stream, nv, L_, n_, rv.ptr, tables.sigma2_atomic, d_gto, d_ortho, ref_val.ptr,
                                       ref_der.ptr);
        }
        evaluate<float>(stream, nv, rv.ptr, net_val.ptr, net_der.ptr);
        launch_column_error(stream, nv, cols, net_val.ptr, ref_val.ptr, err.ptr);
        launch_column_error(stream, nv, cols, net_der.ptr, ref_der.ptr, err.ptr + 2 * cols);
        RA_HIP_CHECK(hipGetLastError());
        std::vector<double> e(4 * static_cast<size_t>(cols));
        to_host(e.data(), err.ptr, e.size(), stream);
        RA_HIP_CHECK(hipEventRecord(stop, stream));
        RA_HIP_CHECK(hipStreamSynchronize(stream));
        report.value_errors.resize(cols);
        report.derivative_errors.resize(cols);
        for (int c = 0; c < cols; c++) {
            const double dv = e[c], rv_max = e[cols + c], dd = e[2 * cols + c], rd_max = e[3 * cols + c];
            report.value_errors[c] = rv_max > 0.0 ? dv / rv_max : dv;
            report.derivative_errors[c] = rd_max > 0.0 ? dd / rd_max : dd;
        }
    }
    float ms = 0.0f;
    RA_HIP_CHECK(hipEventElapsedTime(&ms, start, stop));
    (void)hipEventDestroy(start);
    (void)hipEventDestroy(stop);
    report.train_ms = ms;
    report.max_value_error = *std::max_element(report.value_errors.begin(), report.value_errors.end());
    report.max_derivative_error = *std::max_element(report.derivative_errors.begin(), report.derivative_errors.end());
    report.accepted = std::isfinite(report.max_value_error) && std::isfinite(report.max_derivative_error) &&
                      report.max_value_error <= options.tolerance && report.max_derivative_error <= options.tolerance;
    accepted_ = report.accepted;
    return report;
}

double mlp_loss_gradient_device(int n_out, long long n_samples, const float* params, const float* inputs,
                                const float* targets, float* gradient) {
    if (n_out < 1 || n_out > MLP_MAX_OUTPUTS) throw Error::invalid_parameter("n_out must be in [1, 512]");
    if (n_samples < 1) throw Error::invalid_parameter("n_samples must be positive");
    hipStream_t stream = nullptr;
    const long long P = mlp_n_params(n_out);
    const int ts = mlp_train_tile(n_out);
    const int n_wg = static_cast<int>((n_samples + ts - 1) / ts);
    DeviceArray<float> p(P), x(n_samples), t(static_cast<size_t>(n_samples) * n_out), w2t(MLP_HIDDEN * MLP_HIDDEN),
        w3t(static_cast<size_t>(MLP_HIDDEN) * n_out), partial(static_cast<size_t>(n_wg) * P), g(P);
    DeviceArray<double> loss_partial(n_wg), loss(1);
    StreamSync sync{stream};
    to_device(p.ptr, params, P, stream);
    to_device(x.ptr, inputs, n_samples, stream);
    to_device(t.ptr, targets, static_cast<size_t>(n_samples) * n_out, stream);
    launch_mlp_transpose(stream, n_out, p.ptr, w2t.ptr, w3t.ptr);
    launch_mlp_train(stream, n_out, n_samples, n_samples, x.ptr, t.ptr, 0, 0, p.ptr, w2t.ptr, w3t.ptr, partial.ptr,
                     loss_partial.ptr);
    launch_mlp_update(stream, n_out, n_wg, partial.ptr, loss_partial.ptr, n_samples, 0, 0.0f, 1.0f, nullptr, nullptr,
                      nullptr, nullptr, nullptr, g.ptr, loss.ptr);
    RA_HIP_CHECK(hipGetLastError());
    double l = 0.0;
    to_host(gradient, g.ptr, P, stream);
    to_host(&l, loss.ptr, 1, stream);
    RA_HIP_CHECK(hipStreamSynchronize(stream));
    return l;
}

void mlp_forward_device(int n_out, long long n_samples, const float* params, const float* inputs, float* outputs) {
    if (n_out < 1 || n_out > MLP_MAX_OUTPUTS) throw Error::invalid_parameter("n_out must be in [1, 512]");
    if (n_samples < 1) throw Error::invalid_parameter("n_samples must be positive");
    check_inference_fits(n_out);
    hipStream_t stream = nullptr;
    std::vector<uint16_t> frag2(mlp_frag2_size()), frag3(mlp_frag3_size(n_out));
    std::vector<float> small(mlp_small_size(n_out));
    std::vector<float> ones(n_out, 1.0f);
    const auto range = std::minmax_element(inputs, inputs + n_samples);
    mlp_pack_inference(n_out, params, ones.data(), *range.first, *range.second, frag2.data(), frag3.data(),
                       small.data());
    std::vector<double> r(inputs, inputs + n_samples);
    DeviceArray<uint16_t> f2(frag2.size()), f3(frag3.size());
    DeviceArray<float> sm(small.size()), out(static_cast<size_t>(n_samples) * n_out);
    DeviceArray<double> rd(n_samples);
    StreamSync sync{stream};
    to_device(f2.ptr, frag2.data(), frag2.size(), stream);
    to_device(f3.ptr, frag3.data(), frag3.size(), stream);
    to_device(sm.ptr, small.data(), small.size(), stream);
    to_device(rd.ptr, r.data(), r.size(), stream);
    MlpInference net;
    net.n_out = n_out;
    net.split = n_out;
    net.frag2 = f2.ptr;
    net.frag3 = f3.ptr;
    net.small = sm.ptr;
    launch_mlp_inference<float>(stream, net, n_samples, rd.ptr, out.ptr, nullptr);
    RA_HIP_CHECK(hipGetLastError());
    to_host(outputs, out.ptr, static_cast<size_t>(n_samples) * n_out, stream);
    RA_HIP_CHECK(hipStreamSynchronize(stream));
}

}  // namespace rascaline_amd
