// Learned surrogate of the SOAP radial integral: one small network gives every
// R_nl(r) and dR_nl/dr at once. It fills the slot of the reference's
// SoapRadialIntegral trait (featomic/src/calculators/soap/radial_integral/
// mod.rs:19-58), which the reference fills with the exact GTO integral
// (soap/radial_integral/gto.rs:115-205) or its spline (spline.rs:28-70).
//
// Training runs on the GPU:
//   1. samples r uniform on [0, cutoff), exact GTO integrals in FP64 as targets,
//      each (l, n) column divided by its largest magnitude;
//   2. Adam over all parameters, FP32 forward/backward kernels;
//   3. the output layer re-solved exactly by ridge least squares in FP64
//      (Gram matrix over every sample, Cholesky);
//   4. the judge: the calculator's own radial integral (spline, or exact when
//      no spline is built) against the network's inference kernel on a uniform
//      grid, per (l, n) error max|net - ref| / max|ref| for values and
//      derivatives. The network is used only if every error is within the
//      tolerance; otherwise the calculator keeps its spline.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "device.hpp"
#include "kernels_mlp.hpp"
#include "radial.hpp"

namespace rascaline_amd {

struct SurrogateOptions {
    long long n_samples = 1000000;
    long long steps = 20000;
    int batch = 2048;
    double learning_rate = 3e-3;
    double ridge = 1e-12;
    double tolerance = 1e-3;
    unsigned long long seed = 1;
    long long n_validation = 20001;
};

struct SurrogateReport {
    bool accepted = false;
    double adam_loss = 0.0;  // mean squared error of the last Adam step (normalised outputs)
    double max_value_error = 0.0;
    double max_derivative_error = 0.0;
    double train_ms = 0.0;
    std::vector<double> value_errors;       // [L][n]
    std::vector<double> derivative_errors;  // [L][n]
};

// Throws invalid_parameter for options out of range (checked before any GPU work).
void check_surrogate_options(const SurrogateOptions& options);

class RadialSurrogate {
public:
    RadialSurrogate(int max_angular, int n_radial, double cutoff);
    ~RadialSurrogate();
    RadialSurrogate(const RadialSurrogate&) = delete;
    RadialSurrogate& operator=(const RadialSurrogate&) = delete;

    // Trains on the exact integral of `tables`; `judge` is the calculator's
    // device radial integral in FP32 (spline tables) or nullptr to judge
    // against the exact integral.
    SurrogateReport train(const RadialTables& tables, const double* d_gto, const double* d_ortho,
                          const DeviceRadial<float>* judge, hipStream_t stream, const SurrogateOptions& options);

    bool accepted() const { return accepted_; }
    int n_outputs() const { return n_out_; }
    const std::vector<float>& parameters() const { return params_; }
    const std::vector<float>& output_scale() const { return scale_; }

    // values / derivatives [count][L][n] at the distances r (device arrays)
    template <typename T>
    void evaluate(hipStream_t stream, long long count, const double* r, T* values, T* derivatives) const;

private:
    int L_, n_, n_out_;
    double cutoff_;
    bool accepted_ = false;
    std::vector<float> params_;
    std::vector<float> scale_;
    uint16_t* d_frag2_ = nullptr;
    uint16_t* d_frag3_ = nullptr;
    float* d_small_ = nullptr;

    void upload_inference();
    MlpInference inference() const;
};

// One FP32 forward + backward of the training kernels over all `n_samples`
// (host arrays in and out); returns the loss.
double mlp_loss_gradient_device(int n_out, long long n_samples, const float* params, const float* inputs,
                                const float* targets, float* gradient);
// The inference kernel over host inputs (network inputs used as given).
void mlp_forward_device(int n_out, long long n_samples, const float* params, const float* inputs, float* outputs);

}  // namespace rascaline_amd
