// Host-side FP64 construction of the radial integral tables that the device
// kernels evaluate: GTO orthonormalisation matrices and per-l cubic Hermite
// spline knots (values + derivatives).
#pragma once

#include <vector>

#include "hypers.hpp"

namespace rascaline_amd {

// Orthonormalisation of the GTO radial basis, O = S^{-1/2} diag(norm), row-major
// (featomic/src/calculators/shared/basis/radial/gto.rs:17-77). Throws
// Error::internal when S is singular.
std::vector<double> gto_orthonormalization(double radius, int n);

// Knots of one angular channel: uniform grid of K intervals over [0, stop].
struct SplineChannel {
    std::vector<double> positions;    // K + 1
    std::vector<double> values;       // (K + 1) x n
    std::vector<double> derivatives;  // (K + 1) x n
};

// All channels of one radial integral, with the exact evaluator's parameters.
struct RadialTables {
    bool lode = false;
    int max_angular = 0;
    int n = 0;
    double stop = 0.0;                 // cutoff (SOAP) or k_cutoff (LODE)
    double sigma2_atomic = 0.0;        // SOAP: Gaussian width^2
    std::vector<double> gto;           // SOAP: 1 / (2 sigma_n^2); LODE: sigma_n
    std::vector<double> ortho;         // n x n, values <- values . ortho^T
    bool use_spline = true;
    std::vector<SplineChannel> channels;  // one per l when use_spline

    // exact evaluation (orthonormalised), values / grads of size n
    void exact(int l, double x, double* values, double* grads) const;
};

// Builds the tables, including HermitCubicSpline::with_accuracy
// (featomic/src/math/splines.rs:79-194): 11 uniform knots, then midpoints are
// added until the mean absolute or mean relative error at the midpoints falls
// below the requested accuracy (at most 10 000 knots).
RadialTables build_soap_radial(const SoapParams& params);

// Hermite spline of one channel at x (values / grads of size n; grads may be null),
// with the reference's interval choice (the last knot <= x)
void evaluate_channel(const SplineChannel& channel, int n, double x, double* values, double* grads);
RadialTables build_lode_radial(const LodeParams& params);

// LODE centre contribution (featomic/src/calculators/lode/radial_integral/gto.rs:136-185)
std::vector<double> lode_center_contribution(const LodeParams& params, const RadialTables& tables);

}  // namespace rascaline_amd
