// Calculator hyper-parameters with featomic's JSON schema and validation rules.
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "json.hpp"

namespace rascaline_amd {

// Status codes, featomic/include/featomic.h:19-51
constexpr int STATUS_SUCCESS = 0;
constexpr int STATUS_INVALID_PARAMETER = 1;
constexpr int STATUS_JSON = 2;
constexpr int STATUS_UTF8 = 3;
constexpr int STATUS_SYSTEM = 128;
constexpr int STATUS_BUFFER_SIZE = 254;
constexpr int STATUS_INTERNAL = 255;

// Error carrying a featomic status code; the message follows the reference's
// Display implementation (featomic/src/errors.rs:27-45).
struct Error : std::runtime_error {
    int status;
    Error(int status_, const std::string& message) : std::runtime_error(message), status(status_) {}

    static Error invalid_parameter(const std::string& m) { return Error(STATUS_INVALID_PARAMETER, "invalid parameter: " + m); }
    static Error json(const std::string& m) { return Error(STATUS_JSON, "json error: " + m); }
    static Error internal(const std::string& m) { return Error(STATUS_INTERNAL, "internal featomic error: " + m); }
    static Error buffer_size(const std::string& m) { return Error(STATUS_BUFFER_SIZE, "buffer is not big enough: " + m); }
    static Error system(int status, const std::string& m) {
        return Error(status < 0 ? status : STATUS_SYSTEM, "error from external code (status " + std::to_string(status) + "): " + m);
    }
};

enum class Smoothing { Step, ShiftedCosine };

// featomic/src/calculators/shared/density.rs:27-60
enum class DensityKind { Gaussian, DiracDelta };

// One radial basis of the SOAP expansion (featomic's SoapRadialBasis,
// shared/basis/radial/mod.rs:10-48): GTO with n functions, or a tabulated radial
// integral given as Hermite spline points (shared/basis/radial/tabulated.rs:43-81),
// sorted by position.
struct RadialSpec {
    bool tabulated = false;
    int n = 0;
    std::vector<double> positions;                // tabulated: [points]
    std::vector<double> values, derivatives;      // tabulated: [points][n]
};

// SphericalExpansionParameters / PowerSpectrumParameters
// (featomic/src/calculators/soap/spherical_expansion_pair.rs:37-60,
//  featomic/src/calculators/soap/power_spectrum.rs:35-43)
struct SoapParams {
    double cutoff = 0.0;
    Smoothing smoothing = Smoothing::Step;
    double smoothing_width = 0.0;
    double gaussian_width = 0.0;
    bool has_scaling = false;
    double scaling_scale = 0.0, scaling_rate = 0.0, scaling_exponent = 0.0;
    double center_atom_weight = 1.0;
    int max_angular = 0;
    int n_radial = 0;             // max_radial + 1
    bool use_spline = true;
    double spline_accuracy = 1e-8;

    DensityKind density = DensityKind::Gaussian;
    // the basis: its angular channels (ascending) and the radial basis of each. A
    // TensorProduct basis has channels 0..max_angular sharing one radial basis; an
    // Explicit basis (shared/basis/mod.rs:95-116) lists its channels, each with its
    // own. max_angular and n_radial are the largest channel and radial size: the
    // tensor-product shape the kernels compute.
    bool explicit_basis = false;
    std::vector<int> lambdas;
    std::vector<RadialSpec> radial;
    // the computed shape holds channels or radial functions the basis does not have
    // (an Explicit basis with missing channels or smaller radial bases); the output
    // keeps only the basis' own
    bool padded() const {
        if (static_cast<int>(lambdas.size()) != max_angular + 1) return true;
        for (const auto& r : radial) {
            if (r.n != n_radial) return true;
        }
        return false;
    }
    // radial size of channel l (0 when l is not a channel of the basis)
    int n_of(int l) const {
        for (size_t k = 0; k < lambdas.size(); k++) {
            if (lambdas[k] == l) return radial[k].n;
        }
        return 0;
    }
    const RadialSpec* radial_of(int l) const {
        for (size_t k = 0; k < lambdas.size(); k++) {
            if (lambdas[k] == l) return &radial[k];
        }
        return nullptr;
    }

    int L() const { return max_angular + 1; }
    int L2() const { return (max_angular + 1) * (max_angular + 1); }
};

// LodeSphericalExpansionParameters (featomic/src/calculators/lode/spherical_expansion.rs:36-60)
struct LodeParams {
    double smearing = 0.0;
    int exponent = 0;
    double center_atom_weight = 1.0;
    int max_angular = 0;
    int n_radial = 0;
    double radius = 0.0;
    double k_cutoff = 0.0;        // resolved default 1.2 pi / smearing
    bool use_spline = true;
    double spline_accuracy = 1e-8;

    int L() const { return max_angular + 1; }
    int L2() const { return (max_angular + 1) * (max_angular + 1); }
};

// NeighborList parameters (featomic/src/calculators/neighbor_list.rs:31-44): three
// required fields, unknown fields ignored
struct NeighborListParams {
    double cutoff = 0.0;
    bool full_neighbor_list = false;
    bool self_pairs = false;
};

// `deny_unknown_fields = true` mirrors serde's attribute on the SE parameters;
// the power spectrum parameters do not deny unknown top-level fields
// (featomic/src/calculators/soap/power_spectrum.rs:33-35).
// radial_spectrum: the radial spectrum's parameters (basis {radial, spline_accuracy}),
// parsed as a spherical expansion with max_angular = 0
SoapParams parse_soap_params(const std::string& json, bool deny_unknown_fields, bool radial_spectrum = false);
LodeParams parse_lode_params(const std::string& json);
NeighborListParams parse_neighbor_list_params(const std::string& json);

}  // namespace rascaline_amd
