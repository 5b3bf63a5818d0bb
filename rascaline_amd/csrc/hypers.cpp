#include "hypers.hpp"

#include <cmath>
#include <initializer_list>

namespace rascaline_amd {

namespace {

Json parse_json(const std::string& text) {
    try {
        return Json::parse(text);
    } catch (const JsonError& e) {
        throw Error::json(e.what());
    }
}

const Json& require(const Json& obj, const char* key, const char* context) {
    if (!obj.is_object()) throw Error::json(std::string("invalid type for ") + context + ", expected a map");
    const Json* v = obj.find(key);
    if (!v) throw Error::json(std::string("missing field `") + key + "`");
    return *v;
}

double number(const Json& obj, const char* key, const char* context) {
    const Json& v = require(obj, key, context);
    if (!v.is_number()) throw Error::json(std::string("invalid type for `") + key + "`, expected f64");
    return v.as_number();
}

int unsigned_int(const Json& obj, const char* key, const char* context) {
    double v = number(obj, key, context);
    if (v < 0 || v != std::floor(v) || v > 1e6) {
        throw Error::json(std::string("invalid value for `") + key + "`, expected usize");
    }
    return static_cast<int>(v);
}

void deny_unknown(const Json& obj, std::initializer_list<const char*> allowed) {
    for (auto& kv : obj.as_object()) {
        bool ok = false;
        for (const char* a : allowed) ok |= kv.first == a;
        if (!ok) throw Error::json("unknown field `" + kv.first + "`");
    }
}

std::string tag(const Json& obj, const char* context) {
    const Json& t = require(obj, "type", context);
    if (!t.is_string()) throw Error::json(std::string("invalid type tag for ") + context);
    return t.as_string();
}

// SphericalExpansionBasis<RadialBasis> with a GTO radial basis
// (featomic/src/calculators/shared/basis/mod.rs:9-116, basis/radial/mod.rs:10-88)
// With `radial_only`, the RadialSpectrumBasis of the radial spectrum
// (soap/radial_spectrum.rs:44-58): {radial, spline_accuracy}, l = 0 only.
void parse_basis(const Json& basis, int* max_angular, int* n_radial, double* gto_radius, bool* has_radius,
                 bool* use_spline, double* accuracy, bool radial_only = false) {
    if (radial_only) {
        deny_unknown(basis, {"radial", "spline_accuracy"});
        *max_angular = 0;
    } else {
        std::string type = tag(basis, "basis");
        if (type == "Explicit") {
            throw Error::invalid_parameter("the Explicit basis is not supported by this implementation, use TensorProduct");
        }
        if (type != "TensorProduct") throw Error::json("unknown variant `" + type + "` for basis");
        deny_unknown(basis, {"type", "max_angular", "radial", "spline_accuracy"});
        *max_angular = unsigned_int(basis, "max_angular", "basis");
    }
    const Json& radial = require(basis, "radial", "basis");
    std::string rtype = tag(radial, "radial basis");
    if (rtype == "Tabulated") {
        throw Error::invalid_parameter("Tabulated radial basis is not supported by this implementation");
    }
    if (rtype != "Gto") throw Error::json("unknown variant `" + rtype + "` for radial basis");
    deny_unknown(radial, {"type", "max_radial", "radius"});
    *n_radial = unsigned_int(radial, "max_radial", "radial basis") + 1;
    const Json* radius = radial.find("radius");
    *has_radius = radius && !radius->is_null();
    if (*has_radius) {
        if (!radius->is_number()) throw Error::json("invalid type for `radius`, expected f64");
        *gto_radius = radius->as_number();
    }
    const Json* acc = basis.find("spline_accuracy");
    *use_spline = true;
    *accuracy = 1e-8;  // featomic/src/calculators/shared/basis/mod.rs:39
    if (acc) {
        if (acc->is_null()) {
            *use_spline = false;
        } else if (acc->is_number()) {
            *accuracy = acc->as_number();
            if (*accuracy < 0.0) {
                throw Error::invalid_parameter("got invalid accuracy in spline (" + std::to_string(*accuracy) + "), it must be positive");
            }
        } else {
            throw Error::json("invalid type for `spline_accuracy`");
        }
    }
    if (*n_radial > 64) throw Error::invalid_parameter("max_radial is limited to 63 in this implementation");
    if (*max_angular > 30) throw Error::invalid_parameter("max_angular is limited to 30 in this implementation");
}

}  // namespace

SoapParams parse_soap_params(const std::string& text, bool deny_unknown_fields, bool radial_spectrum) {
    Json root = parse_json(text);
    if (!root.is_object()) throw Error::json("invalid type, expected a map");
    if (deny_unknown_fields) deny_unknown(root, {"cutoff", "density", "basis"});
    SoapParams p;

    // Cutoff, featomic/src/calculators/soap/cutoff.rs:5-46
    const Json& cutoff = require(root, "cutoff", "parameters");
    deny_unknown(cutoff, {"radius", "smoothing"});
    p.cutoff = number(cutoff, "radius", "cutoff");
    const Json& smoothing = require(cutoff, "smoothing", "cutoff");
    std::string st = tag(smoothing, "smoothing");
    if (st == "ShiftedCosine") {
        deny_unknown(smoothing, {"type", "width"});
        p.smoothing = Smoothing::ShiftedCosine;
        p.smoothing_width = number(smoothing, "width", "smoothing");
        if (p.smoothing_width <= 0.0 || !std::isfinite(p.smoothing_width)) {
            throw Error::invalid_parameter("expected positive width for shifted cosine cutoff function, got " +
                                           std::to_string(p.smoothing_width));
        }
    } else if (st == "Step") {
        deny_unknown(smoothing, {"type"});
        p.smoothing = Smoothing::Step;
    } else {
        throw Error::json("unknown variant `" + st + "` for smoothing");
    }
    if (!(p.cutoff > 0.0) || !std::isfinite(p.cutoff)) {
        throw Error::invalid_parameter("cutoff radius must be positive and finite");
    }

    // Density, featomic/src/calculators/shared/density.rs:5-106
    const Json& density = require(root, "density", "parameters");
    std::string dt = tag(density, "density");
    if (dt == "Gaussian") {
        p.gaussian_width = number(density, "width", "density");
    } else if (dt == "DiracDelta" || dt == "SmearedPowerLaw") {
        throw Error::invalid_parameter("this combination of basis and density is not supported in SOAP");
    } else {
        throw Error::json("unknown variant `" + dt + "` for density");
    }
    const Json* scaling = density.find("scaling");
    if (scaling && !scaling->is_null()) {
        std::string sct = tag(*scaling, "scaling");
        if (sct != "Willatt2018") throw Error::json("unknown variant `" + sct + "` for scaling");
        deny_unknown(*scaling, {"type", "scale", "rate", "exponent"});
        p.has_scaling = true;
        p.scaling_scale = number(*scaling, "scale", "scaling");
        p.scaling_rate = number(*scaling, "rate", "scaling");
        p.scaling_exponent = number(*scaling, "exponent", "scaling");
        if (p.scaling_scale <= 0.0) throw Error::invalid_parameter("expected positive scale for Willatt2018 radial scaling function, got " + std::to_string(p.scaling_scale));
        if (p.scaling_rate <= 0.0) throw Error::invalid_parameter("expected positive rate for Willatt2018 radial scaling function, got " + std::to_string(p.scaling_rate));
        if (p.scaling_exponent <= 0.0) throw Error::invalid_parameter("expected positive exponent for Willatt2018 radial scaling function, got " + std::to_string(p.scaling_exponent));
    }
    const Json* caw = density.find("center_atom_weight");
    if (caw) {
        if (!caw->is_number()) throw Error::json("invalid type for `center_atom_weight`, expected f64");
        p.center_atom_weight = caw->as_number();
    }
    if (!(p.gaussian_width > 1e-16) || !std::isfinite(p.gaussian_width)) {
        throw Error::invalid_parameter("Gaussian density width must be positive and finite");
    }

    // Basis
    const Json& basis = require(root, "basis", "parameters");
    double gto_radius = 0.0;
    bool has_radius = false;
    parse_basis(basis, &p.max_angular, &p.n_radial, &gto_radius, &has_radius, &p.use_spline, &p.spline_accuracy,
                radial_spectrum);
    if (has_radius && gto_radius != p.cutoff) {
        // featomic/src/calculators/soap/radial_integral/gto.rs:39-47
        throw Error::internal("GTO radius must be the same as the cutoff radius in SOAP, or should not provided");
    }
    return p;
}

LodeParams parse_lode_params(const std::string& text) {
    Json root = parse_json(text);
    if (!root.is_object()) throw Error::json("invalid type, expected a map");
    LodeParams p;
    const Json& density = require(root, "density", "parameters");
    std::string dt = tag(density, "density");
    if (dt != "SmearedPowerLaw") {
        if (dt == "Gaussian" || dt == "DiracDelta") {
            throw Error::invalid_parameter("only SmearedPowerLaw density can be used with LODE");
        }
        throw Error::json("unknown variant `" + dt + "` for density");
    }
    p.smearing = number(density, "smearing", "density");
    p.exponent = unsigned_int(density, "exponent", "density");
    const Json* scaling = density.find("scaling");
    if (scaling && !scaling->is_null()) throw Error::invalid_parameter("LODE does not support custom density scaling");
    const Json* caw = density.find("center_atom_weight");
    if (caw) {
        if (!caw->is_number()) throw Error::json("invalid type for `center_atom_weight`, expected f64");
        p.center_atom_weight = caw->as_number();
    }
    if (p.exponent >= 10) throw Error::invalid_parameter("LODE is only implemented for potential_exponent < 10");

    const Json& basis = require(root, "basis", "parameters");
    bool has_radius = false;
    parse_basis(basis, &p.max_angular, &p.n_radial, &p.radius, &has_radius, &p.use_spline, &p.spline_accuracy);
    if (!has_radius) throw Error::json("missing field `radius`");
    if (p.radius < 1e-16) throw Error::invalid_parameter("radius of GTO radial basis can not be negative");
    if (!std::isfinite(p.radius)) throw Error::invalid_parameter("radius of GTO radial basis can not be infinite/NaN");

    const Json* k_cutoff = root.find("k_cutoff");
    if (k_cutoff && !k_cutoff->is_null()) {
        if (!k_cutoff->is_number()) throw Error::json("invalid type for `k_cutoff`, expected f64");
        p.k_cutoff = k_cutoff->as_number();
    } else {
        p.k_cutoff = 1.2 * 3.141592653589793 / p.smearing;
    }
    return p;
}

NeighborListParams parse_neighbor_list_params(const std::string& text) {
    Json root = parse_json(text);
    if (!root.is_object()) throw Error::json("invalid type, expected a map");
    NeighborListParams p;
    p.cutoff = number(root, "cutoff", "parameters");
    for (const char* key : {"full_neighbor_list", "self_pairs"}) {
        const Json& v = require(root, key, "parameters");
        bool b = false;
        try {
            b = v.as_bool();
        } catch (const JsonError&) {
            throw Error::json(std::string("invalid type for `") + key + "`, expected a boolean");
        }
        (std::string(key) == "self_pairs" ? p.self_pairs : p.full_neighbor_list) = b;
    }
    return p;
}

}  // namespace rascaline_amd
