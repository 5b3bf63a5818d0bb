// Key, sample and property selections of featomic_calculator_compute
// (featomic/src/calculator.rs:23-158, 287-496).
#pragma once

#include <functional>
#include <vector>

#include "../../include/featomic.h"
#include "tensormap.hpp"

namespace rascaline_amd {

bool has_selections(const mts_labels_t* selected_keys, const featomic_labels_selection_t& samples,
                    const featomic_labels_selection_t& properties);

// The selected part of a complete result (which this call takes over and releases).
mts_tensormap_t* apply_selections(mts_tensormap_t* full, const mts_labels_t* selected_keys,
                                  const featomic_labels_selection_t& samples,
                                  const featomic_labels_selection_t& properties);

// A result computed in the padded tensor-product shape of an Explicit basis, restricted
// to the basis (which this call takes over and releases): blocks whose o3_lambda is
// not a channel are dropped; "n" properties keep n < n_of(o3_lambda), (l, n_1, n_2)
// properties keep l in `lambdas` with n_1, n_2 < n_of(l)
// (soap/spherical_expansion.rs:128-180, soap/power_spectrum.rs:498-523).
mts_tensormap_t* restrict_to_basis(mts_tensormap_t* full, const std::vector<int>& lambdas,
                                   const std::function<int(int)>& n_of);

}  // namespace rascaline_amd
