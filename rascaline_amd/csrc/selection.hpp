// Key, sample and property selections of featomic_calculator_compute
// (featomic/src/calculator.rs:23-158, 287-496).
#pragma once

#include "../../include/featomic.h"
#include "tensormap.hpp"

namespace rascaline_amd {

bool has_selections(const mts_labels_t* selected_keys, const featomic_labels_selection_t& samples,
                    const featomic_labels_selection_t& properties);

// The selected part of a complete result (which this call takes over and releases).
mts_tensormap_t* apply_selections(mts_tensormap_t* full, const mts_labels_t* selected_keys,
                                  const featomic_labels_selection_t& samples,
                                  const featomic_labels_selection_t& properties);

}  // namespace rascaline_amd
