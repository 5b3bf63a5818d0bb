// Device-side data structures shared by the host orchestration and the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

#include "hypers.hpp"

namespace rascaline_amd {

#define RA_HIP_CHECK(expr)                                                                          \
    do {                                                                                            \
        hipError_t err_ = (expr);                                                                   \
        if (err_ != hipSuccess) {                                                                   \
            throw ::rascaline_amd::Error::internal(std::string("HIP error in ") + #expr + ": " +     \
                                                   hipGetErrorString(err_));                        \
        }                                                                                           \
    } while (0)

// Per-system cell-list parameters (featomic/src/systems/neighbors.rs:101-163).
struct NLSystem {
    double cell[9];      // rows are the cell vectors
    double inv[9];       // inverse cell (Matrix3::inverse op order)
    int n_cells[3];
    int n_search[3];
    int periodic;
    int n_atoms;
    long long bin_offset;   // first bin of this system in the global bin array
    long long atom_offset;  // first atom of this system in the batch
};

// Radial integral description for the device (all pointers are device memory).
template <typename T>
struct DeviceRadial {
    int max_angular;
    int n;
    int use_spline;
    double stop;
    double sigma2_atomic;      // SOAP exact mode
    const double* gto;         // exact mode: SOAP 1/(2 sigma_n^2), LODE sigma_n
    const double* ortho;       // exact mode: n x n
    const int* knot_offset;    // per l: first knot
    const int* n_knots;        // per l: number of knots
    const double* knot_grid;   // per l: first knot, (knots - 1) / (last - first knot), as the kernels round them
    const T* knot_pos;         // [total knots]
    const T* knot_val;         // [total knots][n]
    const T* knot_der;         // [total knots][n]
    // exact mode: radial values / derivatives precomputed per neighbour entry,
    // [entry][L][n] (filled by launch_radial_exact before the SOAP kernels)
    const T* exact_val;
    const T* exact_der;
};

// SOAP per-pair scalar functions (cutoff smoothing + Willatt2018 scaling).
struct DeviceSoapScalars {
    double cutoff;
    int smoothing;             // 0 step, 1 shifted cosine
    double width;
    int has_scaling;
    double scale, rate, exponent;
};

}  // namespace rascaline_amd
