// Launchers of the LODE kernels (kernels_lode.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "device.hpp"

namespace rascaline_amd {

constexpr int LODE_MAX_L = 16;  // max_angular < 16
constexpr int LODE_MAX_N = 32;  // radial functions

template <typename T>
struct LodeProjectionArgs {
    int n_k, L, n;
    const double* kvec;             // [n_k][4]: direction, norm
    const double* density_fourier;  // [n_k]
    double global_factor;           // 4 pi / V
    DeviceRadial<T> radial;
    const T* sh_a;
    const T* sh_b;
    T* bv[2];                       // per parity: [n_k][ncol_v]
    T* bg[2];                       // per parity: [n_k][3][ncol_v]
    int ncol_v[2];
};

template <typename T>
struct LodeGemmArgs {
    int n_atoms, n_k, n_species, n_types, parity, ncol_v, n;
    const int* types_present;      // [n_types] type indices present in the system
    const int* type_index;         // [n_atoms]
    const int* type_rank;          // [n_atoms] rank among the system's atoms of the same type
    const int* below;              // [n_atoms][n_species] atoms of each type before this one
    const double* cosv;            // [n_atoms][n_k]
    const double* sinv;
    const float* cosf;             // the same phases rounded to FP32 (FP32 pair rows), or null
    const float* sinf;
    const double* sumc;            // [n_species][n_k]
    const double* sums;
    const T* bv;                   // [n_k][ncol_v]
    const T* bg;                   // [n_k][3][ncol_v]
    const int* col_l;              // [ncol_v] l of the column
    const int* col_in_l;           // [ncol_v] m * n + q within the l segment
    T* values;
    const long long* value_base;   // [L][S][S] element offset of each block
    const long long* value_row0;   // [S][S] first sample row of this system in the block
    const T* value_shift;          // [S][S][n] constant l = 0 terms (centre and k = 0)
    T* grads;
    const long long* grad_base;    // [L][S][S]
    const long long* grad_row0;    // [S][S]
    const long long* rows_per_sample;  // [S][S]
};

// Vector-Jacobian product of the position gradients of one system, without the
// N^2 gradient rows: W = dA/dc in the values layout, then
//   V_p[t][x][i](k) = Σ_col Bg_p[k][x][col] W(i, t, col)            (per parity p)
//   U_p[t][x](k)    = Σ_i (cos_i, sin_i)(k) V_p[t][x][i](k)
//   dA/dr_j = Σ_k of the pair terms (from U, minus the j = i term) and the self
//   terms (from the per-type sums and V_p[.][x][j]).
template <typename T>
struct LodeVjpArgs {
    int n_atoms, n_k, n_species, n_types, n;
    const int* types_present;      // [n_types]
    const int* type_index;         // [n_atoms]
    const int* type_rank;          // [n_atoms]
    const double* cosv;            // [n_atoms][n_k]
    const double* sinv;
    const double* sumc;            // [n_species][n_k]
    const double* sums;
    const T* bg[2];                // [n_k][3][ncol_v]
    int ncol_v[2];
    const int* col_l[2];
    const int* col_in_l[2];
    const T* grad_values;          // dA/dc, values layout
    const long long* value_base;   // [L][S][S]
    const long long* value_row0;   // [S][S] this system's first sample row
    T* V;                          // [2][n_types][3][n_atoms][n_k]
    double* U;                     // [2][n_types][3][2][n_k] (cos, sin)
    T* out;                        // [n_atoms][3], this system's atoms
};
template <typename T>
void launch_lode_vjp(hipStream_t stream, const LodeVjpArgs<T>& v);

// cos / sin (k . r_i) in FP64, plus FP32 copies when cosf / sinf are not null
void launch_lode_phases(hipStream_t stream, int n_atoms, int n_k, const double* positions, const double* kvec, double* cosv,
                        double* sinv, float* cosf, float* sinf);
void launch_lode_type_sums(hipStream_t stream, int n_atoms, int n_k, int S, const int* type_index, const double* cosv,
                           const double* sinv, double* sumc, double* sums);
template <typename T>
void launch_lode_projection(hipStream_t stream, const LodeProjectionArgs<T>& p);
// mode 0: values, 1: the center's own gradient rows, 2: neighbour gradient rows
template <typename T>
void launch_lode_gemm(hipStream_t stream, int mode, const LodeGemmArgs<T>& g);

}  // namespace rascaline_amd
