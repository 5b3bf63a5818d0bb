// Host-side launchers of the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

#include "device.hpp"

namespace rascaline_amd {

// ---- neighbour list and metadata (kernels_nl.hip)
void launch_nl_assign_bins(hipStream_t stream, int n_atoms, const double* positions, const int* atom_system,
                           const NLSystem* systems, long long* atom_bin, int* atom_wrap, int* bin_count);
void launch_nl_fill_bins(hipStream_t stream, int n_atoms, const long long* atom_bin, const long long* bin_start,
                         int* bin_fill, int* bin_atoms);
// counting pass (emit = false) and emitting pass (emit = true); with `stash`
// (nl_stash_keys(n_atoms) keys), the counting pass keeps each center's first keys and
// the emitting pass copies them instead of searching again
size_t nl_stash_keys(int n_atoms);
void launch_nl_candidates(hipStream_t stream, bool emit, int n_atoms, const double* positions, const int* atom_system,
                          const NLSystem* systems, const long long* atom_bin, const int* atom_wrap,
                          const long long* bin_start, const int* bin_len, const int* bin_atoms, double cutoff2,
                          long long* counts, const long long* offsets, unsigned long long* keys, int* error_flag,
                          unsigned long long* stash = nullptr);
void launch_nl_sort_segments(hipStream_t stream, int n_atoms, const long long* offsets, unsigned long long* keys,
                             int* error_flag);
void launch_nl_finalize(hipStream_t stream, long long n_entries, int n_atoms, const long long* offsets,
                        const unsigned long long* keys, const double* positions, const int* atom_system,
                        const NLSystem* systems, int* entry_center, int* entry_j, int* entry_shift, double* entry_vec,
                        double* entry_dist);
void launch_center_meta(hipStream_t stream, int n_atoms, int n_species, const long long* offsets, const int* entry_j,
                        const int* species, unsigned* type_mask, int* unique_count);
void launch_block_membership(hipStream_t stream, int n_atoms, int n_species, int ps_mode, int n_blocks,
                             const int* species, const unsigned* type_mask, const int* unique_count,
                             const int* block_types, long long* sample_flag, long long* row_count);
void launch_gather_block_totals(hipStream_t stream, int n_blocks, long long n_atoms, const long long* sample_scan,
                                const long long* row_scan, long long* totals);
void launch_sample_labels(hipStream_t stream, int n_blocks, int n_atoms, const long long* sample_flag,
                          const long long* sample_scan, const int* atom_system, const int* atom_local, int* labels);

// exclusive prefix sum of `n` values into out[0..n] (out[n] = total)
void exclusive_scan(hipStream_t stream, const long long* in, long long* out, long long n);
// the same, with its temporary storage from `ws` (no stream-ordered allocation: the
// form a captured graph can hold)
class Workspace;
void exclusive_scan(hipStream_t stream, const long long* in, long long* out, long long n, Workspace& ws);
// zero `bytes` (a multiple of 4) with a kernel rather than hipMemsetAsync (graph-safe)
void fill_zero(void* p, size_t bytes, hipStream_t stream);
// replay consistency checks: atomically OR `bit` into *flag when value[0] != expected,
// when value[k] != expected[k] for some k < n, or when *src != 0
void check_equal(hipStream_t stream, const long long* value, long long expected, int bit, int* flag);
void check_equal_array(hipStream_t stream, const long long* value, const long long* expected, long long n, int bit,
                       int* flag);
void or_flag(hipStream_t stream, const int* src, int bit, int* flag);

// ---- SOAP (kernels_soap.hip)
template <typename T>
struct SoapLaunch {
    hipStream_t stream;
    int n_atoms, n_species, L, n;
    const long long* offsets;
    const int* entry_j;
    const int* entry_shift;
    const double* entry_vec;
    const double* entry_dist;
    const int* species;
    DeviceRadial<T> radial;
    DeviceSoapScalars scalars;
    const T* sh_a;
    const T* sh_b;
    const T* sh_c;
    const T* self_contribution;
    int chunk;
    int* error_flag;
};

template <typename T>
size_t soap_chunk_bytes(int L, int n, int chunk);

// One sample of spherical_expansion_by_pair: its neighbour entry in the full list
// (-1: a self pair), its type pair, its row in the l = 0 block and in the l > 0
// blocks (-1 for a self pair), its first position-gradient row and their count
// (0: self pair, 1: an atom and its own periodic image, 2).
struct ByPairSample {
    long long entry;
    int type_pair;
    int row0;
    int row1;
    int grad_row;
    int grad_rows;
    int pad;
};
// values / gradient rows of every sample into the blocks (type pair, l), whose
// element offsets in each output are *_base[type_pair * L + l]; null outputs are skipped
template <typename T>
void launch_by_pair(const SoapLaunch<T>& p, long long n_samples, const ByPairSample* samples, const long long* value_base,
                    const long long* pos_base, const long long* strain_base, const long long* cell_base, T* values, T* pos,
                    T* strain, T* cell);
template <typename T>
void launch_radial_exact(hipStream_t stream, long long n_entries, int L, int n, const double* entry_dist,
                         double sigma2_atomic, const double* gto, const double* ortho, T* out_val, T* out_der);
// radial integral at `count` distances, [i][L][n] (spline tables, or the
// precomputed per-entry arrays with entry index i)
template <typename T>
void launch_radial_eval(hipStream_t stream, long long count, const DeviceRadial<T>& radial, const double* r,
                        T* out_val, T* out_der);
template <typename T>
void launch_soap_center(const SoapLaunch<T>& p, int mode, T* coef, T* grad);
// vector-Jacobian product of the position gradients (no gradient rows): H = dA/dc
// from dA/dPS (power spectrum) or from the per-block dA/dc (spherical expansion),
// g[entry][3] per neighbour entry, then dA/dr per atom
template <typename T>
size_t ps_vjp_coef_shmem(const SoapLaunch<T>& p);
template <typename T>
void launch_ps_vjp_coef(const SoapLaunch<T>& p, const T* coef, const unsigned* type_mask, const long long* sample_flag,
                        const long long* sample_scan, const T* grad_values, T* H);
template <typename T>
void launch_se_vjp_coef(const SoapLaunch<T>& p, const long long* sample_flag, const long long* sample_scan,
                        const long long* block_sample_base, const long long* value_base, const T* grad_values, T* H);
template <typename T>
size_t vjp_pair_shmem(const SoapLaunch<T>& p, int max_entries);
template <typename T>
void launch_vjp_pair(const SoapLaunch<T>& p, int max_entries, const T* entry_data, const T* H, T* g);
template <typename T>
void launch_vjp_gather(hipStream_t stream, int n_atoms, const long long* offsets, const int* entry_j,
                       const int* entry_shift, const T* g, T* out);
// fast (FP32 only): the m-contraction in bfloat16 hi + lo on MFMA when n is a
// multiple of 4, at most 16 types and the per-center Gram matrix fits in lds_limit
template <typename T>
void launch_ps_values(const SoapLaunch<T>& p, const T* coef, const unsigned* type_mask, const long long* sample_flag,
                      const long long* sample_scan, T* out, bool fast, int lds_limit);
template <typename T>
size_t gradient_rows_shmem(const SoapLaunch<T>& p, bool ps);
template <typename T>
void launch_gradient_rows(const SoapLaunch<T>& p, bool ps, const T* coef, const T* dself, const unsigned* type_mask,
                          const long long* sample_flag, const long long* sample_scan, const long long* row_scan,
                          const long long* block_sample_base, const long long* se_grad_base, const int* atom_system,
                          const int* atom_local, T* out, int* labels);
template <typename T>
size_t ps_gradient_tiled_shmem(const SoapLaunch<T>& p, int max_runs, int max_entries, int rows_per_pass);
// dynamic plus static LDS per workgroup of the gradient-row kernel in use
template <typename T>
size_t ps_gradient_tiled_lds(const SoapLaunch<T>& p, int max_runs, int max_entries, int rows_per_pass);
bool ps_gradient_tiled_supported(int L, int n, int n_species, int max_runs, int max_entries, int rows_per_pass);
// Planning pass of the tiled gradient kernel; runs / windows are int4 arrays with
// one slot per neighbour entry, run_rows [entries][n_species], self_rows
// [n_atoms][n_species (n_species + 1) / 2].
void launch_ps_grad_plan(hipStream_t stream, int n_atoms, int n_species, const long long* offsets, const int* entry_j,
                         const int* species, const unsigned* type_mask, const long long* sample_scan,
                         const long long* row_scan, const long long* block_sample_base, const int* atom_system,
                         const int* atom_local, int max_runs, int max_entries, int* runs, int* run_rows, int* windows,
                         int* n_windows, int* self_rows, int* labels, bool power_spectrum, int* split = nullptr);
// spherical-expansion gradient rows from the same plan (power_spectrum = false
// above: one row per run in run_rows[run * S], self rows at self_rows[i * S + t])
template <typename T>
size_t se_gradient_tiled_shmem(const SoapLaunch<T>& p, int max_runs, int max_entries);
template <typename T>
void launch_se_gradient_tiled(const SoapLaunch<T>& p, int max_runs, int max_entries, const T* dself,
                              const unsigned* type_mask, const int* runs, const int* run_rows, const int* windows,
                              const int* n_windows, const int* self_rows, const long long* se_grad_base,
                              const T* entry_data, T* out);
// Per-entry data (Legendre functions, direction, trigonometric and cutoff factors,
// radial values) of the tiled gradient kernel: entry_prep_size values.
template <typename T>
size_t entry_prep_size(const SoapLaunch<T>& p, long long n_entries);
template <typename T>
void launch_entry_prep(const SoapLaunch<T>& p, long long n_entries, T* data);
// c_nlm and self gradients from the per-entry data (n <= 16, l_max < 16)
template <typename T>
size_t expansion_tiled_shmem(const SoapLaunch<T>& p, int max_entries);
template <typename T>
void launch_expansion_tiled(const SoapLaunch<T>& p, int max_entries, const T* entry_data, T* coef, T* dself);
// the same on the matrix cores (FP32, n = 8, l_max < 9), in chunks of 8 or 16 entries;
// with `ps_values` (at most 16 types), it also writes the power-spectrum values of the
// samples that `sample_flag` / `sample_scan` place (as launch_ps_values does), with the
// products over m in FP32 or, with `bf16_values`, in bfloat16 hi + lo
bool expansion_mfma_supported(int L, int n, int n_species);
bool expansion_mfma_values_supported(int n_species);
size_t expansion_mfma_lds(const SoapLaunch<float>& p, int max_entries, bool values);
void launch_expansion_mfma(const SoapLaunch<float>& p, int max_entries, const float* entry_data, float* coef,
                           float* dself, const unsigned* type_mask = nullptr, const long long* sample_flag = nullptr,
                           const long long* sample_scan = nullptr, float* ps_values = nullptr,
                           bool bf16_values = false);
template <typename T>
void launch_ps_gradient_tiled(const SoapLaunch<T>& p, int max_runs, int max_entries, int rows_per_pass, const T* coef,
                              const T* dself, const unsigned* type_mask, const int* runs, const int* run_rows,
                              const int* windows, const int* n_windows, const int* self_rows, const T* entry_data,
                              T* out, bool fast, const int* only = nullptr);
// Rank-2 gradient rows on MFMA (FP32, n = 8) over a plan of whole-run windows of up
// to `max_entries` (8 or 16) images and 8 runs; the centers the plan flags in `split`
// (a run longer than a window) are left to launch_ps_gradient_tiled with only = split.
bool ps_gradient_rank2_supported(int L, int n, int n_species, int max_entries);
size_t ps_gradient_rank2_lds(const SoapLaunch<float>& p, int max_entries);
void launch_ps_gradient_rank2(const SoapLaunch<float>& p, int max_entries, const float* coef, const float* dself,
                              const unsigned* type_mask, const int* runs, const int* run_rows, const int* windows,
                              const int* n_windows, const int* self_rows, const int* split, const float* entry_data,
                              float* out);
template <typename T>
size_t ps_gradient_shmem(const SoapLaunch<T>& p, int chunk);
// register-blocked power-spectrum gradient rows; returns false when n > 16
template <typename T>
bool launch_ps_gradient_fast(const SoapLaunch<T>& p, int chunk, const T* coef, const T* dself, const unsigned* type_mask,
                             const long long* sample_scan, const long long* row_scan, const long long* block_sample_base,
                             const int* atom_system, const int* atom_local, T* out, int* labels);
template <typename T>
void launch_ps_cell_strain(const SoapLaunch<T>& p, const T* coef, const T* dcs, const unsigned* type_mask,
                           const long long* sample_flag, const long long* sample_scan, T* out);
template <typename T>
void launch_se_gather(const SoapLaunch<T>& p, int ncomp, const T* dense, const long long* sample_flag,
                      const long long* sample_scan, const long long* block_sample_base, const long long* region_base, T* out);

}  // namespace rascaline_amd
