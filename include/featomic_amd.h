/*
 * Extension entry points of librascaline_amd.so (no counterpart in the
 * reference C API, which only exposes `featomic_calculator_compute` with host
 * systems and a host TensorMap).
 *
 *  - precision control: the reference computes in float64 only; these
 *    calculators can also run their device kernels in float32;
 *  - a device-resident batch entry point for callers whose positions already
 *    live in GPU memory and who consume the descriptor on the GPU (training
 *    loops, benchmarks); outputs stay in library-owned device buffers;
 *  - the neighbour list itself, in the reference's `featomic_pair_t` layout,
 *    for callers that want the pairs (and for parity testing against the
 *    reference's `NeighborsList`, featomic/src/systems/neighbors.rs:349-403).
 */
#ifndef FEATOMIC_AMD_H
#define FEATOMIC_AMD_H

#include <stdint.h>

#include "featomic.h"

#ifdef __cplusplus
extern "C" {
#endif

#define FEATOMIC_AMD_FLOAT64 64
#define FEATOMIC_AMD_FLOAT32 32

/* Select the arithmetic of the device kernels (default FEATOMIC_AMD_FLOAT64).
 * Outputs of featomic_calculator_compute are float64 arrays either way. */
featomic_status_t featomic_amd_calculator_set_precision(featomic_calculator_t* calculator, int32_t bits);

/* Select the fast form (1) or the plain form (0, default) of the FP32 kernels. The
 * fast form runs the power-spectrum values' m-contraction on the matrix cores with
 * bfloat16 hi + lo operand splitting and FP32 accumulation (three products per
 * block, about 16 significant bits per operand; results within 1e-4 of the FP64
 * computation): inside the expansion kernel where it forms the values (8 radial
 * functions, max_angular < 9, at most 16 types), else in the values kernel for n a
 * multiple of 4 and at most 16 types. Its gradient rows take
 * the fastest contraction the autotuner finds among the matrix-core ones: the
 * rank-2 kernel (FP32 operands, n = 8, the default of both forms) or the bfloat16
 * hi + lo tiled kernel (n = 8, 12 or 16). Other settings and FP64 use the plain
 * kernels. */
featomic_status_t featomic_amd_calculator_set_fast_form(featomic_calculator_t* calculator, int32_t enabled);

/* Half neighbour list of one system (pairs sorted by first, second, then
 * cell shift), computed on the GPU. `positions` [n_atoms][3] and `cell` [9]
 * are host arrays; a zero cell means no periodicity. The pairs are allocated
 * by the library; release them with featomic_amd_free. */
featomic_status_t featomic_amd_neighbors(const double* positions, uintptr_t n_atoms, const double* cell, double cutoff,
                                         featomic_pair_t** pairs, uintptr_t* n_pairs);

void featomic_amd_free(void* pointer);

/* A batch of systems whose positions and types are already in device memory. */
typedef struct featomic_amd_device_systems_t {
    uintptr_t n_systems;
    const int64_t* atom_offsets;    /* host, [n_systems + 1] */
    const double* positions;        /* device, [n_atoms][3] float64 */
    const int32_t* types;           /* host, [n_atoms] atomic types */
    const double* cells;            /* host, [n_systems][9] */
} featomic_amd_device_systems_t;

/* Per-block layout of a device-resident result: block b has n_samples[b]
 * rows of `properties` values starting at values + value_offsets[b], and (when
 * positions gradients were requested) n_gradient_rows[b] rows of
 * 3 * properties values starting at gradients + gradient_offsets[b]. For the
 * power spectrum, `properties` is the same for every block; for the spherical
 * expansion one block per (o3_lambda, center_type, neighbor_type) key is
 * reported and `properties` is (2 o3_lambda + 1) * n for that block. Block b's
 * labels are the n_samples[b] sample labels from sample_label_offsets[b] and the
 * n_gradient_rows[b] gradient labels from gradient_label_offsets[b]; the blocks
 * of one spherical-expansion type pair share their labels. The LODE calculator
 * leaves the labels NULL. */
typedef struct featomic_amd_device_output_t {
    uintptr_t n_blocks;
    const int32_t* keys;               /* host, [n_blocks][key size] */
    uintptr_t key_size;
    const int64_t* n_samples;          /* host */
    const int64_t* n_gradient_rows;    /* host */
    const int64_t* value_offsets;      /* host, in elements */
    const int64_t* gradient_offsets;   /* host, in elements */
    const int64_t* block_properties;   /* host, values per sample of each block */
    const void* values;                /* device, float32 or float64 */
    const void* gradients;             /* device, or NULL */
    const int32_t* sample_labels;      /* device, [label samples][2] (system, atom), or NULL */
    const int32_t* gradient_labels;    /* device, [label rows][3] (sample in block, system, atom), or NULL */
    int32_t precision;                 /* FEATOMIC_AMD_FLOAT32 or FEATOMIC_AMD_FLOAT64 */
    const int64_t* sample_label_offsets;    /* host, first sample label of each block, or NULL */
    const int64_t* gradient_label_offsets;  /* host, first gradient label of each block, or NULL */
    /* cell gradients (when requested): block b holds n_samples[b] x [3 (xyz)][3 (abc)] rows of
     * `properties` values from cell_gradients + cell_gradient_offsets[b], one per sample, in
     * the layout of the host TensorMap's "cell" gradient */
    const void* cell_gradients;             /* device, or NULL */
    const int64_t* cell_gradient_offsets;   /* host, in elements, or NULL */
} featomic_amd_device_output_t;

/* Compute the calculator on a device-resident batch. `gradients` may hold
 * "positions" and (SOAP calculators) "cell". The result stays valid until the next call on this calculator
 * or its release. All work is queued on `stream` (a hipStream_t, NULL for the
 * default stream); the call synchronises the stream once after the metadata
 * pass (to size the outputs) and returns without waiting for the kernels that
 * fill values and gradients. */
featomic_status_t featomic_amd_compute_device(featomic_calculator_t* calculator,
                                              const featomic_amd_device_systems_t* systems,
                                              const char* const* gradients, uintptr_t gradients_count, void* stream,
                                              featomic_amd_device_output_t* output);

/* A pass of SOAP calculators (spherical expansion or power spectrum, positions
 * gradients or none) over many device-resident batches, captured into one
 * hipGraph: thousands of structures of uneven sizes go through one graph launch.
 * No counterpart in the reference, whose calculators run on the host; it serves
 * callers that recompute the same batches many times (benchmarks, repeated
 * inference over a fixed set of structures).
 *
 * Creation runs every batch once with host synchronisations (the record pass: the
 * neighbour count, block sizes and host-derived inputs of each batch), then
 * captures a replay of all of them: batch k runs on calculators[k % n_calculators]
 * on a stream of its own, consecutive batches of one calculator in order, the
 * calculators concurrently. The graph keeps its own copy of the batches' positions
 * (taken at creation), so every replay recomputes the same batches, from the
 * neighbour list to the gradient rows, with the recorded sizes. The calculators
 * belong to the graph until it is freed: using them directly fails meanwhile.
 * `stream` orders the copies of the positions. */
typedef struct featomic_amd_batch_graph_t featomic_amd_batch_graph_t;

featomic_status_t featomic_amd_batch_graph_create(featomic_calculator_t* const* calculators, uintptr_t n_calculators,
                                                  const featomic_amd_device_systems_t* batches, uintptr_t n_batches,
                                                  const char* const* gradients, uintptr_t gradients_count, void* stream,
                                                  featomic_amd_batch_graph_t** graph);

/* Launch one replay of the graph on `stream` (asynchronous). */
featomic_status_t featomic_amd_batch_graph_launch(featomic_amd_batch_graph_t* graph, void* stream);

/* Wait for the device, then fail (FEATOMIC_INTERNAL_ERROR) if a replay found
 * sizes other than the recorded ones or a kernel reported an error. */
featomic_status_t featomic_amd_batch_graph_check(featomic_amd_batch_graph_t* graph);

/* Layout and device pointers of batch `batch`'s output, as
 * featomic_amd_compute_device describes it. Batch k shares its calculator's
 * buffers with batches k + n_calculators, k + 2 n_calculators, ...: after a
 * replay, the outputs of the last batch of each calculator are the ones held. */
featomic_status_t featomic_amd_batch_graph_output(const featomic_amd_batch_graph_t* graph, uintptr_t batch,
                                                  featomic_amd_device_output_t* output);

/* Release the graph and its calculators' binding to it. */
featomic_status_t featomic_amd_batch_graph_free(featomic_amd_batch_graph_t* graph);

/* Kernel-level timing of the last featomic_amd_compute_device call, measured
 * with hipEvents on the compute stream: total device time and the time of the
 * dominant output kernel (power spectrum gradient rows / values), in ms.
 * Requires a stream synchronisation, which this call performs. */
featomic_status_t featomic_amd_last_timings(featomic_calculator_t* calculator, double* total_ms, double* main_kernel_ms,
                                            int64_t* main_kernel_bytes);

/* Vector-Jacobian product of the position gradients of a SOAP or LODE calculator
 * (spherical expansion, power spectrum, LODE spherical expansion) on a
 * device-resident batch:
 * positions_grad[atom][3] = Σ_blocks Σ_rows (dX/dr)[row][xyz][:] · grad_values[sample of row][:],
 * the contraction featomic-torch's backward performs with the gradient rows
 * (featomic-torch/src/autograd.cpp:237-300: PositionsGrad::forward), computed
 * here without forming them. `grad_values` (device) holds dA/dX in the layout of
 * the values that featomic_amd_compute_device returns for the same batch and
 * calculator (per-block value_offsets, the calculator's precision);
 * `positions_grad` (device, [n_atoms][3], the calculator's precision) is
 * overwritten. Work is queued on `stream`; the call synchronises it once after the
 * metadata pass (SOAP) or once per system (LODE), like featomic_amd_compute_device.
 * For SOAP, an atom whose neighbour list is not symmetric (which the library's own
 * lists always are) gets NaN. For LODE the N^2 pair rows reduce to O(N n_k) sums
 * over the structure factors. */
featomic_status_t featomic_amd_positions_vjp(featomic_calculator_t* calculator, const featomic_amd_device_systems_t* systems,
                                             const void* grad_values, void* stream, void* positions_grad);

/* Autotuner of the SOAP launch shapes (no counterpart in the reference, whose CPU
 * kernels have no launch shapes). On a device-resident batch, times every
 * candidate window shape (runs, entries, rows per pass) of the tiled
 * gradient-row kernel within the device's whole LDS, the untiled gradient
 * kernel and the built-in choice, then 4 to 64
 * neighbour entries per LDS chunk of the expansion kernel, each by the best time,
 * over `repeats` runs after a warm-up run, of the phase it changes (the gradient
 * rows with their planning pass; the expansion kernel), measured with HIP events. The fastest choices stay in the
 * calculator for its following computations (also when `report` is too small);
 * `report` (may be NULL) receives a JSON summary of every candidate, with
 * FEATOMIC_BUFFER_SIZE_ERROR when `report_size` is too small. */
featomic_status_t featomic_amd_calculator_autotune(featomic_calculator_t* calculator,
                                                  const featomic_amd_device_systems_t* systems,
                                                  const char* const* gradients, uintptr_t gradients_count, void* stream,
                                                  int32_t repeats, char* report, uintptr_t report_size);

/* Launch shapes in use: `expansion_entries` 0 (built-in), a power of two up to
 * 64 (entries per chunk of the vector-ALU expansion kernel), or -8 / -16 (the
 * matrix-core expansion kernel with 8 or 16 entries per chunk, FP32 with 8
 * radial functions and max_angular < 9; elsewhere the vector-ALU kernel);
 * `gradient_window` -1 (built-in), -2 (untiled gradient kernel) or the index
 * of a window shape of the autotuner's list (as its report numbers them). */
featomic_status_t featomic_amd_calculator_set_tuning(featomic_calculator_t* calculator, int32_t expansion_entries,
                                                     int32_t gradient_window);
featomic_status_t featomic_amd_calculator_get_tuning(featomic_calculator_t* calculator, int32_t* expansion_entries,
                                                     int32_t* gradient_window);

/* ---- learned radial integral
 *
 * A network 1 -> 64 -> 64 -> 2 (max_angular + 1) n_radial with SiLU activations
 * maps a pair distance to every R_nl(r) and dR_nl/dr. It takes the place of the
 * reference's SoapRadialIntegral implementations
 * (featomic/src/calculators/soap/radial_integral/mod.rs:19-58: the exact GTO
 * integral of gto.rs:115-205 or its spline of spline.rs:28-70) inside the SOAP
 * calculators. Training runs on the GPU: exact GTO targets in FP64, Adam with
 * FP32 forward / backward kernels, then the output layer solved by ridge least
 * squares in FP64. Inference multiplies FP16 hi + lo operands on the matrix
 * cores with FP32 accumulation. The calculator's spline (or exact integral when
 * `spline_accuracy` is null) judges the network on a uniform grid; the network
 * is used only when every per-(l, n) error is within the tolerance, and the
 * spline stays in use otherwise. */
typedef struct featomic_amd_surrogate_options_t {
    int64_t n_samples;       /* training distances, uniform on [0, cutoff) (default 1000000) */
    int64_t steps;           /* Adam steps (default 20000) */
    int32_t batch;           /* distances per Adam step (default 2048) */
    double learning_rate;    /* peak Adam rate, cosine decay to 1e-3 of it (default 3e-3) */
    double ridge;            /* output-layer ridge, relative to trace(G) / 65 (default 1e-12) */
    double tolerance;        /* acceptance: max |net - judge| / max |judge| per (l, n) (default 1e-3) */
    uint64_t seed;           /* default 1 */
    int64_t n_validation;    /* judge grid points (default 20001) */
} featomic_amd_surrogate_options_t;

typedef struct featomic_amd_surrogate_report_t {
    int32_t accepted;              /* 1 when the network is now in use */
    double adam_loss;              /* mean squared error of the last Adam step (normalised outputs) */
    double max_value_error;        /* over (l, n): max |net - judge| / max |judge| of R_nl */
    double max_derivative_error;   /* same for dR_nl/dr */
    double train_ms;               /* device time of sampling, training, solve and judge */
    double* value_errors;          /* caller buffer [max_angular + 1][n_radial], or NULL */
    double* derivative_errors;     /* caller buffer [max_angular + 1][n_radial], or NULL */
    int64_t errors_size;           /* entries in each caller buffer */
} featomic_amd_surrogate_report_t;

/* Fill `options` with the defaults above. */
featomic_status_t featomic_amd_surrogate_options_default(featomic_amd_surrogate_options_t* options);

/* Train the radial surrogate of a SOAP calculator (spherical_expansion or
 * soap_power_spectrum) and use it for later computations when it is accepted.
 * `report` may be NULL. */
featomic_status_t featomic_amd_calculator_train_radial_surrogate(featomic_calculator_t* calculator,
                                                                 const featomic_amd_surrogate_options_t* options,
                                                                 featomic_amd_surrogate_report_t* report);

/* Use (1) the accepted surrogate or (0) the spline / exact radial integral. */
featomic_status_t featomic_amd_calculator_use_radial_surrogate(featomic_calculator_t* calculator, int32_t enabled);

/* The trained network's FP32 parameters, flat: W1[64], b1[64], W2[64][64],
 * b2[64], W3[n_out][64], b3[n_out], then the n_out output scales; `count` must
 * be at least *required (set on return; call with params NULL to query). */
featomic_status_t featomic_amd_calculator_radial_surrogate_parameters(featomic_calculator_t* calculator, float* params,
                                                                      int64_t count, int64_t* required);

/* One forward + backward pass of the device training kernels (FP32) over
 * host arrays: inputs [n_samples], targets [n_samples][n_out]; writes the mean
 * squared error and its gradient with respect to the flat parameters
 * (layout above, without the scales). */
featomic_status_t featomic_amd_mlp_loss_gradient(int32_t n_out, int64_t n_samples, const float* params,
                                                 const float* inputs, const float* targets, double* loss,
                                                 float* gradient);

/* The same pass as featomic_amd_mlp_loss_gradient on device-resident arrays, with
 * no device allocation. No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, serving loops that
 * forbid allocations). `params` [n_params], `inputs` [n_samples] and `targets`
 * [n_samples][n_out] are device arrays; the mean squared error goes to `loss` (one
 * double, device) and its gradient to `gradient` ([n_params] floats, device). The
 * work is queued on `stream` (a hipStream_t, NULL for the default stream) and the
 * call returns without waiting. The scratch (transposed weights, one partial
 * gradient per tile of 8 to 32 samples) is a static 32 MiB device array that every
 * GPU the library runs on holds until the process ends; calls whose scratch would
 * exceed it (about 14 000 samples at n_out = 216) return
 * FEATOMIC_INVALID_PARAMETER_ERROR. Every call on a GPU shares that array: calls that
 * overlap on one GPU (different streams or threads) race. */
featomic_status_t featomic_amd_mlp_loss_gradient_noalloc(int32_t n_out, int64_t n_samples, const float* params,
                                                         const float* inputs, const float* targets, double* loss,
                                                         float* gradient, void* stream);

/* The device inference kernel (FP16 hi + lo on the matrix cores, FP32
 * accumulation) over host inputs: outputs [n_samples][n_out]. */
featomic_status_t featomic_amd_mlp_forward(int32_t n_out, int64_t n_samples, const float* params, const float* inputs,
                                           float* outputs);

#ifdef __cplusplus
}
#endif

#endif
