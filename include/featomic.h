/*
 * featomic C API, as exported by librascaline_amd.so.
 *
 * Same names, struct layouts, status codes and calling conventions as the
 * reference header (featomic/include/featomic.h); each declaration below cites
 * the line of that header it replaces. The calculators with the names
 * "spherical_expansion", "soap_power_spectrum" and "lode_spherical_expansion"
 * run on the GPU (HIP, gfx950).
 */
#ifndef FEATOMIC_H
#define FEATOMIC_H

#include <stdbool.h>
#include <stdint.h>

#include "metatensor.h"

/* status codes, featomic.h:19-51 */
#define FEATOMIC_SUCCESS 0
#define FEATOMIC_INVALID_PARAMETER_ERROR 1
#define FEATOMIC_JSON_ERROR 2
#define FEATOMIC_UTF8_ERROR 3
#define FEATOMIC_SYSTEM_ERROR 128
#define FEATOMIC_BUFFER_SIZE_ERROR 254
#define FEATOMIC_INTERNAL_ERROR 255

/* log levels, featomic.h:53-77 */
#define FEATOMIC_LOG_LEVEL_ERROR 1
#define FEATOMIC_LOG_LEVEL_WARN 2
#define FEATOMIC_LOG_LEVEL_INFO 3
#define FEATOMIC_LOG_LEVEL_DEBUG 4
#define FEATOMIC_LOG_LEVEL_TRACE 5

typedef struct featomic_calculator_t featomic_calculator_t;

/* featomic.h:98 */
typedef int32_t featomic_status_t;

/* featomic.h:105 */
typedef void (*featomic_logging_callback_t)(int32_t level, const char* message);

/* featomic.h:114-139 (64 bytes) */
typedef struct featomic_pair_t {
    uintptr_t first;
    uintptr_t second;
    double distance;
    double vector[3];
    int32_t cell_shift_indices[3];
} featomic_pair_t;

/* featomic.h:167-241 */
typedef struct featomic_system_t {
    void* user_data;
    featomic_status_t (*size)(const void* user_data, uintptr_t* size);
    featomic_status_t (*types)(const void* user_data, const int32_t** types);
    featomic_status_t (*positions)(const void* user_data, const double** positions);
    featomic_status_t (*cell)(const void* user_data, double* cell);
    featomic_status_t (*compute_neighbors)(void* user_data, double cutoff);
    featomic_status_t (*pairs)(const void* user_data, const struct featomic_pair_t** pairs, uintptr_t* count);
    featomic_status_t (*pairs_containing)(const void* user_data, uintptr_t atom, const struct featomic_pair_t** pairs,
                                          uintptr_t* count);
} featomic_system_t;

/* featomic.h:249-264 */
typedef struct featomic_labels_selection_t {
    const mts_labels_t* subset;
    const mts_tensormap_t* predefined;
} featomic_labels_selection_t;

/* featomic.h:271-317 */
typedef struct featomic_calculation_options_t {
    const char* const* gradients;
    uintptr_t gradients_count;
    bool use_native_system;
    struct featomic_labels_selection_t selected_samples;
    struct featomic_labels_selection_t selected_properties;
    const mts_labels_t* selected_keys;
} featomic_calculation_options_t;

#ifdef __cplusplus
extern "C" {
#endif

/* featomic.h:361 */
const char* featomic_last_error(void);

/* featomic.h:368 */
featomic_status_t featomic_set_logging_callback(featomic_logging_callback_t callback);

/* featomic.h:392 */
struct featomic_calculator_t* featomic_calculator(const char* name, const char* parameters);

/* featomic.h:406 */
featomic_status_t featomic_calculator_free(struct featomic_calculator_t* calculator);

/* featomic.h:424 */
featomic_status_t featomic_calculator_name(const struct featomic_calculator_t* calculator, char* name, uintptr_t bufflen);

/* featomic.h:445 */
featomic_status_t featomic_calculator_parameters(const struct featomic_calculator_t* calculator, char* parameters,
                                                 uintptr_t bufflen);

/* featomic.h:465 */
featomic_status_t featomic_calculator_cutoffs(const struct featomic_calculator_t* calculator, const double** cutoffs,
                                              uintptr_t* cutoffs_count);

/* featomic.h:487 */
featomic_status_t featomic_calculator_compute(struct featomic_calculator_t* calculator, mts_tensormap_t** descriptor,
                                              struct featomic_system_t* systems, uintptr_t systems_count,
                                              struct featomic_calculation_options_t options);

/* featomic.h:502 */
featomic_status_t featomic_profiling_clear(void);

/* featomic.h:522 */
featomic_status_t featomic_profiling_enable(bool enabled);

/* featomic.h:540 */
featomic_status_t featomic_profiling_get(const char* format, char* buffer, uintptr_t bufflen);

#ifdef __cplusplus
}
#endif

#endif /* FEATOMIC_H */
