/*
 * Subset of the metatensor-core C API (metatensor-core 0.1, `metatensor.h`)
 * that the featomic C API returns and its callers read: labels, blocks with
 * gradients, tensor maps and the `mts_array_t` data interface.
 *
 * featomic hands results to callers as `mts_tensormap_t` (featomic/include/featomic.h:487).
 * metatensor-core is not available to this build, so librascaline_amd.so exports
 * these functions itself with the same names, argument meaning and struct layout.
 * Linking a process against both this library and a real metatensor-core is
 * untested.
 */
#ifndef RASCALINE_AMD_METATENSOR_SUBSET_H
#define RASCALINE_AMD_METATENSOR_SUBSET_H

#include <stdbool.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MTS_SUCCESS 0
#define MTS_INVALID_PARAMETER_ERROR 1
#define MTS_BUFFER_SIZE_ERROR 254
#define MTS_INTERNAL_ERROR 255

typedef int32_t mts_status_t;
typedef uint64_t mts_data_origin_t;

typedef struct mts_block_t mts_block_t;
typedef struct mts_tensormap_t mts_tensormap_t;

/* metatensor-core: mts_labels_t */
typedef struct mts_labels_t {
    const void* internal_ptr_;
    const char* const* names;
    const int32_t* values;
    uintptr_t size;
    uintptr_t count;
} mts_labels_t;

typedef struct mts_sample_mapping_t {
    uintptr_t input;
    uintptr_t output;
} mts_sample_mapping_t;

/* metatensor-core: mts_array_t (data is always a host float64 array here) */
typedef struct mts_array_t {
    void* ptr;
    mts_status_t (*origin)(const void* array, mts_data_origin_t* origin);
    mts_status_t (*data)(void* array, double** data);
    mts_status_t (*shape)(const void* array, const uintptr_t** shape, uintptr_t* shape_count);
    mts_status_t (*reshape)(void* array, const uintptr_t* shape, uintptr_t shape_count);
    mts_status_t (*swap_axes)(void* array, uintptr_t axis_1, uintptr_t axis_2);
    mts_status_t (*create)(const void* array, const uintptr_t* shape, uintptr_t shape_count, struct mts_array_t* new_array);
    mts_status_t (*copy)(const void* array, struct mts_array_t* new_array);
    void (*destroy)(void* array);
    mts_status_t (*move_samples_from)(void* output, const void* input, const mts_sample_mapping_t* samples,
                                      uintptr_t samples_count, uintptr_t property_start, uintptr_t property_end);
} mts_array_t;

const char* mts_last_error(void);

mts_status_t mts_labels_free(mts_labels_t* labels);

mts_status_t mts_tensormap_free(mts_tensormap_t* tensor);
mts_status_t mts_tensormap_keys(const mts_tensormap_t* tensor, mts_labels_t* keys);
mts_status_t mts_tensormap_block_by_id(mts_tensormap_t* tensor, mts_block_t** block, uintptr_t index);

/* axis 0: samples, 1 .. n-2: components, n-1: properties */
mts_status_t mts_block_labels(const mts_block_t* block, uintptr_t axis, mts_labels_t* labels);
mts_status_t mts_block_data(mts_block_t* block, mts_array_t* data);
mts_status_t mts_block_gradient(mts_block_t* block, const char* parameter, mts_block_t** gradient);
mts_status_t mts_block_gradients_list(const mts_block_t* block, const char* const** parameters, uintptr_t* parameters_count);

/* construction: labels are copied; `data` is taken over (its content copied, then
 * released with its `destroy` callback; an array without `data` callback gives zeros);
 * mts_tensormap takes over the blocks and sets the caller's pointers to NULL */
mts_block_t* mts_block(struct mts_array_t data, mts_labels_t samples, const mts_labels_t* components,
                       uintptr_t components_count, mts_labels_t properties);
mts_status_t mts_block_free(mts_block_t* block);
mts_tensormap_t* mts_tensormap(mts_labels_t keys, mts_block_t** blocks, uintptr_t blocks_count);

#ifdef __cplusplus
}
#endif

#endif
