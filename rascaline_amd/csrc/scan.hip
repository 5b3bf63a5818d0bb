// Exclusive prefix sums (rocPRIM through hipCUB) used for neighbour-list
// segment offsets, cell-list bins and output block offsets.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>

#include "device.hpp"
#include "calculator.hpp"
#include "kernels.hpp"

namespace rascaline_amd {

namespace {
__global__ void write_total(const long long* in, long long* out, long long n) {
    // out[n] = out[n - 1] + in[n - 1]
    if (n > 0) out[n] = out[n - 1] + in[n - 1];
    else out[0] = 0;
}
}  // namespace

__global__ void fill_zero_kernel(unsigned* p, long long words) {
    for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < words; k += (long long)gridDim.x * blockDim.x) {
        p[k] = 0u;
    }
}

// replay checks: flag |= bit when value[0] != expected
__global__ void check_equal_kernel(const long long* value, long long expected, int bit, int* flag) {
    if (value[0] != expected) atomicOr(flag, bit);
}

__global__ void check_equal_array_kernel(const long long* value, const long long* expected, long long n, int bit,
                                         int* flag) {
    for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < n; k += (long long)gridDim.x * blockDim.x) {
        if (value[k] != expected[k]) atomicOr(flag, bit);
    }
}

__global__ void or_flag_kernel(const int* src, int bit, int* flag) {
    if (src[0] != 0) atomicOr(flag, bit);
}

void fill_zero(void* p, size_t bytes, hipStream_t stream) {
    // bytes is a multiple of 4 for every buffer cleared here
    const long long words = static_cast<long long>(bytes / 4);
    if (words == 0) return;
    const unsigned blocks = static_cast<unsigned>(std::min<long long>((words + 255) / 256, 4096));
    fill_zero_kernel<<<blocks, 256, 0, stream>>>(static_cast<unsigned*>(p), words);
}

void check_equal(hipStream_t stream, const long long* value, long long expected, int bit, int* flag) {
    check_equal_kernel<<<1, 1, 0, stream>>>(value, expected, bit, flag);
}

void check_equal_array(hipStream_t stream, const long long* value, const long long* expected, long long n, int bit,
                       int* flag) {
    if (n <= 0) return;
    const unsigned blocks = static_cast<unsigned>(std::min<long long>((n + 255) / 256, 1024));
    check_equal_array_kernel<<<blocks, 256, 0, stream>>>(value, expected, n, bit, flag);
}

void or_flag(hipStream_t stream, const int* src, int bit, int* flag) { or_flag_kernel<<<1, 1, 0, stream>>>(src, bit, flag); }

void exclusive_scan(hipStream_t stream, const long long* in, long long* out, long long n, Workspace& ws) {
    if (n > 0) {
        size_t temp_bytes = 0;
        RA_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, temp_bytes, in, out, (int)n, stream));
        void* temp = ws.get("scan_temp", temp_bytes);
        RA_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(temp, temp_bytes, in, out, (int)n, stream));
    }
    write_total<<<1, 1, 0, stream>>>(in, out, n);
}

void exclusive_scan(hipStream_t stream, const long long* in, long long* out, long long n) {
    if (n > 0) {
        size_t temp_bytes = 0;
        RA_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, temp_bytes, in, out, (int)n, stream));
        void* temp = nullptr;
        RA_HIP_CHECK(hipMallocAsync(&temp, temp_bytes, stream));
        RA_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(temp, temp_bytes, in, out, (int)n, stream));
        RA_HIP_CHECK(hipFreeAsync(temp, stream));
    }
    write_total<<<1, 1, 0, stream>>>(in, out, n);
}

}  // namespace rascaline_amd
