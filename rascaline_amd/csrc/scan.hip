// Exclusive prefix sums (rocPRIM through hipCUB) used for neighbour-list
// segment offsets, cell-list bins and output block offsets.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "device.hpp"
#include "kernels.hpp"

namespace rascaline_amd {

namespace {
__global__ void write_total(const long long* in, long long* out, long long n) {
    // out[n] = out[n - 1] + in[n - 1]
    if (n > 0) out[n] = out[n - 1] + in[n - 1];
    else out[0] = 0;
}
}  // namespace

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
