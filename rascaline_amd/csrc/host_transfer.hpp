// Device-to-host transfer of a calculator's results into the float64 arrays of
// host blocks (featomic_calculator_compute's TensorMap output).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <functional>
#include <vector>

namespace rascaline_amd {

// `count` elements of a device buffer, from element `offset`, bound for `dst`.
struct HostSegment {
    long long offset;
    long long count;
    double* dst;
};

// Two pinned host buffers, allocated on first use (hipHostMalloc) and kept by the
// calculator; with their completion events.
class PinnedStaging {
public:
    PinnedStaging() = default;
    PinnedStaging(const PinnedStaging&) = delete;
    PinnedStaging& operator=(const PinnedStaging&) = delete;
    ~PinnedStaging();
    void* buffer(int k) { return buf_[k]; }
    hipEvent_t event(int k) { return ev_[k]; }
    // make both buffers at least `bytes` long
    void reserve(size_t bytes);
    size_t bytes() const { return bytes_; }

private:
    void* buf_[2] = {nullptr, nullptr};
    hipEvent_t ev_[2] = {nullptr, nullptr};
    size_t bytes_ = 0;
};

// Host threads used to convert results (at most 16, or OMP_NUM_THREADS when set).
int host_threads();

// f(begin, end) over [0, n) split between the host threads (inline below `grain`).
void parallel_ranges(size_t n, size_t grain, const std::function<void(size_t, size_t)>& f);

// Copy the segments of `device` (T = float or double) to their host arrays, converting
// to float64: the covered range streams through the pinned buffers in chunks while
// the host threads convert the previous chunk. Returns once every segment is written.
template <typename T>
void download_to_blocks(const T* device, std::vector<HostSegment> segments, hipStream_t stream, PinnedStaging& staging);

}  // namespace rascaline_amd
