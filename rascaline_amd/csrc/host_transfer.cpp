// Chunked, pinned device-to-host transfer with multi-threaded float64 conversion.
#include "host_transfer.hpp"

#include <algorithm>
#include <cstdlib>
#include <thread>

#include "device.hpp"

namespace rascaline_amd {

namespace {
// bytes of device data per staged chunk
constexpr size_t CHUNK_BYTES = 64u << 20;
}  // namespace

PinnedStaging::~PinnedStaging() {
    for (int k = 0; k < 2; k++) {
        if (buf_[k]) (void)hipHostFree(buf_[k]);
        if (ev_[k]) (void)hipEventDestroy(ev_[k]);
    }
}

void PinnedStaging::reserve(size_t bytes) {
    if (!ev_[0]) {
        RA_HIP_CHECK(hipEventCreateWithFlags(&ev_[0], hipEventDisableTiming));
        RA_HIP_CHECK(hipEventCreateWithFlags(&ev_[1], hipEventDisableTiming));
    }
    if (bytes <= bytes_) return;
    for (int k = 0; k < 2; k++) {
        if (buf_[k]) {
            RA_HIP_CHECK(hipHostFree(buf_[k]));
            buf_[k] = nullptr;
        }
        RA_HIP_CHECK(hipHostMalloc(&buf_[k], bytes, hipHostMallocDefault));
    }
    bytes_ = bytes;
}

int host_threads() {
    static const int threads = [] {
        int t = static_cast<int>(std::thread::hardware_concurrency());
        if (const char* env = std::getenv("OMP_NUM_THREADS")) {
            const int e = std::atoi(env);
            if (e > 0) t = e;
        }
        return std::max(1, std::min(t, 16));
    }();
    return threads;
}

void parallel_ranges(size_t n, size_t grain, const std::function<void(size_t, size_t)>& f) {
    const size_t nt = std::min<size_t>(static_cast<size_t>(host_threads()), grain ? (n + grain - 1) / grain : 1);
    if (nt <= 1) {
        if (n) f(0, n);
        return;
    }
    std::vector<std::thread> pool;
    pool.reserve(nt - 1);
    for (size_t k = 1; k < nt; k++) pool.emplace_back(f, k * n / nt, (k + 1) * n / nt);
    f(0, n / nt);
    for (auto& t : pool) t.join();
}

template <typename T>
void download_to_blocks(const T* device, std::vector<HostSegment> segments, hipStream_t stream, PinnedStaging& staging) {
    segments.erase(std::remove_if(segments.begin(), segments.end(), [](const HostSegment& s) { return s.count <= 0; }),
                   segments.end());
    if (segments.empty()) return;
    std::sort(segments.begin(), segments.end(),
              [](const HostSegment& a, const HostSegment& b) { return a.offset < b.offset; });
    const long long lo = segments.front().offset;
    long long hi = lo;
    for (const auto& s : segments) hi = std::max(hi, s.offset + s.count);
    const long long chunk = static_cast<long long>(CHUNK_BYTES / sizeof(T));
    staging.reserve(CHUNK_BYTES);
    const long long n_chunks = (hi - lo + chunk - 1) / chunk;

    auto issue = [&](long long c) {
        const long long a = lo + c * chunk, b = std::min(hi, a + chunk);
        const int k = static_cast<int>(c & 1);
        RA_HIP_CHECK(hipMemcpyAsync(staging.buffer(k), device + a, sizeof(T) * static_cast<size_t>(b - a),
                                    hipMemcpyDeviceToHost, stream));
        RA_HIP_CHECK(hipEventRecord(staging.event(k), stream));
    };
    issue(0);
    if (n_chunks > 1) issue(1);
    for (long long c = 0; c < n_chunks; c++) {
        const int k = static_cast<int>(c & 1);
        RA_HIP_CHECK(hipEventSynchronize(staging.event(k)));
        const long long a = lo + c * chunk, b = std::min(hi, a + chunk);
        const T* src = static_cast<const T*>(staging.buffer(k));
        parallel_ranges(static_cast<size_t>(b - a), 1 << 20, [&](size_t x0, size_t x1) {
            const long long ra = a + static_cast<long long>(x0), rb = a + static_cast<long long>(x1);
            // first segment ending after ra
            auto it = std::upper_bound(segments.begin(), segments.end(), ra,
                                       [](long long v, const HostSegment& s) { return v < s.offset + s.count; });
            for (; it != segments.end() && it->offset < rb; ++it) {
                const long long s0 = std::max(ra, it->offset), s1 = std::min(rb, it->offset + it->count);
                double* dst = it->dst + (s0 - it->offset);
                const T* in = src + (s0 - a);
                for (long long x = 0; x < s1 - s0; x++) dst[x] = static_cast<double>(in[x]);
            }
        });
        // the buffer is free again: stream the chunk after next into it
        if (c + 2 < n_chunks) issue(c + 2);
    }
}

template void download_to_blocks<float>(const float*, std::vector<HostSegment>, hipStream_t, PinnedStaging&);
template void download_to_blocks<double>(const double*, std::vector<HostSegment>, hipStream_t, PinnedStaging&);

}  // namespace rascaline_amd
