// One pass of SOAP calculators over many device batches, captured into a single
// hipGraph (featomic_amd_batch_graph_*).
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <vector>

#include "calculator.hpp"

namespace rascaline_amd {

class BatchGraph {
public:
    // Batch k runs on lanes[k % lanes.size()] and on that lane's stream. The graph
    // keeps its own device copy of every batch's positions, so that each replay has
    // the sizes its record pass measured.
    BatchGraph(std::vector<Calculator*> lanes, const std::vector<HostBatch>& batches, const Gradients& gradients,
               hipStream_t stream);
    ~BatchGraph();
    BatchGraph(const BatchGraph&) = delete;
    BatchGraph& operator=(const BatchGraph&) = delete;

    void launch(hipStream_t stream);
    // waits for the device; throws when a replay saw sizes other than the recorded ones
    // or a kernel reported an error
    void check();
    void output(size_t batch, featomic_amd_device_output_t* out) const;
    size_t n_batches() const { return records_.size(); }
    size_t n_lanes() const { return lanes_.size(); }

private:
    std::vector<Calculator*> lanes_;
    std::vector<hipStream_t> streams_;
    std::vector<hipEvent_t> join_;
    hipEvent_t fork_ = nullptr;
    hipStream_t capture_ = nullptr;  // the stream the graph is captured on (never the null stream)
    std::vector<HostBatch> batches_;
    std::vector<double*> positions_;
    std::vector<std::unique_ptr<BatchRecord>> records_;
    int* flags_ = nullptr;
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t exec_ = nullptr;
    void release();
};

}  // namespace rascaline_amd
