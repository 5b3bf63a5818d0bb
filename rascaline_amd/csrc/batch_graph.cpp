// Record, capture and replay of a pass over many device batches.
//
// The record pass runs every batch once with host synchronisations, keeping its
// neighbour count, block sizes and host-derived inputs (BatchRecord); the lanes'
// workspaces grow to the largest batch. The capture pass then replays every batch
// with the workspaces frozen: no synchronisation, allocation, upload or event record
// reaches the stream, and the kernels of all lanes are captured into one graph
// (lane streams forked from and joined back to the capturing stream). A replay
// recomputes everything from the graph's copy of the positions: neighbour lists,
// expansion, values and gradient rows.
#include "batch_graph.hpp"

#include <string>

#include "device.hpp"
#include "kernels.hpp"

namespace rascaline_amd {

BatchGraph::BatchGraph(std::vector<Calculator*> lanes, const std::vector<HostBatch>& batches, const Gradients& gradients,
                       hipStream_t stream)
    : lanes_(std::move(lanes)) {
    if (lanes_.empty()) throw Error::invalid_parameter("a batch graph needs at least one calculator");
    for (size_t a = 0; a < lanes_.size(); a++) {
        for (size_t b = 0; b < a; b++) {
            if (lanes_[a] == lanes_[b]) throw Error::invalid_parameter("the calculators of a batch graph must be distinct");
        }
        if (lanes_[a]->bound_to_graph) throw Error::invalid_parameter("a calculator can belong to one batch graph only");
        if (lanes_[a]->kind() != Kind::PowerSpectrum && lanes_[a]->kind() != Kind::SphericalExpansion) {
            throw Error::invalid_parameter("batch graphs support the SOAP spherical expansion and power spectrum");
        }
    }
    if (gradients.strain || gradients.cell) throw Error::invalid_parameter("batch graphs support positions gradients only");
    try {
        RA_HIP_CHECK(hipMalloc(&flags_, sizeof(int)));
        fill_zero(flags_, sizeof(int), stream);
        RA_HIP_CHECK(hipEventCreateWithFlags(&fork_, hipEventDisableTiming));
        RA_HIP_CHECK(hipStreamCreateWithFlags(&capture_, hipStreamNonBlocking));
        for (size_t l = 0; l < lanes_.size(); l++) {
            hipStream_t s = nullptr;
            RA_HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
            streams_.push_back(s);
            hipEvent_t e = nullptr;
            RA_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            join_.push_back(e);
        }
        // the graph's own positions
        for (const HostBatch& b : batches) {
            HostBatch copy = b;
            double* p = nullptr;
            const size_t bytes = sizeof(double) * 3 * static_cast<size_t>(std::max<int64_t>(b.n_atoms(), 1));
            RA_HIP_CHECK(hipMalloc(&p, bytes));
            positions_.push_back(p);
            if (b.n_atoms() > 0) {
                if (!b.device_positions) throw Error::invalid_parameter("batch graphs need device-resident positions");
                RA_HIP_CHECK(hipMemcpyAsync(p, b.device_positions, sizeof(double) * 3 * b.n_atoms(),
                                            hipMemcpyDeviceToDevice, stream));
            }
            copy.device_positions = p;
            batches_.push_back(std::move(copy));
            records_.push_back(std::make_unique<BatchRecord>());
        }
        RA_HIP_CHECK(hipStreamSynchronize(stream));

        // record pass (synchronising), lane by lane on the lanes' streams
        for (size_t k = 0; k < batches_.size(); k++) {
            Calculator& c = *lanes_[k % lanes_.size()];
            featomic_amd_device_output_t out{};
            c.set_pass(PassMode::Record, records_[k].get(), nullptr);
            c.compute_device(batches_[k], gradients, streams_[k % lanes_.size()], &out);
            c.set_pass(PassMode::Eager, nullptr, nullptr);
        }
        for (hipStream_t s : streams_) RA_HIP_CHECK(hipStreamSynchronize(s));

        // capture pass
        for (Calculator* c : lanes_) c->bound_to_graph = true;
        RA_HIP_CHECK(hipStreamBeginCapture(capture_, hipStreamCaptureModeRelaxed));
        try {
            RA_HIP_CHECK(hipEventRecord(fork_, capture_));
            for (hipStream_t s : streams_) RA_HIP_CHECK(hipStreamWaitEvent(s, fork_, 0));
            for (size_t k = 0; k < batches_.size(); k++) {
                Calculator& c = *lanes_[k % lanes_.size()];
                featomic_amd_device_output_t out{};
                c.set_pass(PassMode::Replay, records_[k].get(), flags_);
                c.compute_device(batches_[k], gradients, streams_[k % lanes_.size()], &out);
                c.set_pass(PassMode::Eager, nullptr, nullptr);
            }
            for (size_t l = 0; l < streams_.size(); l++) {
                RA_HIP_CHECK(hipEventRecord(join_[l], streams_[l]));
                RA_HIP_CHECK(hipStreamWaitEvent(capture_, join_[l], 0));
            }
        } catch (...) {
            for (Calculator* c : lanes_) c->set_pass(PassMode::Eager, nullptr, nullptr);
            hipGraph_t broken = nullptr;
            (void)hipStreamEndCapture(capture_, &broken);
            if (broken) (void)hipGraphDestroy(broken);
            throw;
        }
        RA_HIP_CHECK(hipStreamEndCapture(capture_, &graph_));
        RA_HIP_CHECK(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0));
    } catch (...) {
        release();
        throw;
    }
}

BatchGraph::~BatchGraph() { release(); }

void BatchGraph::release() {
    if (exec_) (void)hipGraphExecDestroy(exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    exec_ = nullptr;
    graph_ = nullptr;
    (void)hipDeviceSynchronize();
    for (Calculator* c : lanes_) {
        c->bound_to_graph = false;
        c->set_pass(PassMode::Eager, nullptr, nullptr);
    }
    for (double* p : positions_) (void)hipFree(p);
    positions_.clear();
    for (hipStream_t s : streams_) (void)hipStreamDestroy(s);
    streams_.clear();
    for (hipEvent_t e : join_) (void)hipEventDestroy(e);
    join_.clear();
    if (fork_) (void)hipEventDestroy(fork_);
    fork_ = nullptr;
    if (capture_) (void)hipStreamDestroy(capture_);
    capture_ = nullptr;
    if (flags_) (void)hipFree(flags_);
    flags_ = nullptr;
    lanes_.clear();
}

void BatchGraph::launch(hipStream_t stream) { RA_HIP_CHECK(hipGraphLaunch(exec_, stream)); }

void BatchGraph::check() {
    RA_HIP_CHECK(hipDeviceSynchronize());
    int flags = 0;
    RA_HIP_CHECK(hipMemcpy(&flags, flags_, sizeof(int), hipMemcpyDeviceToHost));
    if (flags == 0) return;
    std::string what;
    if (flags & REPLAY_NEIGHBORS_CHANGED) what += " the neighbour lists changed size;";
    if (flags & REPLAY_BLOCKS_CHANGED) what += " the blocks changed size;";
    if (flags & REPLAY_NEIGHBOR_ERROR) what += " the neighbour list reported an error;";
    if (flags & REPLAY_KERNEL_ERROR) what += " a kernel reported an error;";
    throw Error::internal("a batch graph replay is inconsistent with its record:" + what);
}

void BatchGraph::output(size_t batch, featomic_amd_device_output_t* out) const {
    if (batch >= records_.size()) {
        throw Error::invalid_parameter("batch " + std::to_string(batch) + " is out of range for a graph of " +
                                       std::to_string(records_.size()) + " batches");
    }
    records_[batch]->describe(out);
}

}  // namespace rascaline_amd
