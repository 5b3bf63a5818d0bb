// Calculator objects behind the featomic C API: hyper-parameters, radial
// tables, device workspace and the host orchestration of the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/featomic.h"
#include "../../include/featomic_amd.h"
#include "device.hpp"
#include "host_transfer.hpp"
#include "hypers.hpp"
#include "radial.hpp"
#include "surrogate.hpp"
#include "tensormap.hpp"

namespace rascaline_amd {

// Grow-only device allocations reused across compute calls.
class Workspace {
public:
    ~Workspace();
    void* get(const std::string& name, size_t bytes);
    template <typename T>
    T* get(const std::string& name, size_t count) {
        return static_cast<T*>(get(name, count * sizeof(T) + 16));
    }
    void release();
    // a frozen workspace hands out its existing buffers only: a request that would
    // allocate or grow one throws (buffers captured in a graph must not move)
    bool frozen = false;

private:
    struct Slot {
        void* ptr = nullptr;
        size_t bytes = 0;
    };
    std::map<std::string, Slot> slots_;
};

// How a device pass runs. Eager: sizes read back from the device as the pass needs
// them (host synchronisations). Record: the same, keeping the sizes and the host-
// derived inputs of the batch in a BatchRecord. Replay: no synchronisation, upload or
// allocation: sizes and inputs come from the record (the form captured in a graph).
enum class PassMode { Eager, Record, Replay };

// Sizes and host-derived inputs of one device batch, recorded by a synchronising
// pass, and the description of its device output.
struct BatchRecord {
    long long n_entries = -1;                 // neighbour-list entries
    std::vector<long long> block_totals;      // (samples, rows) offsets per block
    Workspace inputs;                         // device copies of the batch's uploads
    std::map<std::string, size_t> input_counts;
    // host description of the output (featomic_amd_device_output_t arrays)
    std::vector<int32_t> keys;
    size_t key_size = 0;
    std::vector<int64_t> n_samples, n_rows, value_offsets, gradient_offsets, block_properties;
    std::vector<int64_t> sample_label_offsets, gradient_label_offsets, cell_gradient_offsets;
    featomic_amd_device_output_t output{};
    void describe(featomic_amd_device_output_t* out) const;
};

// Replay flag bits (featomic_amd_batch_graph_check)
constexpr int REPLAY_NEIGHBORS_CHANGED = 1;
constexpr int REPLAY_BLOCKS_CHANGED = 2;
constexpr int REPLAY_NEIGHBOR_ERROR = 4;
constexpr int REPLAY_KERNEL_ERROR = 8;

// The pass mode of a neighbour-list build (Eager when record is null).
struct PassContext {
    PassMode mode = PassMode::Eager;
    BatchRecord* record = nullptr;
    int* replay_flags = nullptr;  // device word the Replay checks OR into
};

// Systems gathered from the caller, concatenated.
struct HostBatch {
    std::vector<int64_t> offsets;        // n_systems + 1
    std::vector<double> positions;       // host copy (may stay empty when positions are on device)
    const double* device_positions = nullptr;
    std::vector<int32_t> types;          // atomic types
    std::vector<double> cells;           // n_systems x 9
    // user-provided half neighbour lists (use_native_system = false)
    bool has_pairs = false;
    std::vector<std::vector<featomic_pair_t>> pairs;

    size_t n_systems() const { return offsets.empty() ? 0 : offsets.size() - 1; }
    int64_t n_atoms() const { return offsets.empty() ? 0 : offsets.back(); }
};

enum class Kind { SphericalExpansion, PowerSpectrum, LodeSphericalExpansion, RadialSpectrum, NeighborList, ByPair };

struct DeviceNeighbors;
template <typename T>
struct SoapLaunch;

// Launch shapes of the SOAP kernels that the autotuner chooses (0 / -1: the
// built-in choice).
struct SoapTuning {
    // neighbour entries per LDS chunk of the expansion kernel (power of two <= 64)
    int expansion_entries = 0;
    // gradient rows: -1 the first window shape of the built-in table that fits the
    // default LDS budget (else the whole LDS); k >= 0 window shape k of
    // soap_window_shapes() within the whole LDS; -2 the untiled kernel
    // (register-blocked for the power spectrum, one wave per center for the
    // spherical expansion)
    int gradient_window = -1;
};

// (runs, entries, rows per pass, rank-2) window shapes of the gradient kernels that
// the autotuner tries; rank-2 shapes take the rank-2 MFMA kernel where it applies
// (FP32 power spectrum, n = 8, plain form) and the tiled kernel elsewhere
const std::vector<std::array<int, 4>>& soap_window_shapes();

struct Gradients {
    bool positions = false, strain = false, cell = false;
};

class Calculator {
public:
    Calculator(const std::string& registry_name, const std::string& parameters);
    ~Calculator();

    const char* name() const;
    const std::string& parameters() const { return parameters_; }
    const std::vector<double>& cutoffs() const { return cutoffs_; }
    Kind kind() const { return kind_; }
    double nl_cutoff() const { return kind_ == Kind::NeighborList ? nl_.cutoff : soap_.cutoff; }

    int precision = 64;
    // fast form: mixed-precision (bfloat16 hi + lo on MFMA, FP32 accumulation) contractions
    int fast_form = 0;

    // host systems in, host TensorMap out
    mts_tensormap_t* compute(const HostBatch& batch, const Gradients& gradients);

    // device-resident batch (positions on the device), outputs stay on device
    void compute_device(const HostBatch& batch, const Gradients& gradients, hipStream_t stream,
                        featomic_amd_device_output_t* output);
    void last_timings(double* total_ms, double* main_ms, int64_t* main_bytes);

    // vector-Jacobian product of the position gradients on a device-resident batch
    // (SOAP and LODE calculators):
    // out[atom][3] = Σ_blocks Σ_rows (dX/dr)[row] · grad_values[sample of row], with
    // grad_values in the device layout of this calculator's values for the batch
    // (the precision of the calculator); computed without the gradient rows
    void positions_vjp(const HostBatch& batch, hipStream_t stream, const void* grad_values, void* out);

    SoapTuning tuning;

    // record / replay of device batches (batch graphs): set the mode of the following
    // compute_device calls; replays use the record's sizes and inputs and OR
    // inconsistencies into replay_flags. A calculator bound to a graph refuses eager
    // device calls, which could move the buffers the graph uses.
    void set_pass(PassMode mode, BatchRecord* record, int* replay_flags);
    bool bound_to_graph = false;
    Workspace& workspace() { return ws_; }

    // times the candidate launch shapes on a device-resident batch (the best of
    // `repeats` runs each, device time), keeps the fastest in `tuning` and
    // returns a JSON report of every candidate
    std::string autotune(const HostBatch& batch, const Gradients& gradients, hipStream_t stream, int repeats);

    // learned radial integral (SOAP calculators): trains a network on the exact
    // GTO integral, judges it against this calculator's radial integral and
    // uses it for the following computations when it is accepted
    SurrogateReport train_radial_surrogate(const SurrogateOptions& options);
    // switches between the accepted network and the spline / exact integral
    void use_radial_surrogate(bool enabled);
    bool radial_surrogate_active() const { return use_surrogate_ && surrogate_ && surrogate_->accepted(); }
    const RadialSurrogate* radial_surrogate() const { return surrogate_.get(); }

private:
    std::string registry_name_;
    std::string parameters_;
    std::vector<double> cutoffs_;
    Kind kind_;
    SoapParams soap_;
    LodeParams lode_;
    NeighborListParams nl_;
    RadialTables tables_;
    std::vector<double> self_contribution_;  // SOAP l=0 self term / LODE centre contribution
    Workspace ws_;
    PinnedStaging staging_;  // pinned buffers of the host TensorMap transfers

    // host mirror of the last device-resident output
    struct DeviceResult {
        std::vector<int32_t> keys;
        size_t key_size = 0;
        std::vector<int64_t> n_samples, n_rows, value_offsets, gradient_offsets, block_properties;
        std::vector<int64_t> sample_label_offsets, gradient_label_offsets, cell_gradient_offsets;
        const void* values = nullptr;
        const void* gradients = nullptr;
        const int32_t* sample_labels = nullptr;
        const int32_t* gradient_labels = nullptr;
    } device_result_;

    PassMode mode_ = PassMode::Eager;
    BatchRecord* record_ = nullptr;
    int* replay_flags_ = nullptr;
    // device buffer `name` holding `host`: uploaded into the workspace (Eager), into the
    // batch record (Record), or the record's copy (Replay, no upload)
    template <typename U>
    U* batch_input(const std::string& name, const std::vector<U>& host, hipStream_t stream);
    // timing events are left out of replays (a captured graph holds no event records)
    void mark(hipEvent_t event, hipStream_t stream);

    std::unique_ptr<RadialSurrogate> surrogate_;
    bool use_surrogate_ = false;
    const void* vjp_grad_ = nullptr;
    void* vjp_out_ = nullptr;

    hipEvent_t ev_start_ = nullptr, ev_stop_ = nullptr, ev_main_start_ = nullptr, ev_main_stop_ = nullptr;
    // SOAP phases the autotuner times: the expansion kernel, and the gradient rows
    // from their planning pass to the end of ev_main_stop_
    hipEvent_t ev_exp_start_ = nullptr, ev_exp_stop_ = nullptr, ev_grad_start_ = nullptr;
    int64_t main_bytes_ = 0;

    template <typename T>
    mts_tensormap_t* run_soap(const HostBatch& batch, const Gradients& gradients, hipStream_t stream, bool to_host,
                              featomic_amd_device_output_t* output);
    template <typename T>
    void soap_launch(SoapLaunch<T>& lp, const DeviceNeighbors& nl, const int* d_species, int n_atoms, int S,
                     hipStream_t stream);
    // spherical_expansion_by_pair: one sample per directed pair, host TensorMap
    template <typename T>
    mts_tensormap_t* run_by_pair(const HostBatch& batch, const Gradients& gradients);
    // neighbor_list calculator: pair vectors from the GPU half list, host TensorMap
    mts_tensormap_t* run_neighbor_list(const HostBatch& batch, const Gradients& gradients);
    template <typename T>
    mts_tensormap_t* run_lode(const HostBatch& batch, const Gradients& gradients, hipStream_t stream,
                              featomic_amd_device_output_t* output);
    // LODE: device time of the pair-row GEMMs summed over systems (negative when unused)
    double main_ms_sum_ = -1.0;

    template <typename T>
    struct DeviceTables;
    std::shared_ptr<void> tables_f32_, tables_f64_;
    template <typename T>
    const DeviceTables<T>& device_tables();

    // neighbour list on the device, shared by all calculators
    friend struct NeighborRunner;
};

// Full neighbour list in device memory: per center i, entries offsets[i]..offsets[i+1]
struct DeviceNeighbors {
    long long n_entries = 0;
    long long* offsets = nullptr;   // n_atoms + 1
    int* entry_j = nullptr;
    int* entry_shift = nullptr;     // 3 per entry
    double* entry_vec = nullptr;    // 3 per entry
    double* entry_dist = nullptr;
};

// LDS bytes per workgroup on the current device.
int device_lds_limit();

// Sends a message to the logging callback set with featomic_set_logging_callback.
void log_message(int level, const std::string& message);

// The GPU neighbour list of one system (host positions and cell) as a half list.
std::vector<featomic_pair_t> gpu_half_pairs(const double* positions, uintptr_t n_atoms, const double* cell, double cutoff);
// The same for every system of a host batch (one GPU build for the batch).
std::vector<std::vector<featomic_pair_t>> gpu_half_pairs(const HostBatch& batch, double cutoff, Workspace& ws);

// Uploads the batch geometry and builds the full neighbour list (GPU cell list,
// or the caller's half list when provided).
DeviceNeighbors build_neighbors(Workspace& ws, const HostBatch& batch, double cutoff, const double* d_positions,
                                const int* d_atom_system, hipStream_t stream, const PassContext& pass = PassContext());


// Device copies of the radial and spherical-harmonics tables (per precision).
template <typename T>
struct Calculator::DeviceTables {
    DeviceRadial<T> radial;
    T* sh_a = nullptr;
    T* sh_b = nullptr;
    T* sh_c = nullptr;
    T* self_c = nullptr;
    std::vector<void*> owned;

    ~DeviceTables() {
        for (void* p : owned) (void)hipFree(p);
    }

    template <typename U>
    U* put(const std::vector<U>& v) {
        void* p = nullptr;
        RA_HIP_CHECK(hipMalloc(&p, sizeof(U) * std::max<size_t>(v.size(), 1)));
        if (!v.empty()) RA_HIP_CHECK(hipMemcpy(p, v.data(), sizeof(U) * v.size(), hipMemcpyHostToDevice));
        owned.push_back(p);
        return static_cast<U*>(p);
    }
};

// ---- helpers shared by the SOAP and LODE host code

template <typename T>
inline void upload(T* dst, const std::vector<T>& src, hipStream_t stream) {
    if (!src.empty()) RA_HIP_CHECK(hipMemcpyAsync(dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice, stream));
}

template <typename T>
inline void download(std::vector<T>& dst, const T* src, size_t count, hipStream_t stream) {
    dst.resize(count);
    if (count) RA_HIP_CHECK(hipMemcpyAsync(dst.data(), src, sizeof(T) * count, hipMemcpyDeviceToHost, stream));
}

struct Species {
    std::vector<int32_t> types;  // sorted unique atomic types
    std::vector<int32_t> index;  // per atom
};

inline Species species_of(const HostBatch& batch) {
    Species s;
    s.types = batch.types;
    std::sort(s.types.begin(), s.types.end());
    s.types.erase(std::unique(s.types.begin(), s.types.end()), s.types.end());
    if (s.types.size() > 32) throw Error::invalid_parameter("at most 32 different atomic types are supported");
    s.index.resize(batch.types.size());
    for (size_t i = 0; i < batch.types.size(); i++) {
        s.index[i] = static_cast<int32_t>(std::lower_bound(s.types.begin(), s.types.end(), batch.types[i]) - s.types.begin());
    }
    return s;
}

inline std::unique_ptr<mts_block_t> make_block(LabelsData samples, std::vector<LabelsData> components, LabelsData properties,
                                        std::vector<uintptr_t> shape) {
    auto b = std::make_unique<mts_block_t>();
    b->samples = std::move(samples);
    b->components = std::move(components);
    b->properties = std::move(properties);
    b->array.shape = std::move(shape);
    size_t total = 1;
    for (auto s : b->array.shape) total *= s;
    b->array.data.allocate(total);
    return b;
}

inline LabelsData mu_labels(int l) {
    std::vector<int32_t> v;
    for (int m = -l; m <= l; m++) v.push_back(m);
    return LabelsData({"o3_mu"}, v);
}


}  // namespace rascaline_amd
