// featomic C API (featomic/include/featomic.h, implementation in
// featomic/src/c_api/*.rs) over the GPU calculators, plus the extension entry
// points declared in include/featomic_amd.h.
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/featomic.h"
#include "../../include/featomic_amd.h"
#include "batch_graph.hpp"
#include "calculator.hpp"
#include "device.hpp"
#include "kernels.hpp"
#include "kernels_mlp.hpp"
#include "selection.hpp"

struct featomic_calculator_t {
    std::unique_ptr<rascaline_amd::Calculator> impl;
    std::string name_storage;
};

struct featomic_amd_batch_graph_t {
    std::unique_ptr<rascaline_amd::BatchGraph> impl;
};

namespace {
featomic_logging_callback_t LOGGING_CALLBACK = nullptr;
}  // namespace

namespace rascaline_amd {
void log_message(int level, const std::string& message) {
    if (LOGGING_CALLBACK) LOGGING_CALLBACK(level, message.c_str());
}
}  // namespace rascaline_amd

namespace {

using rascaline_amd::Error;
using rascaline_amd::log_message;

thread_local std::string LAST_ERROR;

// time spans for featomic_profiling_*, c_api/profiling.rs:17-91
struct Profiler {
    std::mutex mutex;
    bool enabled = false;
    struct Span {
        uint64_t calls = 0;
        double total_ms = 0.0;
    };
    std::map<std::string, Span> spans;
} PROFILER;

struct ScopedSpan {
    std::string name;
    std::chrono::steady_clock::time_point start;
    explicit ScopedSpan(const char* n) : name(n), start(std::chrono::steady_clock::now()) {}
    ~ScopedSpan() {
        if (!PROFILER.enabled) return;
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - start).count();
        std::lock_guard<std::mutex> lock(PROFILER.mutex);
        auto& s = PROFILER.spans[name];
        s.calls++;
        s.total_ms += ms;
    }
};

// catch every exception at the ABI (c_api/status.rs:82-88)
template <typename F>
featomic_status_t guard(F&& f) {
    try {
        f();
        return FEATOMIC_SUCCESS;
    } catch (const Error& e) {
        LAST_ERROR = e.what();
        return e.status;
    } catch (const std::exception& e) {
        LAST_ERROR = std::string("internal featomic error: ") + e.what();
        return FEATOMIC_INTERNAL_ERROR;
    } catch (...) {
        LAST_ERROR = "internal featomic error: unknown exception";
        return FEATOMIC_INTERNAL_ERROR;
    }
}

template <typename P>
void check_pointer(P p, const char* name) {
    if (!p) throw Error::invalid_parameter(std::string("got invalid NULL pointer for ") + name);
}

// copy a string into a caller buffer (c_api/utils.rs:5-17)
void copy_str(const std::string& s, char* buffer, uintptr_t len) {
    if (s.size() + 1 > len) {
        throw Error::buffer_size("string buffer is not big enough: got " + std::to_string(len) + " bytes, need " +
                                 std::to_string(s.size() + 1));
    }
    std::memcpy(buffer, s.c_str(), s.size() + 1);
}

void system_call(featomic_status_t status, const char* what) {
    if (status != FEATOMIC_SUCCESS) {
        throw Error::system(status, std::string("call to featomic_system_t::") + what + " failed");
    }
}

rascaline_amd::HostBatch gather_systems(featomic_system_t* systems, uintptr_t count, bool use_native, double cutoff,
                                        bool need_neighbors) {
    rascaline_amd::HostBatch batch;
    batch.offsets.push_back(0);
    batch.has_pairs = need_neighbors && !use_native;
    for (uintptr_t s = 0; s < count; s++) {
        featomic_system_t& sys = systems[s];
        check_pointer(sys.size, "featomic_system_t.size");
        check_pointer(sys.types, "featomic_system_t.types");
        check_pointer(sys.positions, "featomic_system_t.positions");
        check_pointer(sys.cell, "featomic_system_t.cell");
        uintptr_t n = 0;
        system_call(sys.size(sys.user_data, &n), "size");
        const int32_t* types = nullptr;
        const double* positions = nullptr;
        double cell[9];
        if (n > 0) {
            system_call(sys.types(sys.user_data, &types), "types");
            system_call(sys.positions(sys.user_data, &positions), "positions");
        }
        system_call(sys.cell(sys.user_data, cell), "cell");
        batch.types.insert(batch.types.end(), types, types + n);
        batch.positions.insert(batch.positions.end(), positions, positions + 3 * n);
        batch.cells.insert(batch.cells.end(), cell, cell + 9);
        batch.offsets.push_back(batch.offsets.back() + static_cast<int64_t>(n));
        if (batch.has_pairs) {
            check_pointer(sys.compute_neighbors, "featomic_system_t.compute_neighbors");
            check_pointer(sys.pairs, "featomic_system_t.pairs");
            system_call(sys.compute_neighbors(sys.user_data, cutoff), "compute_neighbors");
            const featomic_pair_t* pairs = nullptr;
            uintptr_t n_pairs = 0;
            system_call(sys.pairs(sys.user_data, &pairs, &n_pairs), "pairs");
            batch.pairs.emplace_back(pairs, pairs + n_pairs);
        }
    }
    return batch;
}

rascaline_amd::Gradients parse_gradients(const char* const* gradients, uintptr_t count, rascaline_amd::Kind kind) {
    rascaline_amd::Gradients g;
    for (uintptr_t k = 0; k < count; k++) {
        check_pointer(gradients[k], "gradients");
        const std::string name = gradients[k];
        if (name == "positions") {
            g.positions = true;
        } else if (name == "strain") {
            g.strain = true;
        } else if (name == "cell") {
            g.cell = true;
        } else {
            throw Error::invalid_parameter("unexpected gradient \"" + name +
                                           "\", should be one of \"positions\", \"cell\", or \"strain\"");
        }
    }
    // calculator.rs:342-380: only the SOAP calculators have cell and strain gradients
    // (cell is checked first)
    if ((kind == rascaline_amd::Kind::LodeSphericalExpansion || kind == rascaline_amd::Kind::NeighborList) && (g.strain || g.cell)) {
        const std::string name = kind == rascaline_amd::Kind::NeighborList ? "neighbors list" : "lode spherical expansion";
        throw Error::invalid_parameter("the " + name + " calculator does not support gradients with respect to " +
                                       (g.cell ? "cell" : "strain"));
    }
    return g;
}

}  // namespace

extern "C" {

const char* featomic_last_error(void) { return LAST_ERROR.c_str(); }

featomic_status_t featomic_set_logging_callback(featomic_logging_callback_t callback) {
    return guard([&] { LOGGING_CALLBACK = callback; });
}

featomic_calculator_t* featomic_calculator(const char* name, const char* parameters) {
    featomic_calculator_t* result = nullptr;
    const featomic_status_t status = guard([&] {
        check_pointer(name, "name");
        check_pointer(parameters, "parameters");
        auto calc = std::make_unique<featomic_calculator_t>();
        calc->impl = std::make_unique<rascaline_amd::Calculator>(name, parameters);
        result = calc.release();
    });
    return status == FEATOMIC_SUCCESS ? result : nullptr;
}

featomic_status_t featomic_calculator_free(featomic_calculator_t* calculator) {
    return guard([&] { delete calculator; });
}

featomic_status_t featomic_calculator_name(const featomic_calculator_t* calculator, char* name, uintptr_t bufflen) {
    return guard([&] {
        check_pointer(calculator, "calculator");
        check_pointer(name, "name");
        copy_str(calculator->impl->name(), name, bufflen);
    });
}

featomic_status_t featomic_calculator_parameters(const featomic_calculator_t* calculator, char* parameters, uintptr_t bufflen) {
    return guard([&] {
        check_pointer(calculator, "calculator");
        check_pointer(parameters, "parameters");
        copy_str(calculator->impl->parameters(), parameters, bufflen);
    });
}

featomic_status_t featomic_calculator_cutoffs(const featomic_calculator_t* calculator, const double** cutoffs,
                                              uintptr_t* cutoffs_count) {
    return guard([&] {
        check_pointer(calculator, "calculator");
        check_pointer(cutoffs, "cutoffs");
        check_pointer(cutoffs_count, "cutoffs_count");
        *cutoffs = calculator->impl->cutoffs().data();
        *cutoffs_count = calculator->impl->cutoffs().size();
    });
}

featomic_status_t featomic_calculator_compute(featomic_calculator_t* calculator, mts_tensormap_t** descriptor,
                                              featomic_system_t* systems, uintptr_t systems_count,
                                              featomic_calculation_options_t options) {
    return guard([&] {
        ScopedSpan span("Calculator::compute");
        check_pointer(calculator, "calculator");
        check_pointer(descriptor, "descriptor");
        if (systems_count == 0) {
            // c_api/calculator.rs:400-403
            log_message(FEATOMIC_LOG_LEVEL_WARN, "0 systems given to featomic_calculator_compute, nothing to do");
            return;
        }
        check_pointer(systems, "systems");
        if (options.gradients_count > 0) check_pointer(options.gradients, "options.gradients");
        auto& impl = *calculator->impl;
        const auto grads = parse_gradients(options.gradients, options.gradients_count, impl.kind());
        if (grads.cell && !std::getenv("FEATOMIC_NO_WARN_CELL_GRADIENTS")) {
            // calculator.rs:350-358
            log_message(FEATOMIC_LOG_LEVEL_WARN,
                        "The meaning of \"cell\" gradients has changed recently, you likely want \"strain\" gradients "
                        "instead. Please review the documentation carefully.");
        }
        const bool soap = impl.kind() != rascaline_amd::Kind::LodeSphericalExpansion;
        if (impl.kind() == rascaline_amd::Kind::NeighborList && !(impl.nl_cutoff() > 0.0 && std::isfinite(impl.nl_cutoff()))) {
            // neighbor_list.rs:67-68 asserts this before anything else
            throw Error::internal("assertion failed: self.cutoff > 0.0 && self.cutoff.is_finite()");
        }
        auto batch = gather_systems(systems, systems_count, options.use_native_system, soap ? impl.nl_cutoff() : 0.0, soap);
        mts_tensormap_t* full = impl.compute(batch, grads);
        if (rascaline_amd::has_selections(options.selected_keys, options.selected_samples, options.selected_properties)) {
            // the complete result, then the selected keys, samples and properties of it
            *descriptor = rascaline_amd::apply_selections(full, options.selected_keys, options.selected_samples,
                                                          options.selected_properties);
        } else {
            *descriptor = full;
        }
    });
}

featomic_status_t featomic_profiling_clear(void) {
    return guard([&] {
        std::lock_guard<std::mutex> lock(PROFILER.mutex);
        PROFILER.spans.clear();
    });
}

featomic_status_t featomic_profiling_enable(bool enabled) {
    return guard([&] { PROFILER.enabled = enabled; });
}

featomic_status_t featomic_profiling_get(const char* format, char* buffer, uintptr_t bufflen) {
    return guard([&] {
        check_pointer(format, "format");
        check_pointer(buffer, "buffer");
        const std::string f = format;
        std::string out;
        std::lock_guard<std::mutex> lock(PROFILER.mutex);
        if (f == "json") {
            out = "{";
            bool first = true;
            for (auto& kv : PROFILER.spans) {
                if (!first) out += ",";
                first = false;
                out += "\"" + kv.first + "\":{\"calls\":" + std::to_string(kv.second.calls) +
                       ",\"total_ms\":" + std::to_string(kv.second.total_ms) + "}";
            }
            out += "}";
        } else if (f == "table" || f == "short_table") {
            out = "span | calls | total (ms)\n";
            for (auto& kv : PROFILER.spans) {
                out += kv.first + " | " + std::to_string(kv.second.calls) + " | " + std::to_string(kv.second.total_ms) + "\n";
            }
        } else {
            throw Error::invalid_parameter("invalid format for profiling data: expected 'table', 'short_table' or 'json', got '" + f + "'");
        }
        copy_str(out, buffer, bufflen);
    });
}

// ---------------------------------------------------------------------------- extensions

featomic_status_t featomic_amd_calculator_set_precision(featomic_calculator_t* calculator, int32_t bits) {
    return guard([&] {
        check_pointer(calculator, "calculator");
        if (bits != 32 && bits != 64) throw Error::invalid_parameter("precision must be 32 or 64 bits");
        calculator->impl->precision = bits;
    });
}

featomic_status_t featomic_amd_calculator_set_fast_form(featomic_calculator_t* calculator, int32_t enabled) {
    return guard([&] {
        check_pointer(calculator, "calculator");
        if (enabled != 0 && enabled != 1) throw Error::invalid_parameter("fast form must be 0 (plain) or 1 (fast)");
        calculator->impl->fast_form = enabled;
    });
}

featomic_status_t featomic_amd_surrogate_options_default(featomic_amd_surrogate_options_t* options) {
    return guard([&] {
        check_pointer(options, "options");
        const rascaline_amd::SurrogateOptions d;
        options->n_samples = d.n_samples;
        options->steps = d.steps;
        options->batch = d.batch;
        options->learning_rate = d.learning_rate;
        options->ridge = d.ridge;
        options->tolerance = d.tolerance;
        options->seed = d.seed;
        options->n_validation = d.n_validation;
    });
}

featomic_status_t featomic_amd_calculator_train_radial_surrogate(featomic_calculator_t* calculator,
                                                                 const featomic_amd_surrogate_options_t* options,
                                                                 featomic_amd_surrogate_report_t* report) {
    return guard([&] {
        check_pointer(calculator, "calculator");
        check_pointer(options, "options");
        rascaline_amd::SurrogateOptions o;
        o.n_samples = options->n_samples;
        o.steps = options->steps;
        o.batch = options->batch;
        o.learning_rate = options->learning_rate;
        o.ridge = options->ridge;
        o.tolerance = options->tolerance;
        o.seed = options->seed;
        o.n_validation = options->n_validation;
        if (report && (report->value_errors || report->derivative_errors) && report->errors_size < 0) {
            throw Error::invalid_parameter("errors_size must not be negative");
        }
        const rascaline_amd::SurrogateReport r = calculator->impl->train_radial_surrogate(o);
        if (report) {
            report->accepted = r.accepted ? 1 : 0;
            report->adam_loss = r.adam_loss;
            report->max_value_error = r.max_value_error;
            report->max_derivative_error = r.max_derivative_error;
            report->train_ms = r.train_ms;
            const int64_t count = static_cast<int64_t>(r.value_errors.size());
            if (report->value_errors || report->derivative_errors) {
                if (report->errors_size < count) {
                    throw Error::buffer_size("error buffers hold " + std::to_string(report->errors_size) +
                                             " entries, " + std::to_string(count) + " are needed");
                }
                for (int64_t k = 0; k < count; k++) {
                    if (report->value_errors) report->value_errors[k] = r.value_errors[k];
                    if (report->derivative_errors) report->derivative_errors[k] = r.derivative_errors[k];
                }
            }
        }
    });
}

featomic_status_t featomic_amd_calculator_use_radial_surrogate(featomic_calculator_t* calculator, int32_t enabled) {
    return guard([&] {
        check_pointer(calculator, "calculator");
        if (enabled != 0 && enabled != 1) throw Error::invalid_parameter("enabled must be 0 or 1");
        calculator->impl->use_radial_surrogate(enabled == 1);
    });
}

featomic_status_t featomic_amd_calculator_radial_surrogate_parameters(featomic_calculator_t* calculator, float* params,
                                                                      int64_t count, int64_t* required) {
    return guard([&] {
        check_pointer(calculator, "calculator");
        check_pointer(required, "required");
        const rascaline_amd::RadialSurrogate* s = calculator->impl->radial_surrogate();
        if (!s) throw Error::invalid_parameter("this calculator has no trained radial surrogate");
        const std::vector<float>& p = s->parameters();
        const std::vector<float>& sc = s->output_scale();
        *required = static_cast<int64_t>(p.size() + sc.size());
        if (!params) return;
        if (count < *required) {
            throw Error::buffer_size("parameter buffer holds " + std::to_string(count) + " values, " +
                                     std::to_string(*required) + " are needed");
        }
        std::copy(p.begin(), p.end(), params);
        std::copy(sc.begin(), sc.end(), params + p.size());
    });
}

featomic_status_t featomic_amd_mlp_loss_gradient(int32_t n_out, int64_t n_samples, const float* params,
                                                 const float* inputs, const float* targets, double* loss,
                                                 float* gradient) {
    return guard([&] {
        check_pointer(params, "params");
        check_pointer(inputs, "inputs");
        check_pointer(targets, "targets");
        check_pointer(loss, "loss");
        check_pointer(gradient, "gradient");
        *loss = rascaline_amd::mlp_loss_gradient_device(n_out, n_samples, params, inputs, targets, gradient);
    });
}

featomic_status_t featomic_amd_mlp_loss_gradient_noalloc(int32_t n_out, int64_t n_samples, const float* params,
                                                         const float* inputs, const float* targets, double* loss,
                                                         float* gradient, void* stream) {
    return guard([&] {
        check_pointer(params, "params");
        check_pointer(inputs, "inputs");
        check_pointer(targets, "targets");
        check_pointer(loss, "loss");
        check_pointer(gradient, "gradient");
        rascaline_amd::mlp_loss_gradient_noalloc(n_out, n_samples, params, inputs, targets, loss, gradient,
                                                 static_cast<hipStream_t>(stream));
    });
}

featomic_status_t featomic_amd_mlp_forward(int32_t n_out, int64_t n_samples, const float* params, const float* inputs,
                                           float* outputs) {
    return guard([&] {
        check_pointer(params, "params");
        check_pointer(inputs, "inputs");
        check_pointer(outputs, "outputs");
        rascaline_amd::mlp_forward_device(n_out, n_samples, params, inputs, outputs);
    });
}

featomic_status_t featomic_amd_neighbors(const double* positions, uintptr_t n_atoms, const double* cell, double cutoff,
                                         featomic_pair_t** pairs, uintptr_t* n_pairs) {
    return guard([&] {
        check_pointer(cell, "cell");
        check_pointer(pairs, "pairs");
        check_pointer(n_pairs, "n_pairs");
        if (n_atoms > 0) check_pointer(positions, "positions");
        if (!(cutoff > 0.0)) throw Error::invalid_parameter("cutoff must be positive");
        std::vector<featomic_pair_t> out = rascaline_amd::gpu_half_pairs(positions, n_atoms, cell, cutoff);
        *n_pairs = out.size();
        *pairs = static_cast<featomic_pair_t*>(std::malloc(sizeof(featomic_pair_t) * std::max<size_t>(out.size(), 1)));
        if (!out.empty()) std::memcpy(*pairs, out.data(), sizeof(featomic_pair_t) * out.size());
    });
}

void featomic_amd_free(void* pointer) { std::free(pointer); }

// host description of a device-resident batch
static rascaline_amd::HostBatch device_batch(const featomic_amd_device_systems_t* systems, const char* what) {
    check_pointer(systems, what);
    check_pointer(systems->atom_offsets, "systems.atom_offsets");
    check_pointer(systems->cells, "systems.cells");
    rascaline_amd::HostBatch batch;
    batch.offsets.assign(systems->atom_offsets, systems->atom_offsets + systems->n_systems + 1);
    const int64_t n_atoms = batch.n_atoms();
    if (n_atoms > 0) {
        check_pointer(systems->positions, "systems.positions");
        check_pointer(systems->types, "systems.types");
    }
    batch.device_positions = systems->positions;
    batch.types.assign(systems->types, systems->types + n_atoms);
    batch.cells.assign(systems->cells, systems->cells + 9 * systems->n_systems);
    return batch;
}

featomic_status_t featomic_amd_compute_device(featomic_calculator_t* calculator, const featomic_amd_device_systems_t* systems,
                                              const char* const* gradients, uintptr_t gradients_count, void* stream,
                                              featomic_amd_device_output_t* output) {
    return guard([&] {
        check_pointer(calculator, "calculator");
        check_pointer(systems, "systems");
        check_pointer(output, "output");
        const rascaline_amd::HostBatch batch = device_batch(systems, "systems");
        auto& impl = *calculator->impl;
        const auto grads = parse_gradients(gradients, gradients_count, impl.kind());
        if (grads.strain) throw Error::invalid_parameter("featomic_amd_compute_device supports positions and cell gradients");
        impl.compute_device(batch, grads, static_cast<hipStream_t>(stream), output);
    });
}

featomic_status_t featomic_amd_batch_graph_create(featomic_calculator_t* const* calculators, uintptr_t n_calculators,
                                                  const featomic_amd_device_systems_t* batches, uintptr_t n_batches,
                                                  const char* const* gradients, uintptr_t gradients_count, void* stream,
                                                  featomic_amd_batch_graph_t** graph) {
    return guard([&] {
        check_pointer(calculators, "calculators");
        check_pointer(graph, "graph");
        if (n_calculators == 0) throw Error::invalid_parameter("a batch graph needs at least one calculator");
        if (n_batches > 0) check_pointer(batches, "batches");
        *graph = nullptr;
        std::vector<rascaline_amd::Calculator*> lanes;
        for (uintptr_t l = 0; l < n_calculators; l++) {
            check_pointer(calculators[l], "calculators[i]");
            lanes.push_back(calculators[l]->impl.get());
        }
        const auto grads = parse_gradients(gradients, gradients_count, lanes[0]->kind());
        std::vector<rascaline_amd::HostBatch> host;
        for (uintptr_t k = 0; k < n_batches; k++) host.push_back(device_batch(&batches[k], "batches[i]"));
        auto g = std::make_unique<featomic_amd_batch_graph_t>();
        g->impl = std::make_unique<rascaline_amd::BatchGraph>(lanes, host, grads, static_cast<hipStream_t>(stream));
        *graph = g.release();
    });
}

featomic_status_t featomic_amd_batch_graph_launch(featomic_amd_batch_graph_t* graph, void* stream) {
    return guard([&] {
        check_pointer(graph, "graph");
        graph->impl->launch(static_cast<hipStream_t>(stream));
    });
}

featomic_status_t featomic_amd_batch_graph_check(featomic_amd_batch_graph_t* graph) {
    return guard([&] {
        check_pointer(graph, "graph");
        graph->impl->check();
    });
}

featomic_status_t featomic_amd_batch_graph_output(const featomic_amd_batch_graph_t* graph, uintptr_t batch,
                                                  featomic_amd_device_output_t* output) {
    return guard([&] {
        check_pointer(graph, "graph");
        check_pointer(output, "output");
        graph->impl->output(batch, output);
    });
}

featomic_status_t featomic_amd_batch_graph_free(featomic_amd_batch_graph_t* graph) {
    return guard([&] { delete graph; });
}

featomic_status_t featomic_amd_positions_vjp(featomic_calculator_t* calculator, const featomic_amd_device_systems_t* systems,
                                             const void* grad_values, void* stream, void* positions_grad) {
    return guard([&] {
        check_pointer(calculator, "calculator");
        const rascaline_amd::HostBatch batch = device_batch(systems, "systems");
        check_pointer(grad_values, "grad_values");
        check_pointer(positions_grad, "positions_grad");
        calculator->impl->positions_vjp(batch, static_cast<hipStream_t>(stream), grad_values, positions_grad);
    });
}

featomic_status_t featomic_amd_calculator_autotune(featomic_calculator_t* calculator,
                                                  const featomic_amd_device_systems_t* systems,
                                                  const char* const* gradients, uintptr_t gradients_count, void* stream,
                                                  int32_t repeats, char* report, uintptr_t report_size) {
    return guard([&] {
        check_pointer(calculator, "calculator");
        check_pointer(systems, "systems");
        check_pointer(systems->atom_offsets, "systems.atom_offsets");
        check_pointer(systems->cells, "systems.cells");
        auto& impl = *calculator->impl;
        const auto grads = parse_gradients(gradients, gradients_count, impl.kind());
        if (grads.strain || grads.cell) throw Error::invalid_parameter("the autotuner supports positions gradients only");
        rascaline_amd::HostBatch batch;
        batch.offsets.assign(systems->atom_offsets, systems->atom_offsets + systems->n_systems + 1);
        const int64_t n_atoms = batch.n_atoms();
        if (n_atoms > 0) {
            check_pointer(systems->positions, "systems.positions");
            check_pointer(systems->types, "systems.types");
        }
        batch.device_positions = systems->positions;
        batch.types.assign(systems->types, systems->types + n_atoms);
        batch.cells.assign(systems->cells, systems->cells + 9 * systems->n_systems);
        const std::string text = impl.autotune(batch, grads, static_cast<hipStream_t>(stream), repeats);
        if (report) copy_str(text, report, report_size);
    });
}

featomic_status_t featomic_amd_calculator_set_tuning(featomic_calculator_t* calculator, int32_t expansion_entries,
                                                     int32_t gradient_window) {
    return guard([&] {
        check_pointer(calculator, "calculator");
        const bool tiled_count =
            expansion_entries >= 0 && expansion_entries <= 64 && (expansion_entries & (expansion_entries - 1)) == 0;
        if (!tiled_count && expansion_entries != -8 && expansion_entries != -16) {
            throw Error::invalid_parameter(
                "expansion_entries must be 0 (built-in), a power of two up to 64, or -8 / -16 (matrix-core kernel)");
        }
        if (gradient_window < -2 || gradient_window >= static_cast<int32_t>(rascaline_amd::soap_window_shapes().size())) {
            throw Error::invalid_parameter("gradient_window must be -2, -1 or a window shape index below " +
                                           std::to_string(rascaline_amd::soap_window_shapes().size()));
        }
        calculator->impl->tuning.expansion_entries = expansion_entries;
        calculator->impl->tuning.gradient_window = gradient_window;
    });
}

featomic_status_t featomic_amd_calculator_get_tuning(featomic_calculator_t* calculator, int32_t* expansion_entries,
                                                     int32_t* gradient_window) {
    return guard([&] {
        check_pointer(calculator, "calculator");
        if (expansion_entries) *expansion_entries = calculator->impl->tuning.expansion_entries;
        if (gradient_window) *gradient_window = calculator->impl->tuning.gradient_window;
    });
}

featomic_status_t featomic_amd_last_timings(featomic_calculator_t* calculator, double* total_ms, double* main_kernel_ms,
                                            int64_t* main_kernel_bytes) {
    return guard([&] {
        check_pointer(calculator, "calculator");
        calculator->impl->last_timings(total_ms, main_kernel_ms, main_kernel_bytes);
    });
}

}  // extern "C"
