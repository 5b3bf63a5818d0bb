// Calculator construction and the SOAP device pipeline.
#include "calculator.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>

#include "device.hpp"
#include "kernels.hpp"
#include "selection.hpp"
#include "special.hpp"

namespace rascaline_amd {

static void check_unbound(const Calculator& c);

// ---------------------------------------------------------------------------- workspace

Workspace::~Workspace() { release(); }

void* Workspace::get(const std::string& name, size_t bytes) {
    Slot& slot = slots_[name];
    if (bytes == 0) bytes = 16;
    if (slot.bytes < bytes) {
        if (frozen) throw Error::internal("the buffer '" + name + "' of a frozen workspace would have to grow");
        // kernels queued by an earlier call may still use the old buffer
        if (slot.ptr) {
            RA_HIP_CHECK(hipDeviceSynchronize());
            RA_HIP_CHECK(hipFree(slot.ptr));
        }
        slot.ptr = nullptr;
        // grow with some slack so that repeated calls with similar sizes reuse the buffer
        const size_t target = bytes + bytes / 8;
        RA_HIP_CHECK(hipMalloc(&slot.ptr, target));
        slot.bytes = target;
    }
    return slot.ptr;
}

void Workspace::release() {
    for (auto& kv : slots_) {
        if (kv.second.ptr) (void)hipFree(kv.second.ptr);
    }
    slots_.clear();
}

int device_lds_limit() {
    static int limit = -1;
    if (limit < 0) {
        int dev = 0;
        RA_HIP_CHECK(hipGetDevice(&dev));
        hipDeviceProp_t prop;
        RA_HIP_CHECK(hipGetDeviceProperties(&prop, dev));
        limit = static_cast<int>(prop.sharedMemPerBlock);
        if (limit <= 0) limit = 65536;
    }
    return limit;
}

// GPU neighbour lists of the systems of a host batch as half lists of featomic_pair_t:
// j > i, or j == i with the shift in the positive half-space
// (featomic/src/systems/neighbors.rs:253-286). Pairs come sorted by (first, second,
// shift a, b, c); the reference sorts by (first, second) only
// (neighbors.rs:391-396), leaving pairs of one atom pair with several shifts in an
// unspecified order.
std::vector<std::vector<featomic_pair_t>> gpu_half_pairs(const HostBatch& batch, double cutoff, Workspace& ws) {
    const int64_t n_atoms = batch.n_atoms();
    const size_t n_systems = batch.n_systems();
    std::vector<std::vector<featomic_pair_t>> result(n_systems);
    if (n_atoms == 0) return result;
    hipStream_t stream = nullptr;
    double* d_pos = ws.get<double>("positions", 3 * n_atoms);
    int* d_sys = ws.get<int>("atom_system", n_atoms);
    std::vector<int> atom_system(n_atoms);
    for (size_t s = 0; s < n_systems; s++) {
        for (int64_t a = batch.offsets[s]; a < batch.offsets[s + 1]; a++) atom_system[a] = static_cast<int>(s);
    }
    RA_HIP_CHECK(hipMemcpy(d_pos, batch.positions.data(), sizeof(double) * 3 * n_atoms, hipMemcpyHostToDevice));
    RA_HIP_CHECK(hipMemcpy(d_sys, atom_system.data(), sizeof(int) * n_atoms, hipMemcpyHostToDevice));
    auto nl = build_neighbors(ws, batch, cutoff, d_pos, d_sys, stream);
    std::vector<long long> offsets(n_atoms + 1);
    std::vector<int> j(nl.n_entries), sh(3 * nl.n_entries);
    std::vector<double> v(3 * nl.n_entries), d(nl.n_entries);
    RA_HIP_CHECK(hipMemcpy(offsets.data(), nl.offsets, sizeof(long long) * (n_atoms + 1), hipMemcpyDeviceToHost));
    if (nl.n_entries) {
        RA_HIP_CHECK(hipMemcpy(j.data(), nl.entry_j, sizeof(int) * nl.n_entries, hipMemcpyDeviceToHost));
        RA_HIP_CHECK(hipMemcpy(sh.data(), nl.entry_shift, sizeof(int) * 3 * nl.n_entries, hipMemcpyDeviceToHost));
        RA_HIP_CHECK(hipMemcpy(v.data(), nl.entry_vec, sizeof(double) * 3 * nl.n_entries, hipMemcpyDeviceToHost));
        RA_HIP_CHECK(hipMemcpy(d.data(), nl.entry_dist, sizeof(double) * nl.n_entries, hipMemcpyDeviceToHost));
    }
    for (size_t s = 0; s < n_systems; s++) {
        std::vector<featomic_pair_t>& out = result[s];
        const int64_t first_atom = batch.offsets[s];
        for (int64_t gi = first_atom; gi < batch.offsets[s + 1]; gi++) {
            // entry_j holds batch-global atom indices
            const int i = static_cast<int>(gi - first_atom);
            for (long long e = offsets[gi]; e < offsets[gi + 1]; e++) {
                const int jj = static_cast<int>(j[e] - first_atom);
                const int* t = &sh[3 * e];
                bool keep = jj > i;
                if (jj == i) {
                    const int sum = t[0] + t[1] + t[2];
                    keep = !(sum < 0 || (sum == 0 && (t[2] < 0 || (t[2] == 0 && t[1] < 0))));
                }
                if (!keep) continue;
                featomic_pair_t p;
                p.first = static_cast<uintptr_t>(i);
                p.second = static_cast<uintptr_t>(jj);
                p.distance = d[e];
                p.vector[0] = v[3 * e];
                p.vector[1] = v[3 * e + 1];
                p.vector[2] = v[3 * e + 2];
                p.cell_shift_indices[0] = t[0];
                p.cell_shift_indices[1] = t[1];
                p.cell_shift_indices[2] = t[2];
                out.push_back(p);
                if (p.distance * p.distance < 1e-3) {
                    log_message(FEATOMIC_LOG_LEVEL_WARN, "atoms " + std::to_string(p.first) + " and " +
                                                             std::to_string(p.second) + " are very close to one another (" +
                                                             std::to_string(p.distance) + " A)");
                }
            }
        }
    }
    return result;
}

std::vector<featomic_pair_t> gpu_half_pairs(const double* positions, uintptr_t n_atoms, const double* cell, double cutoff) {
    HostBatch batch;
    batch.offsets = {0, static_cast<int64_t>(n_atoms)};
    if (n_atoms) batch.positions.assign(positions, positions + 3 * n_atoms);
    batch.types.assign(n_atoms, 0);
    batch.cells.assign(cell, cell + 9);
    Workspace ws;
    return std::move(gpu_half_pairs(batch, cutoff, ws)[0]);
}

// ---------------------------------------------------------------------------- construction

Calculator::Calculator(const std::string& registry_name, const std::string& parameters)
    : registry_name_(registry_name), parameters_(parameters) {
    if (registry_name == "spherical_expansion") {
        kind_ = Kind::SphericalExpansion;
        soap_ = parse_soap_params(parameters, true);
    } else if (registry_name == "soap_power_spectrum") {
        kind_ = Kind::PowerSpectrum;
        soap_ = parse_soap_params(parameters, false);
    } else if (registry_name == "lode_spherical_expansion") {
        kind_ = Kind::LodeSphericalExpansion;
        lode_ = parse_lode_params(parameters);
    } else if (registry_name == "soap_radial_spectrum") {
        // the l = 0 spherical expansion, relabelled (soap/radial_spectrum.rs:75-135)
        kind_ = Kind::RadialSpectrum;
        soap_ = parse_soap_params(parameters, true, true);
    } else if (registry_name == "spherical_expansion_by_pair") {
        // the expansion's parameters, one sample per pair (soap/spherical_expansion_pair.rs:547-840)
        kind_ = Kind::ByPair;
        soap_ = parse_soap_params(parameters, true);
    } else if (registry_name == "neighbor_list") {
        // pair vectors of the neighbour list (neighbor_list.rs:31-44)
        kind_ = Kind::NeighborList;
        nl_ = parse_neighbor_list_params(parameters);
        cutoffs_.push_back(nl_.cutoff);
        return;
    } else if (registry_name == "atomic_composition" || registry_name == "dummy_calculator" ||
               registry_name == "sorted_distances") {
        throw Error::invalid_parameter("the '" + registry_name + "' calculator is not provided by this GPU implementation");
    } else {
        throw Error::invalid_parameter("unknown calculator with name '" + registry_name + "'");
    }

    if (kind_ == Kind::LodeSphericalExpansion) {
        tables_ = build_lode_radial(lode_);
        self_contribution_ = lode_center_contribution(lode_, tables_);
    } else {
        cutoffs_.push_back(soap_.cutoff);
        tables_ = build_soap_radial(soap_);
        // self contribution, soap/spherical_expansion_pair.rs:222-252:
        // w * f(0) * Y_00 * I_n0(0)
        std::vector<double> ri(soap_.n_radial), rg(soap_.n_radial);
        if (tables_.use_spline) {
            evaluate_channel(tables_.channels[0], soap_.n_radial, 0.0, ri.data(), nullptr);
        } else {
            tables_.exact(0, 0.0, ri.data(), rg.data());
        }
        double f0 = 1.0;
        if (soap_.has_scaling) {
            f0 = soap_.scaling_rate / (soap_.scaling_rate + std::pow(0.0 / soap_.scaling_scale, soap_.scaling_exponent));
        }
        // smoothing(0) is 1 for both cutoff functions (cutoff > width)
        double fc = 1.0;
        if (soap_.smoothing == Smoothing::ShiftedCosine && 0.0 > soap_.cutoff - soap_.smoothing_width) {
            fc = 0.5 * (1.0 + std::cos(RA_PI * (0.0 - soap_.cutoff + soap_.smoothing_width) / soap_.smoothing_width));
        }
        const double y00 = 0.3989422804014327 / 1.4142135623730951;
        self_contribution_.resize(soap_.n_radial);
        for (int k = 0; k < soap_.n_radial; k++) {
            self_contribution_[k] = soap_.center_atom_weight * (f0 * fc) * y00 * ri[k];
        }
    }
}

Calculator::~Calculator() {
    if (ev_start_) (void)hipEventDestroy(ev_start_);
    if (ev_stop_) (void)hipEventDestroy(ev_stop_);
    if (ev_main_start_) (void)hipEventDestroy(ev_main_start_);
    if (ev_main_stop_) (void)hipEventDestroy(ev_main_stop_);
    if (ev_exp_start_) (void)hipEventDestroy(ev_exp_start_);
    if (ev_exp_stop_) (void)hipEventDestroy(ev_exp_stop_);
    if (ev_grad_start_) (void)hipEventDestroy(ev_grad_start_);
}

const char* Calculator::name() const {
    switch (kind_) {
    case Kind::SphericalExpansion: return "spherical expansion";
    case Kind::PowerSpectrum: return "SOAP power spectrum";
    case Kind::LodeSphericalExpansion: return "lode spherical expansion";
    case Kind::RadialSpectrum: return "radial spectrum";
    case Kind::NeighborList: return "neighbors list";
    case Kind::ByPair: return "spherical expansion by pair";
    }
    return "";
}

// ---------------------------------------------------------------------------- device tables


template <typename T>
const Calculator::DeviceTables<T>& Calculator::device_tables() {
    std::shared_ptr<void>& slot = sizeof(T) == 4 ? tables_f32_ : tables_f64_;
    if (!slot) {
        auto t = std::make_shared<DeviceTables<T>>();
        const RadialTables& rt = tables_;
        const int L = rt.max_angular + 1, n = rt.n;
        std::vector<int> koff(L, 0), nk(L, 0);
        std::vector<T> kpos, kval, kder;
        std::vector<double> kgrid(2 * L, 0.0);
        if (rt.use_spline) {
            int off = 0;
            for (int l = 0; l < L; l++) {
                const SplineChannel& ch = rt.channels[l];
                koff[l] = off;
                nk[l] = static_cast<int>(ch.positions.size());
                off += nk[l];
                for (double x : ch.positions) kpos.push_back(static_cast<T>(x));
                for (double x : ch.values) kval.push_back(static_cast<T>(x));
                for (double x : ch.derivatives) kder.push_back(static_cast<T>(x));
                // the interval estimate of the uniform grid, from the knots as stored
                const double first = static_cast<double>(kpos[koff[l]]);
                const double last = static_cast<double>(kpos[koff[l] + nk[l] - 1]);
                kgrid[2 * l] = first;
                kgrid[2 * l + 1] = last > first ? (nk[l] - 1) / (last - first) : 0.0;
            }
        }
        t->radial.max_angular = rt.max_angular;
        t->radial.n = n;
        t->radial.use_spline = rt.use_spline ? 1 : 0;
        t->radial.stop = rt.stop;
        t->radial.sigma2_atomic = rt.sigma2_atomic;
        t->radial.gto = t->put(rt.gto);
        t->radial.ortho = t->put(rt.ortho);
        t->radial.knot_offset = t->put(koff);
        t->radial.n_knots = t->put(nk);
        t->radial.knot_grid = t->put(kgrid);
        t->radial.knot_pos = t->put(kpos);
        t->radial.knot_val = t->put(kval);
        t->radial.knot_der = t->put(kder);
        t->radial.exact_val = nullptr;
        t->radial.exact_der = nullptr;
        // Legendre recursion coefficients, featomic/src/math/spherical_harmonics.rs:206-216
        std::vector<T> A(L * (L + 1) / 2 + 1, 0), B(L * (L + 1) / 2 + 1, 0);
        for (int l = 2; l < L; l++) {
            const double ls = l * l, lm1s = (l - 1) * (l - 1);
            for (int m = 0; m < l - 1; m++) {
                const double ms = m * m;
                A[l * (l + 1) / 2 + m] = static_cast<T>(std::sqrt((4.0 * ls - 1.0) / (ls - ms)));
                B[l * (l + 1) / 2 + m] = static_cast<T>(-std::sqrt((lm1s - ms) / (4.0 * lm1s - 1.0)));
            }
        }
        t->sh_a = t->put(A);
        t->sh_b = t->put(B);
        // per-(l, m) and per-l factors of the lane-parallel evaluation, each rounded
        // to T the way the serial evaluation rounds it (spherical_harmonics.rs:228-298):
        // [leg] sqrt((l+m)(l-m+1)), [leg] sqrt((l-m)(l+m+1)), [L] sqrt(2l+1),
        // [L] sqrt(1 + 1/(2l)) (the l = 1 value is sqrt(3/2)), [L] sqrt(l(l+1)/2),
        // [L] -1/(l(l+1))
        {
            const int leg = L * (L + 1) / 2;
            std::vector<T> c(2 * leg + 4 * L, static_cast<T>(0));
            for (int l = 0; l < L; l++) {
                for (int m = 0; m <= l; m++) {
                    const int k = l * (l + 1) / 2 + m;
                    c[k] = static_cast<T>(std::sqrt(static_cast<double>((l + m) * (l - m + 1))));
                    c[leg + k] = static_cast<T>(std::sqrt(static_cast<double>((l - m) * (l + m + 1))));
                }
                c[2 * leg + l] = l == 1 ? static_cast<T>(1.7320508075688772) : static_cast<T>(std::sqrt(2.0 * l + 1.0));
                if (l == 1) {
                    c[2 * leg + L + l] = static_cast<T>(1.224744871391589);
                } else if (l >= 2) {
                    c[2 * leg + L + l] = static_cast<T>(std::sqrt(1.0 + 0.5 / l));
                }
                c[2 * leg + 2 * L + l] = static_cast<T>(std::sqrt(0.5 * static_cast<double>(l * (l + 1))));
                if (l >= 1) c[2 * leg + 3 * L + l] = -static_cast<T>(1) / static_cast<T>(l * l + l);
            }
            t->sh_c = t->put(c);
        }
        std::vector<T> sc;
        for (double x : self_contribution_) sc.push_back(static_cast<T>(x));
        t->self_c = t->put(sc);
        slot = t;
    }
    return *static_cast<DeviceTables<T>*>(slot.get());
}

// ---------------------------------------------------------------------------- radial surrogate

SurrogateReport Calculator::train_radial_surrogate(const SurrogateOptions& options) {
    check_unbound(*this);
    if (kind_ == Kind::LodeSphericalExpansion || kind_ == Kind::NeighborList) {
        throw Error::invalid_parameter("the radial surrogate is only available for the SOAP calculators");
    }
    if (soap_.padded() || soap_.radial[0].tabulated) {
        throw Error::invalid_parameter("the radial surrogate learns one GTO basis shared by every angular channel");
    }
    check_surrogate_options(options);
    auto surrogate = std::make_unique<RadialSurrogate>(soap_.max_angular, soap_.n_radial, soap_.cutoff);
    const DeviceTables<float>& tab = device_tables<float>();
    SurrogateReport report = surrogate->train(tables_, tab.radial.gto, tab.radial.ortho,
                                              tables_.use_spline ? &tab.radial : nullptr, nullptr, options);
    surrogate_ = std::move(surrogate);
    use_surrogate_ = report.accepted;
    return report;
}

void Calculator::use_radial_surrogate(bool enabled) {
    if (enabled && !(surrogate_ && surrogate_->accepted())) {
        throw Error::invalid_parameter("no accepted radial surrogate: train one within its tolerance first");
    }
    use_surrogate_ = enabled;
}

// ---------------------------------------------------------------------------- SOAP pipeline

#ifndef RA_TILED_LDS_BUDGET
#define RA_TILED_LDS_BUDGET 53248
#endif
// neighbour entries per LDS chunk of the expansion kernel: small chunks leave
// room for more resident workgroups, which hides the latency of the per-entry
// LDS reads (8 entries: 27 ms per step on the 4096-structure workload, against
// 48 ms with 64)
#ifndef RA_EXPANSION_ENTRIES
#define RA_EXPANSION_ENTRIES 8
#endif
static_assert(RA_EXPANSION_ENTRIES >= 1 && RA_EXPANSION_ENTRIES <= 64, "expansion chunks hold at most one wave of entries");
// entries per chunk of the matrix-core expansion kernel, the built-in choice where it
// applies (8 or 16; 0 keeps the vector-ALU kernel). On the 4096-structure workload a
// 56 000-atom chunk takes 1.59 ms with 8, 1.68 ms with 16, and 2.37 ms on the
// vector-ALU kernel (8 entries)
#ifndef RA_EXPANSION_MFMA_ENTRIES
#define RA_EXPANSION_MFMA_ENTRIES 8
#endif
static_assert(RA_EXPANSION_MFMA_ENTRIES == 0 || RA_EXPANSION_MFMA_ENTRIES == 8 || RA_EXPANSION_MFMA_ENTRIES == 16,
              "the matrix-core expansion kernel takes chunks of 8 or 16 entries");
// (neighbour runs, entries, staged tile rows, rank-2) per window of the gradient
// kernels, tried in order until one applies and fits the LDS budget
#ifndef RA_TILED_WINDOWS
#define RA_TILED_WINDOWS {{8, 8, 1, 1}, {4, 16, 1, 0}, {4, 8, 1, 0}, {2, 8, 1, 0}, {1, 4, 1, 0}, {1, 1, 1, 0}}
#endif

// Kernel arguments of the SOAP kernels for a built neighbour list: radial tables
// (spline, exact per-entry values, or the learned integral), cutoff and scaling
// parameters, spherical-harmonics coefficients and the cleared error flag.
template <typename T>
void Calculator::soap_launch(SoapLaunch<T>& lp, const DeviceNeighbors& nl, const int* d_species, int n_atoms, int S,
                             hipStream_t stream) {
    const int L = soap_.L(), n = soap_.n_radial;
    const DeviceTables<T>& tab = device_tables<T>();
    lp.stream = stream;
    lp.n_atoms = n_atoms;
    lp.n_species = S;
    lp.L = L;
    lp.n = n;
    lp.offsets = nl.offsets;
    lp.entry_j = nl.entry_j;
    lp.entry_shift = nl.entry_shift;
    lp.entry_vec = nl.entry_vec;
    lp.entry_dist = nl.entry_dist;
    lp.species = d_species;
    lp.radial = tab.radial;
    if (radial_surrogate_active()) {
        // learned radial integral: every (l, n) value and derivative per neighbour entry
        T* ev = ws_.get<T>("exact_val", std::max<long long>(nl.n_entries, 1) * L * n);
        T* ed = ws_.get<T>("exact_der", std::max<long long>(nl.n_entries, 1) * L * n);
        surrogate_->evaluate<T>(stream, nl.n_entries, nl.entry_dist, ev, ed);
        lp.radial.use_spline = 0;
        lp.radial.exact_val = ev;
        lp.radial.exact_der = ed;
    } else if (!tab.radial.use_spline) {
        // exact radial integrals: evaluated once per neighbour entry in FP64
        T* ev = ws_.get<T>("exact_val", std::max<long long>(nl.n_entries, 1) * L * n);
        T* ed = ws_.get<T>("exact_der", std::max<long long>(nl.n_entries, 1) * L * n);
        launch_radial_exact<T>(stream, nl.n_entries, L, n, nl.entry_dist, tables_.sigma2_atomic, tab.radial.gto,
                               tab.radial.ortho, ev, ed);
        lp.radial.exact_val = ev;
        lp.radial.exact_der = ed;
    }
    lp.scalars.cutoff = soap_.cutoff;
    lp.scalars.smoothing = soap_.smoothing == Smoothing::ShiftedCosine ? 1 : 0;
    lp.scalars.width = soap_.smoothing_width;
    lp.scalars.has_scaling = soap_.has_scaling ? 1 : 0;
    lp.scalars.scale = soap_.scaling_scale;
    lp.scalars.rate = soap_.scaling_rate;
    lp.scalars.exponent = soap_.scaling_exponent;
    lp.sh_a = tab.sh_a;
    lp.sh_b = tab.sh_b;
    lp.sh_c = tab.sh_c;
    lp.self_contribution = tab.self_c;
    int* error_flag = ws_.get<int>("soap_error", 1);
    fill_zero(error_flag, sizeof(int), stream);
    lp.error_flag = error_flag;
}

template void Calculator::soap_launch<float>(SoapLaunch<float>&, const DeviceNeighbors&, const int*, int, int, hipStream_t);
template void Calculator::soap_launch<double>(SoapLaunch<double>&, const DeviceNeighbors&, const int*, int, int, hipStream_t);

template <typename T>
mts_tensormap_t* Calculator::run_soap(const HostBatch& batch, const Gradients& grads, hipStream_t stream, bool to_host,
                                      featomic_amd_device_output_t* output) {
    const bool ps = kind_ == Kind::PowerSpectrum;
    // the radial spectrum runs as the l = 0 spherical expansion, with its own labels
    const bool rs = kind_ == Kind::RadialSpectrum;
    const int n_atoms = static_cast<int>(batch.n_atoms());
    const Species sp = species_of(batch);
    const int S = static_cast<int>(sp.types.size());
    const int L = soap_.L(), n = soap_.n_radial, L2 = soap_.L2();
    const size_t slots = static_cast<size_t>(L2) * n;
    const int P = L * n * n;

    if (!ev_start_) {
        RA_HIP_CHECK(hipEventCreate(&ev_start_));
        RA_HIP_CHECK(hipEventCreate(&ev_stop_));
        RA_HIP_CHECK(hipEventCreate(&ev_main_start_));
        RA_HIP_CHECK(hipEventCreate(&ev_main_stop_));
        RA_HIP_CHECK(hipEventCreate(&ev_exp_start_));
        RA_HIP_CHECK(hipEventCreate(&ev_exp_stop_));
        RA_HIP_CHECK(hipEventCreate(&ev_grad_start_));
    }
    mark(ev_start_, stream);

    auto tm = to_host ? std::make_unique<mts_tensormap_t>() : nullptr;
    if (tm) {
        tm->keys = ps   ? LabelsData({"center_type", "neighbor_1_type", "neighbor_2_type"}, {})
                   : rs ? LabelsData({"center_type", "neighbor_type"}, {})
                        : LabelsData({"o3_lambda", "o3_sigma", "center_type", "neighbor_type"}, {});
    }
    if (n_atoms == 0) {
        mark(ev_exp_start_, stream);
        mark(ev_exp_stop_, stream);
        mark(ev_grad_start_, stream);
        mark(ev_main_start_, stream);
        mark(ev_main_stop_, stream);
        mark(ev_stop_, stream);
        main_bytes_ = 0;
        if (output) {
            device_result_ = DeviceResult();
            std::memset(output, 0, sizeof(*output));
            output->key_size = ps ? 3 : (rs ? 2 : 4);
            output->precision = sizeof(T) == 4 ? 32 : 64;
        }
        return tm.release();
    }

    // ---- geometry upload
    std::vector<int32_t> atom_system(n_atoms), atom_local(n_atoms);
    for (size_t s = 0; s < batch.n_systems(); s++) {
        for (int64_t a = batch.offsets[s]; a < batch.offsets[s + 1]; a++) {
            atom_system[a] = static_cast<int32_t>(s);
            atom_local[a] = static_cast<int32_t>(a - batch.offsets[s]);
        }
    }
    const double* d_positions = batch.device_positions;
    if (!d_positions) {
        if (mode_ != PassMode::Eager) throw Error::invalid_parameter("batch graphs need device-resident positions");
        double* p = ws_.get<double>("positions", 3 * static_cast<size_t>(n_atoms));
        upload(p, batch.positions, stream);
        d_positions = p;
    }
    const int* d_atom_system = batch_input("atom_system", atom_system, stream);
    const int* d_atom_local = batch_input("atom_local", atom_local, stream);
    const int* d_species = batch_input("species", sp.index, stream);

    // ---- neighbour list
    PassContext pass;
    pass.mode = mode_;
    pass.record = record_;
    pass.replay_flags = replay_flags_;
    const DeviceNeighbors nl = build_neighbors(ws_, batch, soap_.cutoff, d_positions, d_atom_system, stream, pass);

    // ---- metadata: block membership, sample and gradient-row offsets
    unsigned* type_mask = ws_.get<unsigned>("type_mask", n_atoms);
    int* unique_count = ws_.get<int>("unique_count", static_cast<size_t>(n_atoms) * S);
    launch_center_meta(stream, n_atoms, S, nl.offsets, nl.entry_j, d_species, type_mask, unique_count);
    const int npairs = S * (S + 1) / 2;
    const int NB = ps ? S * npairs : S * S;
    std::vector<int32_t> block_types;  // (ct, t1, t2) per block, species indices
    for (int ct = 0; ct < S; ct++) {
        if (ps) {
            for (int t1 = 0; t1 < S; t1++)
                for (int t2 = t1; t2 < S; t2++) block_types.insert(block_types.end(), {ct, t1, t2});
        } else {
            for (int nt = 0; nt < S; nt++) block_types.insert(block_types.end(), {ct, nt, nt});
        }
    }
    const int* d_block_types = batch_input("block_types", block_types, stream);
    const size_t NBN = static_cast<size_t>(NB) * n_atoms;
    long long* sample_flag = ws_.get<long long>("sample_flag", NBN);
    long long* row_count = ws_.get<long long>("row_count", NBN);
    long long* sample_scan = ws_.get<long long>("sample_scan", NBN + 1);
    long long* row_scan = ws_.get<long long>("row_scan", NBN + 1);
    launch_block_membership(stream, n_atoms, S, ps ? 1 : 0, NB, d_species, type_mask, unique_count, d_block_types,
                            sample_flag, row_count);
    exclusive_scan(stream, sample_flag, sample_scan, static_cast<long long>(NBN), ws_);
    exclusive_scan(stream, row_count, row_scan, static_cast<long long>(NBN), ws_);
    long long* d_totals = ws_.get<long long>("block_totals", 2 * (NB + 1));
    launch_gather_block_totals(stream, NB, n_atoms, sample_scan, row_scan, d_totals);
    std::vector<long long> totals;
    if (mode_ == PassMode::Replay) {
        // the recorded block sizes; different ones are reported by the graph's check
        totals = record_->block_totals;
        if (totals.size() != static_cast<size_t>(2 * (NB + 1))) throw Error::internal("replayed batch differs from its record (blocks)");
        check_equal_array(stream, d_totals, batch_input("block_totals", totals, stream), 2 * (NB + 1), REPLAY_BLOCKS_CHANGED,
                          replay_flags_);
    } else {
        download(totals, d_totals, 2 * (NB + 1), stream);
        RA_HIP_CHECK(hipStreamSynchronize(stream));
        if (mode_ == PassMode::Record) {
            record_->block_totals = totals;
            batch_input("block_totals", totals, stream);
        }
    }
    std::vector<long long> block_sample_base(NB + 1), block_row_base(NB + 1);
    for (int b = 0; b <= NB; b++) {
        block_sample_base[b] = totals[2 * b];
        block_row_base[b] = totals[2 * b + 1];
    }
    const long long total_samples = block_sample_base[NB];
    const long long total_rows = block_row_base[NB];
    const long long* d_block_sample_base = batch_input("block_sample_base", block_sample_base, stream);

    // ---- kernel launch parameters
    SoapLaunch<T> lp;
    soap_launch<T>(lp, nl, d_species, n_atoms, S, stream);
    const DeviceTables<T>& tab = device_tables<T>();
    int* error_flag = lp.error_flag;

    const int lds = device_lds_limit();
    const size_t per_entry = soap_chunk_bytes<T>(L, n, 1);
    const SoapTuning tune = tuning;
    auto chunk_for = [&](size_t fixed) -> int {
        const long long budget = static_cast<long long>(lds) - static_cast<long long>(fixed) - 1024;
        const long long c = budget / static_cast<long long>(per_entry);
        if (c < 1) {
            throw Error::invalid_parameter("max_angular / max_radial too large for the GPU kernels' local memory (" +
                                           std::to_string(fixed + per_entry) + " bytes needed, " + std::to_string(lds) +
                                           " available)");
        }
        return static_cast<int>(std::min<long long>(c, 64));
    };

    // ---- spherical expansion per center (dense)
    T* coef = ws_.get<T>("coef", static_cast<size_t>(n_atoms) * S * slots);
    T* dself = grads.positions ? ws_.get<T>("dself", static_cast<size_t>(n_atoms) * S * 3 * slots) : nullptr;
    {
        // entries per chunk: up to 64, with at most ~48 KB of LDS per workgroup
        const long long target = std::min<long long>(lds, 49152) - (2048 + 64) * 4 - 1024;
        const long long c = std::max<long long>(1, std::min<long long>(64, target / static_cast<long long>(per_entry)));
        lp.chunk = static_cast<int>(c);
    }
    // per-entry data (Legendre functions, direction, cutoff factors, radial values)
    // shared by the expansion and the power-spectrum gradient kernels
    const bool tiled = n <= 16 && L <= 16 && S <= 32;
    T* entry_data = nullptr;
    // the power spectrum's values come out of the matrix-core expansion kernel itself
    bool fused_values = false;
    if (tiled) {
        entry_data = ws_.get<T>("entry_data", std::max<size_t>(entry_prep_size<T>(lp, nl.n_entries), 1));
        launch_entry_prep<T>(lp, nl.n_entries, entry_data);
        // entries per chunk: the tuned or built-in count, halved until two workgroups fit
        // per CU (large bases: the per-type sums alone take ~62 KB at n = 12, l <= 8)
        // a negative count asks for the matrix-core kernel with that many entries per
        // chunk (FP32, n = 8, l_max < 9; elsewhere the tiled kernel with as many)
        const int asked = tune.expansion_entries < 0 ? -tune.expansion_entries : tune.expansion_entries;
        const bool mfma = std::is_same<T, float>::value && expansion_mfma_supported(L, n, S) &&
                          (tune.expansion_entries < 0 || (tune.expansion_entries == 0 && RA_EXPANSION_MFMA_ENTRIES > 0));
        int max_entries = asked > 0 ? asked : RA_EXPANSION_ENTRIES;
        while (max_entries > 1 && expansion_tiled_shmem<T>(lp, max_entries) + 1024 > static_cast<size_t>(lds) / 2) {
            max_entries /= 2;
        }
        mark(ev_exp_start_, stream);
        if (mfma) {
            if constexpr (std::is_same<T, float>::value) {
                fused_values = ps && !vjp_grad_ && expansion_mfma_values_supported(S);
                T* values = fused_values ? ws_.get<T>("out_values", std::max<long long>(total_samples * P, 1)) : nullptr;
                // the fast form's values: products over m in bfloat16 hi + lo
                launch_expansion_mfma(lp, asked > 0 ? asked : RA_EXPANSION_MFMA_ENTRIES, entry_data, coef, dself,
                                      type_mask, sample_flag, sample_scan, values, fast_form != 0);
            }
        } else if (expansion_tiled_shmem<T>(lp, max_entries) + 1024 <= static_cast<size_t>(lds)) {
            launch_expansion_tiled<T>(lp, max_entries, entry_data, coef, dself);
        } else {
            launch_soap_center<T>(lp, 0, coef, dself);
        }
    } else {
        mark(ev_exp_start_, stream);
        launch_soap_center<T>(lp, 0, coef, dself);
    }
    mark(ev_exp_stop_, stream);
    T* dstrain = nullptr;
    T* dcell = nullptr;
    if (grads.strain) {
        dstrain = ws_.get<T>("dstrain", static_cast<size_t>(n_atoms) * S * 9 * slots);
        launch_soap_center<T>(lp, 1, nullptr, dstrain);
    }
    if (grads.cell) {
        dcell = ws_.get<T>("dcell", static_cast<size_t>(n_atoms) * S * 9 * slots);
        launch_soap_center<T>(lp, 2, nullptr, dcell);
    }
    RA_HIP_CHECK(hipGetLastError());

    if (vjp_grad_) {
        if (mode_ != PassMode::Eager) throw Error::invalid_parameter("batch graphs compute values and gradient rows only");
        // vector-Jacobian product of the position gradients: dA/dc, then per-entry
        // contributions and their per-atom sums; no gradient rows
        if (!entry_data) throw Error::invalid_parameter("the positions vector-Jacobian product needs max_radial < 16, max_angular < 16 and at most 32 atomic types");
        const T* W = static_cast<const T*>(vjp_grad_);
        T* H = ws_.get<T>("vjp_h", static_cast<size_t>(n_atoms) * S * slots);
        if (ps) {
            if (ps_vjp_coef_shmem<T>(lp) + 1024 > static_cast<size_t>(lds)) {
                throw Error::invalid_parameter("too many atomic types or radial / angular functions for the GPU kernels' local memory");
            }
            launch_ps_vjp_coef<T>(lp, coef, type_mask, sample_flag, sample_scan, W, H);
        } else {
            std::vector<long long> vb(static_cast<size_t>(NB) * L);
            long long size = 0;
            for (int tb = 0; tb < NB; tb++) {
                const long long ns = block_sample_base[tb + 1] - block_sample_base[tb];
                for (int l = 0; l < L; l++) {
                    vb[tb * L + l] = size;
                    size += ns * (2LL * l + 1) * n;
                }
            }
            long long* d_vb = ws_.get<long long>("vjp_value_base", vb.size());
            upload(d_vb, vb, stream);
            launch_se_vjp_coef<T>(lp, sample_flag, sample_scan, d_block_sample_base, d_vb, W, H);
        }
        int vjp_entries = 8;
        while (vjp_entries > 1 && vjp_pair_shmem<T>(lp, vjp_entries) + 1024 > static_cast<size_t>(std::min(lds, 65536))) vjp_entries /= 2;
        if (vjp_pair_shmem<T>(lp, vjp_entries) + 1024 > static_cast<size_t>(lds)) {
            throw Error::invalid_parameter("too many atomic types or radial / angular functions for the GPU kernels' local memory");
        }
        T* g = ws_.get<T>("vjp_g", 3 * static_cast<size_t>(std::max<long long>(nl.n_entries, 1)));
        mark(ev_main_start_, stream);
        launch_vjp_pair<T>(lp, vjp_entries, entry_data, H, g);
        launch_vjp_gather<T>(stream, n_atoms, nl.offsets, nl.entry_j, nl.entry_shift, g, static_cast<T*>(vjp_out_));
        mark(ev_main_stop_, stream);
        RA_HIP_CHECK(hipGetLastError());
        mark(ev_stop_, stream);
        // algorithmic bytes: dA/dX read once, dA/dr written once, neighbour entries read once
        main_bytes_ = (ps ? total_samples * P : total_samples * static_cast<long long>(L2) * n) * static_cast<long long>(sizeof(T)) +
                      3LL * n_atoms * static_cast<long long>(sizeof(T)) + nl.n_entries * 48;
        return nullptr;
    }

    // ---- outputs
    int* sample_labels = ws_.get<int>("sample_labels", 2 * static_cast<size_t>(std::max<long long>(total_samples, 1)));
    launch_sample_labels(stream, NB, n_atoms, sample_flag, sample_scan, d_atom_system, d_atom_local, sample_labels);
    int* grad_labels = grads.positions ? ws_.get<int>("grad_labels", 3 * static_cast<size_t>(std::max<long long>(total_rows, 1))) : nullptr;

    // value / gradient regions (elements)
    std::vector<long long> value_base, grad_base, cs_base;  // PS: per block; SE: per (tb, l)
    long long values_size = 0, grads_size = 0, cs_size = 0;
    std::vector<long long> se_grad_adj;  // SE: per (tb, l), base - row_base(tb) * 3w
    if (ps) {
        values_size = total_samples * P;
        grads_size = total_rows * 3 * P;
        cs_size = total_samples * 9 * P;
    } else {
        value_base.resize(static_cast<size_t>(NB) * L);
        grad_base.resize(static_cast<size_t>(NB) * L);
        cs_base.resize(static_cast<size_t>(NB) * L);
        se_grad_adj.resize(static_cast<size_t>(NB) * L);
        for (int tb = 0; tb < NB; tb++) {
            const long long ns = block_sample_base[tb + 1] - block_sample_base[tb];
            const long long nr = block_row_base[tb + 1] - block_row_base[tb];
            for (int l = 0; l < L; l++) {
                const long long w = (2LL * l + 1) * n;
                value_base[tb * L + l] = values_size;
                values_size += ns * w;
                grad_base[tb * L + l] = grads_size;
                se_grad_adj[tb * L + l] = grads_size - block_row_base[tb] * 3 * w;
                grads_size += nr * 3 * w;
                cs_base[tb * L + l] = cs_size;
                cs_size += ns * 9 * w;
            }
        }
    }
    T* d_values = ws_.get<T>("out_values", std::max<long long>(values_size, 1));
    T* d_grads = grads.positions ? ws_.get<T>("out_grads", std::max<long long>(grads_size, 1)) : nullptr;
    T* d_strain = grads.strain ? ws_.get<T>("out_strain", std::max<long long>(cs_size, 1)) : nullptr;
    T* d_cell = grads.cell ? ws_.get<T>("out_cell", std::max<long long>(cs_size, 1)) : nullptr;
    long long* d_region = nullptr;
    long long* d_se_grad = nullptr;
    if (!ps) {
        d_region = batch_input("se_value_base", value_base, stream);
        d_se_grad = batch_input("se_grad_base", se_grad_adj, stream);
    }

    const size_t slots_bytes = slots * sizeof(T);
    if (ps) {
        if (static_cast<long long>(S * slots_bytes) > lds) throw Error::invalid_parameter("too many atomic types for the GPU kernels' local memory");
        if (!fused_values) launch_ps_values<T>(lp, coef, type_mask, sample_flag, sample_scan, d_values, fast_form != 0, lds);
    } else {
        launch_se_gather<T>(lp, 1, coef, sample_flag, sample_scan, d_block_sample_base, d_region, d_values);
    }
    bool timed_start = false;
    mark(ev_grad_start_, stream);
    if (!grads.positions) {
        mark(ev_main_start_, stream);
        timed_start = true;
    }
    if (grads.positions) {
        bool done = false;
        if (total_rows < (1ll << 31) && tune.gradient_window != -2 && entry_data) {
            // planned, windowed tiled kernel: the first (runs, entries, staged rows)
            // configuration that keeps the workgroup within the LDS budget (4
            // workgroups per CU in FP32, 2 in FP64), else the first within the whole
            // LDS (large bases: one or two workgroups per CU, still well ahead of the
            // register-blocked kernel), or the tuned shape within the whole LDS
            // (the rank-2 MFMA kernel where it applies: FP32 power spectrum with n = 8 in
            // the plain form, the tiled kernel taking the centers with a run longer
            // than a window)
            static const int default_windows[][4] = RA_TILED_WINDOWS;
            const size_t whole = lds > 1024 ? static_cast<size_t>(lds) - 1024 : 0;
            std::vector<std::pair<std::array<int, 4>, size_t>> windows;
            const bool explicit_shape =
                tune.gradient_window >= 0 && tune.gradient_window < static_cast<int>(soap_window_shapes().size());
            if (explicit_shape) {
                windows.push_back({soap_window_shapes()[tune.gradient_window], whole});
            } else {
                const size_t cap = std::min<size_t>(whole, sizeof(T) == 4 ? RA_TILED_LDS_BUDGET : 81920);
                for (const auto& w : default_windows) windows.push_back({{w[0], w[1], w[2], w[3]}, cap});
                for (const auto& w : default_windows) windows.push_back({{w[0], w[1], w[2], w[3]}, whole});
            }
            for (const auto& [w, budget] : windows) {
                // (in both forms: its contraction is on MFMA with FP32 operands)
                const bool rank2 = w[3] != 0 && ps && sizeof(T) == 4 && ps_gradient_rank2_supported(L, n, S, w[1]);
                // a rank-2 shape of the built-in table is for the rank-2 kernel only; chosen
                // explicitly, it runs the tiled kernel where the rank-2 one does not apply
                if (w[3] != 0 && !rank2 && !explicit_shape) continue;
                if (!ps_gradient_tiled_supported(L, n, S, w[0], w[1], w[2])) continue;
                // the spherical expansion's rows are the stage-2 sums themselves: no
                // coefficients or store staging in LDS, one staged row per pass
                if (!ps && w[2] != 1) continue;
                const size_t shmem = ps ? ps_gradient_tiled_lds<T>(lp, w[0], w[1], w[2])
                                        : se_gradient_tiled_shmem<T>(lp, w[0], w[1]);
                if (rank2) {
                    // the tiled kernel only takes the few centers with long runs
                    if (shmem > whole) continue;
                    size_t rank2_lds = 0;
                    if constexpr (std::is_same<T, float>::value) rank2_lds = ps_gradient_rank2_lds(lp, w[1]);
                    if (rank2_lds > budget) continue;
                } else if (shmem > budget) {
                    continue;
                }
                const long long ne = std::max<long long>(nl.n_entries, 1);
                const int nloc = ps ? S * (S + 1) / 2 : S;
                int* d_runs = ws_.get<int>("plan_runs", ne * 4);
                int* d_run_rows = ws_.get<int>("plan_run_rows", ne * S);
                int* d_windows = ws_.get<int>("plan_windows", ne * 4);
                int* d_n_windows = ws_.get<int>("plan_n_windows", std::max<long long>(n_atoms, 1));
                int* d_self_rows = ws_.get<int>("plan_self_rows", std::max<long long>(n_atoms, 1) * nloc);
                int* d_split = rank2 ? ws_.get<int>("plan_split", std::max<long long>(n_atoms, 1)) : nullptr;
                launch_ps_grad_plan(stream, n_atoms, S, nl.offsets, nl.entry_j, d_species, type_mask, sample_scan,
                                    row_scan, d_block_sample_base, d_atom_system, d_atom_local, w[0], w[1], d_runs,
                                    d_run_rows, d_windows, d_n_windows, d_self_rows, grad_labels, ps, d_split);
                mark(ev_main_start_, stream);
                timed_start = true;
                if (rank2) {
                    if constexpr (std::is_same<T, float>::value) {
                        launch_ps_gradient_tiled<T>(lp, w[0], w[1], w[2], coef, dself, type_mask, d_runs, d_run_rows,
                                                    d_windows, d_n_windows, d_self_rows, entry_data, d_grads, false,
                                                    d_split);
                        launch_ps_gradient_rank2(lp, w[1], coef, dself, type_mask, d_runs, d_run_rows, d_windows,
                                                 d_n_windows, d_self_rows, d_split, entry_data, d_grads);
                    }
                } else if (ps) {
                    launch_ps_gradient_tiled<T>(lp, w[0], w[1], w[2], coef, dself, type_mask, d_runs, d_run_rows,
                                                d_windows, d_n_windows, d_self_rows, entry_data, d_grads, fast_form != 0);
                } else {
                    launch_se_gradient_tiled<T>(lp, w[0], w[1], dself, type_mask, d_runs, d_run_rows, d_windows,
                                                d_n_windows, d_self_rows, d_se_grad, entry_data, d_grads);
                }
                done = true;
                break;
            }
        }
        if (!timed_start) mark(ev_main_start_, stream);
        if (ps && !done) {
            // register-blocked kernel; chunk sized for ~48 KB of LDS per workgroup
            const size_t fixed = ps_gradient_shmem<T>(lp, 0);
            const size_t per = ps_gradient_shmem<T>(lp, 1) - fixed;
            // chunk: at most 8 entries (so that ~5 workgroups fit per CU) and L * chunk
            // radial threads in the last three waves
            const long long target = std::min<long long>(lds, 32768) - 1024;
            long long c = (target - static_cast<long long>(fixed)) / static_cast<long long>(per);
            c = std::max<long long>(1, std::min<long long>(c, std::min<long long>(8, 192 / L)));
            if (static_cast<long long>(fixed + per * c) + 1024 <= lds) {
                done = launch_ps_gradient_fast<T>(lp, static_cast<int>(c), coef, dself, type_mask, sample_scan, row_scan,
                                                  d_block_sample_base, d_atom_system, d_atom_local, d_grads, grad_labels);
            }
        }
        if (!done) {
            lp.chunk = chunk_for((ps ? S * slots_bytes : 0) + 6 * slots_bytes + 2 * 528 * 4);
            launch_gradient_rows<T>(lp, ps, coef, dself, type_mask, sample_flag, sample_scan, row_scan, d_block_sample_base,
                                    d_se_grad, d_atom_system, d_atom_local, d_grads, grad_labels);
        }
    }
    mark(ev_main_stop_, stream);
    // algorithmic bytes of the gradient-row kernel: gradient values and labels written once,
    // per-center coefficients and self gradients read once, neighbour entries read once
    main_bytes_ = grads.positions ? grads_size * static_cast<long long>(sizeof(T)) + total_rows * 12 +
                                        static_cast<long long>(n_atoms) * S * static_cast<long long>(slots) * 4 * sizeof(T) +
                                        nl.n_entries * 48
                                  : 0;
    for (int which = 0; which < 2; which++) {
        const T* dense = which == 0 ? dstrain : dcell;
        T* out = which == 0 ? d_strain : d_cell;
        if (!dense) continue;
        if (ps) {
            if (static_cast<long long>((S + 18) * slots_bytes) > lds) throw Error::invalid_parameter("strain/cell gradients too large for the GPU kernels' local memory");
            launch_ps_cell_strain<T>(lp, coef, dense, type_mask, sample_flag, sample_scan, out);
        } else {
            long long* d_cs = ws_.get<long long>(which == 0 ? "se_strain_base" : "se_cell_base", cs_base.size());
            upload(d_cs, cs_base, stream);
            launch_se_gather<T>(lp, 9, dense, sample_flag, sample_scan, d_block_sample_base, d_cs, out);
        }
    }
    RA_HIP_CHECK(hipGetLastError());
    if (mode_ == PassMode::Replay) or_flag(stream, error_flag, REPLAY_KERNEL_ERROR, replay_flags_);
    mark(ev_stop_, stream);

    // ---- describe the blocks
    std::vector<int32_t> keys;
    std::vector<int64_t> out_ns, out_nr, out_voff, out_goff, out_props, out_slab, out_glab, out_coff;
    struct BlockRef {
        int b, l;
    };
    std::vector<BlockRef> refs;
    for (int b = 0; b < NB; b++) {
        const long long ns = block_sample_base[b + 1] - block_sample_base[b];
        if (ns == 0) continue;
        const long long nr = block_row_base[b + 1] - block_row_base[b];
        const int ct = block_types[3 * b], t1 = block_types[3 * b + 1], t2 = block_types[3 * b + 2];
        if (ps) {
            keys.insert(keys.end(), {sp.types[ct], sp.types[t1], sp.types[t2]});
            out_ns.push_back(ns);
            out_nr.push_back(nr);
            out_voff.push_back(block_sample_base[b] * P);
            out_goff.push_back(block_row_base[b] * 3 * P);
            out_props.push_back(P);
            out_slab.push_back(block_sample_base[b]);
            out_glab.push_back(block_row_base[b]);
            out_coff.push_back(block_sample_base[b] * 9 * P);
            refs.push_back({b, 0});
        } else {
            for (int l = 0; l < L; l++) {
                if (rs) {
                    keys.insert(keys.end(), {sp.types[ct], sp.types[t1]});
                } else {
                    keys.insert(keys.end(), {l, 1, sp.types[ct], sp.types[t1]});
                }
                out_ns.push_back(ns);
                out_nr.push_back(nr);
                out_voff.push_back(value_base[b * L + l]);
                out_goff.push_back(grad_base[b * L + l]);
                out_props.push_back((2LL * l + 1) * n);
                // the L blocks of one (center type, neighbour type) pair share its labels
                out_slab.push_back(block_sample_base[b]);
                out_glab.push_back(block_row_base[b]);
                out_coff.push_back(cs_base.empty() ? 0 : cs_base[b * L + l]);
                refs.push_back({b, l});
            }
        }
    }

    if (output) {
        device_result_.keys = keys;
        device_result_.key_size = ps ? 3 : (rs ? 2 : 4);
        device_result_.n_samples = out_ns;
        device_result_.n_rows = out_nr;
        device_result_.value_offsets = out_voff;
        device_result_.gradient_offsets = out_goff;
        device_result_.block_properties = out_props;
        device_result_.sample_label_offsets = out_slab;
        device_result_.gradient_label_offsets = out_glab;
        device_result_.cell_gradient_offsets = out_coff;
        output->n_blocks = refs.size();
        output->keys = device_result_.keys.data();
        output->key_size = device_result_.key_size;
        output->n_samples = device_result_.n_samples.data();
        output->n_gradient_rows = device_result_.n_rows.data();
        output->value_offsets = device_result_.value_offsets.data();
        output->gradient_offsets = device_result_.gradient_offsets.data();
        output->block_properties = device_result_.block_properties.data();
        output->values = d_values;
        output->gradients = d_grads;
        output->sample_labels = sample_labels;
        output->gradient_labels = grad_labels;
        output->sample_label_offsets = device_result_.sample_label_offsets.data();
        output->gradient_label_offsets = grad_labels ? device_result_.gradient_label_offsets.data() : nullptr;
        output->cell_gradients = d_cell;
        output->cell_gradient_offsets = d_cell ? device_result_.cell_gradient_offsets.data() : nullptr;
        output->precision = sizeof(T) == 4 ? 32 : 64;
        if (record_) {
            // the batch's output description, kept for the graph's replays (the
            // pointers of the capture pass are the ones replays write through)
            record_->keys = device_result_.keys;
            record_->key_size = device_result_.key_size;
            record_->n_samples = device_result_.n_samples;
            record_->n_rows = device_result_.n_rows;
            record_->value_offsets = device_result_.value_offsets;
            record_->gradient_offsets = device_result_.gradient_offsets;
            record_->block_properties = device_result_.block_properties;
            record_->sample_label_offsets = device_result_.sample_label_offsets;
            record_->gradient_label_offsets = device_result_.gradient_label_offsets;
            record_->cell_gradient_offsets = device_result_.cell_gradient_offsets;
            record_->output = *output;
        }
    }
    if (!to_host) return nullptr;

    // ---- download the labels, then assemble the TensorMap: the blocks' float64 arrays
    // are written by the staged transfers below
    std::vector<int32_t> h_slabels, h_glabels;
    download(h_slabels, sample_labels, 2 * total_samples, stream);
    if (grad_labels) download(h_glabels, grad_labels, 3 * total_rows, stream);
    int flag = 0;
    RA_HIP_CHECK(hipMemcpyAsync(&flag, error_flag, sizeof(int), hipMemcpyDeviceToHost, stream));
    RA_HIP_CHECK(hipStreamSynchronize(stream));
    if (flag) throw Error::invalid_parameter("an atom has more than 2048 neighbours within the cutoff");

    const size_t key_size = ps ? 3 : (rs ? 2 : 4);
    tm->keys.values = keys;
    std::vector<int32_t> prop_values;
    if (ps) {
        for (int l = 0; l < L; l++)
            for (int n1 = 0; n1 < n; n1++)
                for (int n2 = 0; n2 < n; n2++) prop_values.insert(prop_values.end(), {l, n1, n2});
    } else {
        for (int k = 0; k < n; k++) prop_values.push_back(k);
    }
    const LabelsData properties = ps ? LabelsData({"l", "n_1", "n_2"}, prop_values) : LabelsData({"n"}, prop_values);
    const LabelsData xyz = LabelsData::range("xyz", 3);
    std::vector<HostSegment> seg_values, seg_grads, seg_strain, seg_cell;

    for (size_t k = 0; k < refs.size(); k++) {
        const int b = refs[k].b, l = refs[k].l;
        const long long ns = out_ns[k], nr = out_nr[k];
        const long long s0 = block_sample_base[b], r0 = block_row_base[b];
        const long long w = out_props[k];
        std::vector<int32_t> sl(h_slabels.begin() + 2 * s0, h_slabels.begin() + 2 * (s0 + ns));
        std::vector<LabelsData> comps;
        std::vector<uintptr_t> shape;
        if (ps) {
            shape = {static_cast<uintptr_t>(ns), static_cast<uintptr_t>(P)};
        } else if (rs) {
            shape = {static_cast<uintptr_t>(ns), static_cast<uintptr_t>(n)};
        } else {
            comps.push_back(mu_labels(l));
            shape = {static_cast<uintptr_t>(ns), static_cast<uintptr_t>(2 * l + 1), static_cast<uintptr_t>(n)};
        }
        auto block = make_block(LabelsData({"system", "atom"}, sl), comps, properties, shape);
        seg_values.push_back({out_voff[k], ns * w, block->array.data.data()});

        if (grads.positions) {
            std::vector<int32_t> gl(h_glabels.begin() + 3 * r0, h_glabels.begin() + 3 * (r0 + nr));
            std::vector<LabelsData> gcomps = {xyz};
            for (auto& c : comps) gcomps.push_back(c);
            std::vector<uintptr_t> gshape = {static_cast<uintptr_t>(nr), 3};
            for (size_t d = 1; d < shape.size(); d++) gshape.push_back(shape[d]);
            auto g = make_block(LabelsData({"sample", "system", "atom"}, gl), gcomps, properties, gshape);
            seg_grads.push_back({out_goff[k], nr * 3 * w, g->array.data.data()});
            block->add_gradient("positions", std::move(g));
        }
        for (int which = 0; which < 2; which++) {
            const bool want = which == 0 ? grads.cell : grads.strain;
            if (!want) continue;
            std::vector<LabelsData> gcomps;
            if (which == 0) {
                gcomps = {xyz, LabelsData::range("abc", 3)};
            } else {
                gcomps = {LabelsData::range("xyz_2", 3), LabelsData::range("xyz_1", 3)};
            }
            for (auto& c : comps) gcomps.push_back(c);
            std::vector<uintptr_t> gshape = {static_cast<uintptr_t>(ns), 3, 3};
            for (size_t d = 1; d < shape.size(); d++) gshape.push_back(shape[d]);
            auto g = make_block(LabelsData::range("sample", static_cast<int>(ns)), gcomps, properties, gshape);
            const long long base = ps ? s0 * 9 * P : cs_base[b * L + l];
            (which == 0 ? seg_cell : seg_strain).push_back({base, ns * 9 * w, g->array.data.data()});
            block->add_gradient(which == 0 ? "cell" : "strain", std::move(g));
        }
        tm->blocks.push_back(std::move(block));
    }
    download_to_blocks<T>(d_values, std::move(seg_values), stream, staging_);
    if (d_grads) download_to_blocks<T>(d_grads, std::move(seg_grads), stream, staging_);
    if (d_strain) download_to_blocks<T>(d_strain, std::move(seg_strain), stream, staging_);
    if (d_cell) download_to_blocks<T>(d_cell, std::move(seg_cell), stream, staging_);
    (void)key_size;
    return tm.release();
}

// ---------------------------------------------------------------------------- record / replay

void Calculator::set_pass(PassMode mode, BatchRecord* record, int* replay_flags) {
    if (mode != PassMode::Eager && !record) throw Error::internal("record and replay passes need a batch record");
    mode_ = mode;
    record_ = mode == PassMode::Eager ? nullptr : record;
    replay_flags_ = mode == PassMode::Replay ? replay_flags : nullptr;
    ws_.frozen = mode == PassMode::Replay;
}

template <typename U>
U* Calculator::batch_input(const std::string& name, const std::vector<U>& host, hipStream_t stream) {
    if (mode_ == PassMode::Eager) {
        U* p = ws_.get<U>(name, host.size());
        upload(p, host, stream);
        return p;
    }
    U* p = record_->inputs.get<U>(name, host.size());
    if (mode_ == PassMode::Record) {
        record_->input_counts[name] = host.size();
        upload(p, host, stream);
    } else {
        const auto it = record_->input_counts.find(name);
        if (it == record_->input_counts.end() || it->second != host.size()) {
            throw Error::internal("replayed batch differs from its record (" + name + ")");
        }
    }
    return p;
}

void Calculator::mark(hipEvent_t event, hipStream_t stream) {
    if (mode_ != PassMode::Replay) RA_HIP_CHECK(hipEventRecord(event, stream));
}

void BatchRecord::describe(featomic_amd_device_output_t* out) const {
    *out = output;
    out->n_blocks = n_samples.size();
    out->keys = keys.data();
    out->key_size = key_size;
    out->n_samples = n_samples.data();
    out->n_gradient_rows = n_rows.data();
    out->value_offsets = value_offsets.data();
    out->gradient_offsets = gradient_offsets.data();
    out->block_properties = block_properties.data();
    out->sample_label_offsets = sample_label_offsets.data();
    out->gradient_label_offsets = output.gradient_label_offsets ? gradient_label_offsets.data() : nullptr;
    out->cell_gradient_offsets = output.cell_gradients ? cell_gradient_offsets.data() : nullptr;
}

static void check_unbound(const Calculator& c) {
    if (c.bound_to_graph) {
        throw Error::invalid_parameter("this calculator runs inside a batch graph: free the graph before using it directly");
    }
}

// ---------------------------------------------------------------------------- entry points

mts_tensormap_t* Calculator::compute(const HostBatch& batch, const Gradients& grads) {
    check_unbound(*this);
    hipStream_t stream = nullptr;
    if (kind_ == Kind::NeighborList) return run_neighbor_list(batch, grads);
    if (kind_ == Kind::ByPair) {
        mts_tensormap_t* tm = precision == 32 ? run_by_pair<float>(batch, grads) : run_by_pair<double>(batch, grads);
        if (soap_.padded()) tm = restrict_to_basis(tm, soap_.lambdas, [this](int l) { return soap_.n_of(l); });
        return tm;
    }
    if (kind_ == Kind::LodeSphericalExpansion) {
        if (grads.strain || grads.cell) {
            throw Error::invalid_parameter("the lode spherical expansion calculator does not support gradients with respect to " +
                                           std::string(grads.cell ? "cell" : "strain"));
        }
        return precision == 32 ? run_lode<float>(batch, grads, stream, nullptr)
                               : run_lode<double>(batch, grads, stream, nullptr);
    }
    mts_tensormap_t* tm = precision == 32 ? run_soap<float>(batch, grads, stream, true, nullptr)
                                          : run_soap<double>(batch, grads, stream, true, nullptr);
    if (soap_.padded()) tm = restrict_to_basis(tm, soap_.lambdas, [this](int l) { return soap_.n_of(l); });
    return tm;
}

void Calculator::compute_device(const HostBatch& batch, const Gradients& grads, hipStream_t stream,
                                featomic_amd_device_output_t* output) {
    if (mode_ == PassMode::Eager) check_unbound(*this);
    if (mode_ != PassMode::Eager && kind_ != Kind::SphericalExpansion && kind_ != Kind::PowerSpectrum) {
        throw Error::invalid_parameter("batch graphs support the SOAP spherical expansion and power spectrum");
    }
    if (kind_ == Kind::NeighborList || kind_ == Kind::ByPair) {
        throw Error::invalid_parameter(std::string("the ") + name() +
                                       " calculator returns host TensorMaps only (use featomic_calculator_compute)");
    }
    if (kind_ == Kind::LodeSphericalExpansion) {
        if (grads.strain || grads.cell) {
            throw Error::invalid_parameter("the lode spherical expansion calculator does not support gradients with respect to " +
                                           std::string(grads.cell ? "cell" : "strain"));
        }
        if (precision == 32) {
            run_lode<float>(batch, grads, stream, output);
        } else {
            run_lode<double>(batch, grads, stream, output);
        }
        return;
    }
    if (soap_.padded()) {
        throw Error::invalid_parameter(
            "an Explicit basis whose channels differ from a tensor product is computed in the padded tensor-product "
            "layout: use featomic_calculator_compute for its TensorMap");
    }
    main_ms_sum_ = -1.0;
    if (precision == 32) {
        run_soap<float>(batch, grads, stream, false, output);
    } else {
        run_soap<double>(batch, grads, stream, false, output);
    }
}

const std::vector<std::array<int, 4>>& soap_window_shapes() {
    static const std::vector<std::array<int, 4>> shapes = {
        {4, 16, 1, 0}, {4, 8, 1, 0}, {2, 8, 1, 0}, {1, 4, 1, 0}, {1, 1, 1, 0},  {8, 16, 1, 0}, {4, 32, 1, 0},
        {2, 16, 1, 0}, {4, 16, 2, 0}, {4, 4, 1, 0}, {8, 32, 1, 0}, {8, 8, 1, 1}, {8, 16, 1, 1},
    };
    return shapes;
}

std::string Calculator::autotune(const HostBatch& batch, const Gradients& grads, hipStream_t stream, int repeats) {
    check_unbound(*this);
    if (kind_ == Kind::LodeSphericalExpansion || kind_ == Kind::NeighborList || kind_ == Kind::ByPair) {
        throw Error::invalid_parameter("the autotuner covers the SOAP calculators");
    }
    if (repeats < 1 || repeats > 100) throw Error::invalid_parameter("autotune repeats must be in [1, 100]");
    featomic_amd_device_output_t out{};
    // best time of one phase over `repeats` runs after a warm-up run (which also sizes
    // the workspace): 0 the expansion kernel, 1 the gradient rows with their planning
    // pass (device time between events, free of the host's share of the call)
    auto measure = [&](int phase) -> double {
        double best = 1e300;
        for (int r = 0; r <= repeats; r++) {
            compute_device(batch, grads, stream, &out);
            RA_HIP_CHECK(hipEventSynchronize(ev_stop_));
            float ms = 0.f;
            if (phase == 0) {
                RA_HIP_CHECK(hipEventElapsedTime(&ms, ev_exp_start_, ev_exp_stop_));
            } else {
                RA_HIP_CHECK(hipEventElapsedTime(&ms, ev_grad_start_, ev_main_stop_));
            }
            if (r > 0) best = std::min(best, static_cast<double>(ms));
        }
        return best;
    };
    const SoapTuning saved = tuning;
    SoapTuning best = saved;
    std::string report = "{\"lds_bytes\": " + std::to_string(device_lds_limit());
    auto fmt = [](double v) {
        char buf[32];
        std::snprintf(buf, sizeof(buf), "%.4f", v);
        return std::string(buf);
    };
    try {
        // 1. gradient rows: the built-in choice, every window shape of the tiled kernels
        //    within the whole LDS, the untiled kernel
        if (grads.positions) {
            // the fast form keeps an MFMA contraction: the rank-2 shapes and the shapes the
            // mixed-precision tiled kernel covers (n = 8: windows of up to 4 runs; n = 12,
            // 16: all)
            const int n = soap_.n_radial;
            const bool mixed = kind_ == Kind::PowerSpectrum && fast_form != 0 && precision == 32 &&
                               (n == 8 || n == 12 || n == 16);
            std::vector<int> candidates = {-1};
            for (int k = 0; k < static_cast<int>(soap_window_shapes().size()); k++) {
                if (mixed && n == 8 && soap_window_shapes()[k][0] > 4 && soap_window_shapes()[k][3] == 0) continue;
                // the spherical-expansion rows leave without store staging
                if (kind_ != Kind::PowerSpectrum && soap_window_shapes()[k][2] != 1) continue;
                candidates.push_back(k);
            }
            if (!mixed) candidates.push_back(-2);
            double best_ms = 1e300;
            report += ", \"gradient_window\": [";
            bool first = true;
            for (int c : candidates) {
                tuning = saved;
                tuning.gradient_window = c;
                std::string entry = "{\"choice\": " + std::to_string(c);
                if (c >= 0) {
                    const auto& w = soap_window_shapes()[c];
                    entry += ", \"shape\": [" + std::to_string(w[0]) + ", " + std::to_string(w[1]) + ", " +
                             std::to_string(w[2]) + "]";
                }
                try {
                    const double ms = measure(1);
                    entry += ", \"ms\": " + fmt(ms);
                    if (ms < best_ms) {
                        best_ms = ms;
                        best.gradient_window = c;
                    }
                } catch (const std::exception& e) {
                    // a shape the device cannot launch (e.g. beyond its LDS) is skipped
                    (void)hipGetLastError();
                    entry += ", \"error\": \"launch failed\"";
                }
                report += (first ? "" : ", ") + entry + "}";
                first = false;
            }
            report += "]";
        }
        // 2. neighbour entries per LDS chunk of the expansion kernel, and the matrix-core
        //    expansion kernel (negative counts) where it applies
        report += ", \"expansion_entries\": [";
        double best_ms = 1e300;
        bool first = true;
        std::vector<int> entry_counts = {4, 8, 16, 32, 64};
        // (the compute path falls back to the vector-ALU kernel where it does not apply)
        if (precision == 32 && expansion_mfma_supported(soap_.max_angular + 1, soap_.n_radial, 1)) {
            entry_counts.push_back(-8);
            entry_counts.push_back(-16);
        }
        for (int e : entry_counts) {
            tuning = best;
            tuning.expansion_entries = e;
            const double ms = measure(0);
            report += (first ? "" : ", ") + std::string("{\"entries\": ") + std::to_string(e) + ", \"ms\": " + fmt(ms) + "}";
            first = false;
            if (ms < best_ms) {
                best_ms = ms;
                best.expansion_entries = e;
            }
        }
        report += "]";
    } catch (...) {
        tuning = saved;
        throw;
    }
    tuning = best;
    report += ", \"chosen\": {\"gradient_window\": " + std::to_string(best.gradient_window) +
              ", \"expansion_entries\": " + std::to_string(best.expansion_entries) + "}}";
    return report;
}

void Calculator::positions_vjp(const HostBatch& batch, hipStream_t stream, const void* grad_values, void* out) {
    check_unbound(*this);
    if (kind_ != Kind::LodeSphericalExpansion && soap_.padded()) {
        throw Error::invalid_parameter("the positions vector-Jacobian product needs a tensor-product basis");
    }
    if (!grad_values || !out) throw Error::invalid_parameter("grad_values and out must not be NULL");
    if (kind_ == Kind::NeighborList || kind_ == Kind::ByPair) {
        throw Error::invalid_parameter("the positions VJP covers the SOAP and LODE calculators");
    }
    vjp_grad_ = grad_values;
    vjp_out_ = out;
    main_ms_sum_ = -1.0;
    try {
        if (kind_ == Kind::LodeSphericalExpansion) {
            if (precision == 32) {
                run_lode<float>(batch, Gradients{}, stream, nullptr);
            } else {
                run_lode<double>(batch, Gradients{}, stream, nullptr);
            }
        } else if (precision == 32) {
            run_soap<float>(batch, Gradients{}, stream, false, nullptr);
        } else {
            run_soap<double>(batch, Gradients{}, stream, false, nullptr);
        }
    } catch (...) {
        vjp_grad_ = nullptr;
        vjp_out_ = nullptr;
        throw;
    }
    vjp_grad_ = nullptr;
    vjp_out_ = nullptr;
}

void Calculator::last_timings(double* total_ms, double* main_ms, int64_t* main_bytes) {
    if (!ev_start_) throw Error::invalid_parameter("no device computation was run yet");
    RA_HIP_CHECK(hipEventSynchronize(ev_stop_));
    float t = 0.f, m = 0.f;
    RA_HIP_CHECK(hipEventElapsedTime(&t, ev_start_, ev_stop_));
    RA_HIP_CHECK(hipEventElapsedTime(&m, ev_main_start_, ev_main_stop_));
    if (total_ms) *total_ms = t;
    if (main_ms) *main_ms = main_ms_sum_ >= 0.0 ? main_ms_sum_ : m;
    if (main_bytes) *main_bytes = main_bytes_;
}

template mts_tensormap_t* Calculator::run_soap<float>(const HostBatch&, const Gradients&, hipStream_t, bool, featomic_amd_device_output_t*);
template mts_tensormap_t* Calculator::run_soap<double>(const HostBatch&, const Gradients&, hipStream_t, bool, featomic_amd_device_output_t*);
template const Calculator::DeviceTables<float>& Calculator::device_tables<float>();
template const Calculator::DeviceTables<double>& Calculator::device_tables<double>();

}  // namespace rascaline_amd
