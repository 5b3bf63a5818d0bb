// spherical_expansion_by_pair: the spherical expansion split per directed pair
// (featomic/src/calculators/soap/spherical_expansion_pair.rs:547-840). The GPU
// builds the full neighbour list and writes every sample's values and gradient rows
// (by_pair_kernel); the host lays out the samples in the reference's order and
// assembles the TensorMap.

#include <algorithm>
#include <array>
#include <map>

#include "calculator.hpp"
#include "kernels.hpp"

namespace rascaline_amd {

namespace {

struct PairSample {
    std::array<int32_t, 6> labels;  // system, first_atom, second_atom, cell_shift_a, b, c
    long long entry;                // full-list entry of (first; second, shift), -1 for a self pair
};

bool is_self_pair(const std::array<int32_t, 6>& l) { return l[1] == l[2] && l[3] == 0 && l[4] == 0 && l[5] == 0; }

}  // namespace

template <typename T>
mts_tensormap_t* Calculator::run_by_pair(const HostBatch& batch, const Gradients& grads) {
    hipStream_t stream = nullptr;
    const int n_atoms = static_cast<int>(batch.n_atoms());
    const int L = soap_.L(), n = soap_.n_radial;
    auto tm = std::make_unique<mts_tensormap_t>();
    tm->keys = LabelsData({"o3_lambda", "o3_sigma", "first_atom_type", "second_atom_type"}, {});
    // timings read by last_timings: the whole call, and by_pair_kernel
    for (hipEvent_t* ev : {&ev_start_, &ev_stop_, &ev_main_start_, &ev_main_stop_, &ev_exp_start_, &ev_exp_stop_,
                           &ev_grad_start_}) {
        if (!*ev) RA_HIP_CHECK(hipEventCreate(ev));
    }
    main_ms_sum_ = -1.0;
    main_bytes_ = 0;
    RA_HIP_CHECK(hipEventRecord(ev_start_, stream));
    if (n_atoms == 0) {
        RA_HIP_CHECK(hipEventRecord(ev_main_start_, stream));
        RA_HIP_CHECK(hipEventRecord(ev_main_stop_, stream));
        RA_HIP_CHECK(hipEventRecord(ev_stop_, stream));
        return tm.release();
    }
    const Species sp = species_of(batch);
    const int S = static_cast<int>(sp.types.size());

    // ---- full neighbour list on the device
    std::vector<int32_t> atom_system(n_atoms);
    for (size_t s = 0; s < batch.n_systems(); s++) {
        for (int64_t a = batch.offsets[s]; a < batch.offsets[s + 1]; a++) atom_system[a] = static_cast<int32_t>(s);
    }
    double* d_positions = ws_.get<double>("positions", 3 * static_cast<size_t>(n_atoms));
    upload(d_positions, batch.positions, stream);
    int* d_atom_system = ws_.get<int>("atom_system", n_atoms);
    int* d_species = ws_.get<int>("species", n_atoms);
    upload(d_atom_system, atom_system, stream);
    upload(d_species, sp.index, stream);
    const DeviceNeighbors nl = build_neighbors(ws_, batch, soap_.cutoff, d_positions, d_atom_system, stream);
    std::vector<long long> offsets;
    std::vector<int> entry_j, entry_shift;
    download(offsets, nl.offsets, n_atoms + 1, stream);
    download(entry_j, nl.entry_j, nl.n_entries, stream);
    download(entry_shift, nl.entry_shift, 3 * nl.n_entries, stream);
    RA_HIP_CHECK(hipStreamSynchronize(stream));

    // entries of a center are sorted by (neighbour, shift a, b, c)
    auto find_entry = [&](int64_t i, int64_t j, const int32_t* shift) -> long long {
        long long lo = offsets[i], hi = offsets[i + 1];
        auto less = [&](long long e) {
            if (entry_j[e] != j) return entry_j[e] < j;
            for (int k = 0; k < 3; k++) {
                if (entry_shift[3 * e + k] != shift[k]) return entry_shift[3 * e + k] < shift[k];
            }
            return false;
        };
        while (lo < hi) {
            const long long mid = (lo + hi) / 2;
            if (less(mid)) lo = mid + 1; else hi = mid;
        }
        if (lo == offsets[i + 1] || entry_j[lo] != j || entry_shift[3 * lo] != shift[0] || entry_shift[3 * lo + 1] != shift[1] ||
            entry_shift[3 * lo + 2] != shift[2]) {
            throw Error::internal("a pair of the neighbour list has no entry in the full list");
        }
        return lo;
    };

    // ---- keys: the full-list type pairs with self pairs (neighbor_list.rs:388-414),
    // each with every l (spherical_expansion_pair.rs:560-583)
    std::map<std::pair<int32_t, int32_t>, std::vector<PairSample>> groups;
    for (int32_t t : sp.types) groups[{t, t}];
    // ---- samples: per system its half-list pairs in list order, both directions
    // (neighbor_list.rs:416-489), then its self pairs
    for (size_t s = 0; s < batch.n_systems(); s++) {
        const int64_t base = batch.offsets[s];
        const int64_t count = batch.offsets[s + 1] - base;
        std::vector<featomic_pair_t> derived;
        const std::vector<featomic_pair_t>* pairs = nullptr;
        if (batch.has_pairs) {
            pairs = &batch.pairs[s];
        } else {
            // half list from the full one: j > i, or j == i with the shift in the
            // positive half-space (systems/neighbors.rs:253-286)
            for (int64_t i = 0; i < count; i++) {
                for (long long e = offsets[base + i]; e < offsets[base + i + 1]; e++) {
                    const int64_t j = entry_j[e] - base;
                    const int* t = &entry_shift[3 * e];
                    bool keep = j > i;
                    if (j == i) {
                        const int sum = t[0] + t[1] + t[2];
                        keep = !(sum < 0 || (sum == 0 && (t[2] < 0 || (t[2] == 0 && t[1] < 0))));
                    }
                    if (!keep) continue;
                    featomic_pair_t p{};
                    p.first = static_cast<uintptr_t>(i);
                    p.second = static_cast<uintptr_t>(j);
                    for (int k = 0; k < 3; k++) p.cell_shift_indices[k] = t[k];
                    derived.push_back(p);
                }
            }
            pairs = &derived;
        }
        for (const featomic_pair_t& p : *pairs) {
            const int64_t i = static_cast<int64_t>(p.first), j = static_cast<int64_t>(p.second);
            const int32_t* sh = p.cell_shift_indices;
            const int32_t minus[3] = {-sh[0], -sh[1], -sh[2]};
            const int32_t a = batch.types[base + i], b = batch.types[base + j];
            PairSample direct{{static_cast<int32_t>(s), static_cast<int32_t>(i), static_cast<int32_t>(j), sh[0], sh[1], sh[2]},
                              find_entry(base + i, base + j, sh)};
            PairSample reverse{{static_cast<int32_t>(s), static_cast<int32_t>(j), static_cast<int32_t>(i), minus[0], minus[1], minus[2]},
                               find_entry(base + j, base + i, minus)};
            groups[{a, b}].push_back(direct);
            groups[{b, a}].push_back(reverse);
        }
        for (int64_t i = 0; i < count; i++) {
            const int32_t t = batch.types[base + i];
            groups[{t, t}].push_back({{static_cast<int32_t>(s), static_cast<int32_t>(i), static_cast<int32_t>(i), 0, 0, 0}, -1});
        }
    }

    // ---- rows and block offsets; keys (type pair)-major, then l
    const int NTP = static_cast<int>(groups.size());
    std::vector<ByPairSample> dev_samples;
    std::vector<long long> value_base(static_cast<size_t>(NTP) * L), pos_base(value_base.size()),
        strain_base(value_base.size()), cell_base(value_base.size());
    std::vector<long long> n_all(NTP), n_pairs(NTP), n_rows(NTP);
    long long value_total = 0, pos_total = 0, cs_total = 0;
    int tp = 0;
    for (auto& kv : groups) {
        int row1 = 0, grad_row = 0;
        for (size_t k = 0; k < kv.second.size(); k++) {
            const PairSample& ps = kv.second[k];
            ByPairSample d{};
            d.entry = ps.entry;
            d.type_pair = tp;
            d.row0 = static_cast<int>(k);
            const bool self = is_self_pair(ps.labels);
            d.row1 = self ? -1 : row1++;
            d.grad_row = grad_row;
            d.grad_rows = self ? 0 : (ps.labels[1] == ps.labels[2] ? 1 : 2);
            grad_row += d.grad_rows;
            dev_samples.push_back(d);
        }
        n_all[tp] = static_cast<long long>(kv.second.size());
        n_pairs[tp] = row1;
        n_rows[tp] = grad_row;
        for (int l = 0; l < L; l++) {
            const long long w = static_cast<long long>(2 * l + 1) * n;
            const long long ns = l == 0 ? n_all[tp] : n_pairs[tp];
            const size_t b = static_cast<size_t>(tp) * L + l;
            value_base[b] = value_total;
            value_total += ns * w;
            pos_base[b] = pos_total;
            pos_total += grads.positions ? n_rows[tp] * 3 * w : 0;
            strain_base[b] = cell_base[b] = cs_total;
            cs_total += (grads.strain || grads.cell) ? ns * 9 * w : 0;
        }
        tp++;
    }

    // ---- GPU: per-entry data and every sample's rows
    SoapLaunch<T> lp;
    soap_launch<T>(lp, nl, d_species, n_atoms, S, stream);
    {
        const long long per_entry = static_cast<long long>(soap_chunk_bytes<T>(L, n, 1));
        const long long budget = std::min<long long>(device_lds_limit(), 49152) - 1024;
        lp.chunk = static_cast<int>(std::max<long long>(1, std::min<long long>(64, budget / per_entry)));
    }
    const long long n_samples = static_cast<long long>(dev_samples.size());
    ByPairSample* d_samples = ws_.get<ByPairSample>("by_pair_samples", std::max<long long>(n_samples, 1));
    long long* d_bases = ws_.get<long long>("by_pair_bases", 4 * value_base.size());
    upload(d_samples, dev_samples, stream);
    std::vector<long long> bases;
    for (auto* v : {&value_base, &pos_base, &strain_base, &cell_base}) bases.insert(bases.end(), v->begin(), v->end());
    upload(d_bases, bases, stream);
    T* d_values = ws_.get<T>("by_pair_values", std::max<long long>(value_total, 1));
    T* d_pos = grads.positions ? ws_.get<T>("by_pair_positions", std::max<long long>(pos_total, 1)) : nullptr;
    T* d_strain = grads.strain ? ws_.get<T>("by_pair_strain", std::max<long long>(cs_total, 1)) : nullptr;
    T* d_cell = grads.cell ? ws_.get<T>("by_pair_cell", std::max<long long>(cs_total, 1)) : nullptr;
    const size_t nb = value_base.size();
    RA_HIP_CHECK(hipEventRecord(ev_main_start_, stream));
    launch_by_pair<T>(lp, n_samples, d_samples, d_bases, d_bases + nb, d_bases + 2 * nb, d_bases + 3 * nb, d_values,
                      d_pos, d_strain, d_cell);
    RA_HIP_CHECK(hipGetLastError());
    RA_HIP_CHECK(hipEventRecord(ev_main_stop_, stream));
    // algorithmic bytes of by_pair_kernel: every output element written once, and per
    // sample its descriptor (32 B) and neighbour entry (vector, distance, index and
    // shift: 48 B) read once
    {
        const long long written = value_total + (d_pos ? pos_total : 0) + (d_strain ? cs_total : 0) + (d_cell ? cs_total : 0);
        main_bytes_ = written * static_cast<long long>(sizeof(T)) + n_samples * 80;
    }
    std::vector<T> h_values, h_pos, h_strain, h_cell;
    download(h_values, d_values, value_total, stream);
    if (d_pos) download(h_pos, d_pos, pos_total, stream);
    if (d_strain) download(h_strain, d_strain, cs_total, stream);
    if (d_cell) download(h_cell, d_cell, cs_total, stream);
    RA_HIP_CHECK(hipEventRecord(ev_stop_, stream));
    RA_HIP_CHECK(hipStreamSynchronize(stream));

    // ---- TensorMap
    std::vector<int32_t> keys;
    std::vector<int32_t> prop_values(n);
    for (int k = 0; k < n; k++) prop_values[k] = k;
    const LabelsData properties({"n"}, prop_values);
    const LabelsData xyz = LabelsData::range("xyz", 3);
    tp = 0;
    for (auto& kv : groups) {
        const std::vector<PairSample>& samples = kv.second;
        for (int l = 0; l < L; l++) {
            keys.insert(keys.end(), {l, 1, kv.first.first, kv.first.second});
            const long long w = static_cast<long long>(2 * l + 1) * n;
            const size_t b = static_cast<size_t>(tp) * L + l;
            std::vector<int32_t> sl;
            std::vector<int32_t> gl;
            for (const PairSample& ps : samples) {
                const bool self = is_self_pair(ps.labels);
                if (l > 0 && self) continue;
                const int32_t row = static_cast<int32_t>(sl.size() / 6);
                sl.insert(sl.end(), ps.labels.begin(), ps.labels.end());
                if (self) continue;
                gl.insert(gl.end(), {row, ps.labels[0], ps.labels[1]});
                if (ps.labels[1] != ps.labels[2]) gl.insert(gl.end(), {row, ps.labels[0], ps.labels[2]});
            }
            const uintptr_t ns = sl.size() / 6;
            auto block = make_block(LabelsData({"system", "first_atom", "second_atom", "cell_shift_a", "cell_shift_b", "cell_shift_c"}, sl),
                                    {mu_labels(l)}, properties, {ns, static_cast<uintptr_t>(2 * l + 1), static_cast<uintptr_t>(n)});
            for (long long x = 0; x < static_cast<long long>(ns) * w; x++) {
                block->array.data[x] = static_cast<double>(h_values[value_base[b] + x]);
            }
            if (grads.positions) {
                const uintptr_t nr = gl.size() / 3;
                auto g = make_block(LabelsData({"sample", "system", "atom"}, gl), {xyz, mu_labels(l)}, properties,
                                    {nr, 3, static_cast<uintptr_t>(2 * l + 1), static_cast<uintptr_t>(n)});
                for (long long x = 0; x < static_cast<long long>(nr) * 3 * w; x++) {
                    g->array.data[x] = static_cast<double>(h_pos[pos_base[b] + x]);
                }
                block->add_gradient("positions", std::move(g));
            }
            for (int which = 0; which < 2; which++) {
                const bool want = which == 0 ? grads.cell : grads.strain;
                if (!want) continue;
                const std::vector<T>& src = which == 0 ? h_cell : h_strain;
                std::vector<LabelsData> gcomps;
                if (which == 0) {
                    gcomps = {xyz, LabelsData::range("abc", 3), mu_labels(l)};
                } else {
                    gcomps = {LabelsData::range("xyz_2", 3), LabelsData::range("xyz_1", 3), mu_labels(l)};
                }
                auto g = make_block(LabelsData::range("sample", static_cast<int>(ns)), gcomps, properties,
                                    {ns, 3, 3, static_cast<uintptr_t>(2 * l + 1), static_cast<uintptr_t>(n)});
                const long long base = which == 0 ? cell_base[b] : strain_base[b];
                for (long long x = 0; x < static_cast<long long>(ns) * 9 * w; x++) {
                    g->array.data[x] = static_cast<double>(src[base + x]);
                }
                block->add_gradient(which == 0 ? "cell" : "strain", std::move(g));
            }
            tm->blocks.push_back(std::move(block));
        }
        tp++;
    }
    tm->keys.values = keys;
    return tm.release();
}

template mts_tensormap_t* Calculator::run_by_pair<float>(const HostBatch&, const Gradients&);
template mts_tensormap_t* Calculator::run_by_pair<double>(const HostBatch&, const Gradients&);

}  // namespace rascaline_amd
