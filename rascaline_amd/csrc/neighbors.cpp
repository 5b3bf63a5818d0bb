// Host orchestration of the GPU neighbour list.
#include <algorithm>
#include <cmath>
#include <vector>

#include "calculator.hpp"
#include "device.hpp"
#include "kernels.hpp"

namespace rascaline_amd {

namespace {

// featomic/src/types/matrix.rs:210-247
void inverse3(const double* m, double* r) {
    const double det = m[0] * (m[4] * m[8] - m[7] * m[5]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
                       m[2] * (m[3] * m[7] - m[4] * m[6]);
    if (!(std::fabs(det) > 1e-30)) throw Error::invalid_parameter("the cell matrix is not invertible");
    const double inv = 1.0 / det;
    r[0] = (m[4] * m[8] - m[7] * m[5]) * inv;
    r[1] = (m[2] * m[7] - m[1] * m[8]) * inv;
    r[2] = (m[1] * m[5] - m[2] * m[4]) * inv;
    r[3] = (m[5] * m[6] - m[3] * m[8]) * inv;
    r[4] = (m[0] * m[8] - m[2] * m[6]) * inv;
    r[5] = (m[3] * m[2] - m[0] * m[5]) * inv;
    r[6] = (m[3] * m[7] - m[6] * m[4]) * inv;
    r[7] = (m[6] * m[1] - m[0] * m[7]) * inv;
    r[8] = (m[0] * m[4] - m[3] * m[1]) * inv;
}

// featomic/src/systems/cell.rs:154-169
void face_distances(const double* m, double* out) {
    const double* a = m;
    const double* b = m + 3;
    const double* c = m + 6;
    auto cross = [](const double* u, const double* v, double* w) {
        w[0] = u[1] * v[2] - u[2] * v[1];
        w[1] = u[2] * v[0] - u[0] * v[2];
        w[2] = u[0] * v[1] - u[1] * v[0];
    };
    auto dot = [](const double* u, const double* v) { return u[0] * v[0] + u[1] * v[1] + u[2] * v[2]; };
    double na[3], nb[3], nc[3];
    cross(b, c, na);
    cross(c, a, nb);
    cross(a, b, nc);
    const double la = std::sqrt(dot(na, na)), lb = std::sqrt(dot(nb, nb)), lc = std::sqrt(dot(nc, nc));
    for (int k = 0; k < 3; k++) {
        na[k] /= la;
        nb[k] /= lb;
        nc[k] /= lc;
    }
    out[0] = std::fabs(dot(na, a));
    out[1] = std::fabs(dot(nb, b));
    out[2] = std::fabs(dot(nc, c));
}

__global__ void int_to_i64(const int* in, long long* out, long long n) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) out[k] = in[k];
}

}  // namespace

DeviceNeighbors build_neighbors(Workspace& ws, const HostBatch& batch, double cutoff, const double* d_positions,
                                const int* d_atom_system, hipStream_t stream, const PassContext& pass) {
    DeviceNeighbors nl;
    const int n_atoms = static_cast<int>(batch.n_atoms());
    const size_t n_systems = batch.n_systems();
    nl.offsets = ws.get<long long>("nl_offsets", n_atoms + 1);

    if (batch.has_pairs) {
        if (pass.mode != PassMode::Eager) throw Error::invalid_parameter("batch graphs need the GPU neighbour list");
        // caller-provided half lists: expand to the sorted full list on the host
        struct Entry {
            int j;
            int s[3];
            double v[3];
            double d;
        };
        std::vector<std::vector<Entry>> per_center(n_atoms);
        for (size_t sys = 0; sys < n_systems; sys++) {
            const int64_t base = batch.offsets[sys];
            for (const auto& p : batch.pairs[sys]) {
                const int a = static_cast<int>(base + p.first), b = static_cast<int>(base + p.second);
                Entry fwd{b, {p.cell_shift_indices[0], p.cell_shift_indices[1], p.cell_shift_indices[2]},
                          {p.vector[0], p.vector[1], p.vector[2]}, p.distance};
                Entry bwd{a, {-p.cell_shift_indices[0], -p.cell_shift_indices[1], -p.cell_shift_indices[2]},
                          {-p.vector[0], -p.vector[1], -p.vector[2]}, p.distance};
                per_center[a].push_back(fwd);
                per_center[b].push_back(bwd);
            }
        }
        std::vector<long long> offsets(n_atoms + 1, 0);
        for (int i = 0; i < n_atoms; i++) {
            auto& v = per_center[i];
            std::sort(v.begin(), v.end(), [](const Entry& x, const Entry& y) {
                if (x.j != y.j) return x.j < y.j;
                for (int k = 0; k < 3; k++) {
                    if (x.s[k] != y.s[k]) return x.s[k] < y.s[k];
                }
                return false;
            });
            offsets[i + 1] = offsets[i] + static_cast<long long>(v.size());
        }
        nl.n_entries = offsets[n_atoms];
        std::vector<int> ej, es;
        std::vector<double> ev, ed;
        ej.reserve(nl.n_entries);
        for (int i = 0; i < n_atoms; i++) {
            for (auto& e : per_center[i]) {
                ej.push_back(e.j);
                es.insert(es.end(), e.s, e.s + 3);
                ev.insert(ev.end(), e.v, e.v + 3);
                ed.push_back(e.d);
            }
        }
        const size_t ne = static_cast<size_t>(nl.n_entries);
        nl.entry_j = ws.get<int>("nl_entry_j", ne);
        nl.entry_shift = ws.get<int>("nl_entry_shift", 3 * ne);
        nl.entry_vec = ws.get<double>("nl_entry_vec", 3 * ne);
        nl.entry_dist = ws.get<double>("nl_entry_dist", ne);
        RA_HIP_CHECK(hipMemcpyAsync(nl.offsets, offsets.data(), sizeof(long long) * (n_atoms + 1), hipMemcpyHostToDevice, stream));
        if (ne) {
            RA_HIP_CHECK(hipMemcpyAsync(nl.entry_j, ej.data(), sizeof(int) * ne, hipMemcpyHostToDevice, stream));
            RA_HIP_CHECK(hipMemcpyAsync(nl.entry_shift, es.data(), sizeof(int) * 3 * ne, hipMemcpyHostToDevice, stream));
            RA_HIP_CHECK(hipMemcpyAsync(nl.entry_vec, ev.data(), sizeof(double) * 3 * ne, hipMemcpyHostToDevice, stream));
            RA_HIP_CHECK(hipMemcpyAsync(nl.entry_dist, ed.data(), sizeof(double) * ne, hipMemcpyHostToDevice, stream));
        }
        RA_HIP_CHECK(hipStreamSynchronize(stream));
        return nl;
    }

    // cell-list parameters per system, featomic/src/systems/neighbors.rs:101-163
    std::vector<NLSystem> params(n_systems);
    long long n_bins = 0;
    for (size_t s = 0; s < n_systems; s++) {
        NLSystem& p = params[s];
        const double* cell = &batch.cells[9 * s];
        bool periodic = false;
        for (int k = 0; k < 9; k++) {
            p.cell[k] = cell[k];
            periodic |= cell[k] != 0.0;
            p.inv[k] = 0.0;
        }
        double faces[3] = {1.0, 1.0, 1.0};
        if (periodic) {
            inverse3(cell, p.inv);
            face_distances(cell, faces);
        }
        double nc[3];
        for (int k = 0; k < 3; k++) nc[k] = std::max(1.0, std::trunc(faces[k] / cutoff));
        if (nc[0] * nc[1] * nc[2] > 1e5) {
            const double rxy = nc[0] / nc[1], ryz = nc[1] / nc[2];
            nc[2] = std::trunc(std::cbrt(1e5 / (rxy * ryz * ryz)));
            nc[1] = std::trunc(ryz * nc[2]);
            nc[0] = std::trunc(rxy * nc[1]);
        }
        for (int k = 0; k < 3; k++) {
            p.n_cells[k] = static_cast<int>(nc[k]);
            p.n_search[k] = static_cast<int>(std::ceil(cutoff * nc[k] / faces[k]));
            if (p.n_search[k] < 1) p.n_search[k] = 1;
            if (p.n_cells[k] == 1 && !periodic) p.n_search[k] = 0;
        }
        p.periodic = periodic ? 1 : 0;
        p.n_atoms = static_cast<int>(batch.offsets[s + 1] - batch.offsets[s]);
        p.bin_offset = n_bins;
        p.atom_offset = batch.offsets[s];
        n_bins += static_cast<long long>(p.n_cells[0]) * p.n_cells[1] * p.n_cells[2];
    }

    // cell-list parameters: uploaded (Eager), kept in the batch record (Record), or the
    // record's copy (Replay)
    NLSystem* d_systems = nullptr;
    if (pass.mode == PassMode::Eager) {
        d_systems = ws.get<NLSystem>("nl_systems", n_systems);
    } else {
        d_systems = pass.record->inputs.get<NLSystem>("nl_systems", n_systems);
        if (pass.mode == PassMode::Record) pass.record->input_counts["nl_systems"] = n_systems;
        if (pass.mode == PassMode::Replay && pass.record->input_counts["nl_systems"] != n_systems) {
            throw Error::internal("replayed batch differs from its record (systems)");
        }
    }
    if (pass.mode != PassMode::Replay) {
        RA_HIP_CHECK(hipMemcpyAsync(d_systems, params.data(), sizeof(NLSystem) * n_systems, hipMemcpyHostToDevice, stream));
    }
    long long* atom_bin = ws.get<long long>("nl_atom_bin", n_atoms);
    int* atom_wrap = ws.get<int>("nl_atom_wrap", 3 * static_cast<size_t>(n_atoms));
    int* bin_count = ws.get<int>("nl_bin_count", n_bins);
    int* bin_fill = ws.get<int>("nl_bin_fill", n_bins);
    long long* bin_count64 = ws.get<long long>("nl_bin_count64", n_bins);
    long long* bin_start = ws.get<long long>("nl_bin_start", n_bins + 1);
    int* bin_atoms = ws.get<int>("nl_bin_atoms", n_atoms);
    long long* counts = ws.get<long long>("nl_counts", n_atoms);
    int* error_flag = ws.get<int>("nl_error", 1);
    fill_zero(bin_count, sizeof(int) * n_bins, stream);
    fill_zero(bin_fill, sizeof(int) * n_bins, stream);
    fill_zero(error_flag, sizeof(int), stream);

    launch_nl_assign_bins(stream, n_atoms, d_positions, d_atom_system, d_systems, atom_bin, atom_wrap, bin_count);
    if (n_bins) int_to_i64<<<(unsigned)((n_bins + 255) / 256), 256, 0, stream>>>(bin_count, bin_count64, n_bins);
    exclusive_scan(stream, bin_count64, bin_start, n_bins, ws);
    launch_nl_fill_bins(stream, n_atoms, atom_bin, bin_start, bin_fill, bin_atoms);
    const double cutoff2 = cutoff * cutoff;
    unsigned long long* stash = ws.get<unsigned long long>("nl_stash", std::max<size_t>(nl_stash_keys(n_atoms), 1));
    launch_nl_candidates(stream, false, n_atoms, d_positions, d_atom_system, d_systems, atom_bin, atom_wrap, bin_start,
                         bin_count, bin_atoms, cutoff2, counts, nullptr, nullptr, error_flag, stash);
    exclusive_scan(stream, counts, nl.offsets, n_atoms, ws);
    long long n_entries = 0;
    if (pass.mode == PassMode::Replay) {
        // the recorded count; a different one is reported by the graph's check
        n_entries = pass.record->n_entries;
        check_equal(stream, nl.offsets + n_atoms, n_entries, REPLAY_NEIGHBORS_CHANGED, pass.replay_flags);
    } else {
        RA_HIP_CHECK(hipMemcpyAsync(&n_entries, nl.offsets + n_atoms, sizeof(long long), hipMemcpyDeviceToHost, stream));
        RA_HIP_CHECK(hipStreamSynchronize(stream));
        if (pass.mode == PassMode::Record) pass.record->n_entries = n_entries;
    }
    nl.n_entries = n_entries;
    const size_t ne = static_cast<size_t>(n_entries);
    unsigned long long* keys = ws.get<unsigned long long>("nl_keys", ne);
    launch_nl_candidates(stream, true, n_atoms, d_positions, d_atom_system, d_systems, atom_bin, atom_wrap, bin_start,
                         bin_count, bin_atoms, cutoff2, counts, nl.offsets, keys, error_flag, stash);
    launch_nl_sort_segments(stream, n_atoms, nl.offsets, keys, error_flag);
    nl.entry_j = ws.get<int>("nl_entry_j", ne);
    nl.entry_shift = ws.get<int>("nl_entry_shift", 3 * ne);
    nl.entry_vec = ws.get<double>("nl_entry_vec", 3 * ne);
    nl.entry_dist = ws.get<double>("nl_entry_dist", ne);
    int* entry_center = ws.get<int>("nl_entry_center", ne);
    launch_nl_finalize(stream, n_entries, n_atoms, nl.offsets, keys, d_positions, d_atom_system, d_systems, entry_center,
                       nl.entry_j, nl.entry_shift, nl.entry_vec, nl.entry_dist);
    RA_HIP_CHECK(hipGetLastError());
    if (pass.mode == PassMode::Replay) {
        or_flag(stream, error_flag, REPLAY_NEIGHBOR_ERROR, pass.replay_flags);
        return nl;
    }
    int flag = 0;
    RA_HIP_CHECK(hipMemcpyAsync(&flag, error_flag, sizeof(int), hipMemcpyDeviceToHost, stream));
    RA_HIP_CHECK(hipStreamSynchronize(stream));
    if (flag & 1) throw Error::invalid_parameter("cell shifts beyond +/-511 cells are not supported");
    if (flag & 2) throw Error::invalid_parameter("an atom has more than 2048 neighbours within the cutoff");
    return nl;
}

}  // namespace rascaline_amd
