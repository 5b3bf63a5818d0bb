// neighbor_list calculator: the pair vectors of the neighbour list as a TensorMap
// (featomic/src/calculators/neighbor_list.rs). The half list comes from the GPU
// cell list (or from the caller's systems, use_native_system = false); the
// relabelling into per-type-pair blocks is host work over the pairs.

#include <map>
#include <utility>

#include "calculator.hpp"

namespace rascaline_amd {

namespace {

// one sample of a block: the pair's labels and its vector
struct PairSample {
    int32_t labels[6];  // system, first_atom, second_atom, cell_shift_a, b, c
    double vector[3];
};

void add_sample(std::vector<PairSample>& block, int system, uintptr_t first, uintptr_t second, const int32_t* shift,
                int sign, const double* vector) {
    PairSample s;
    s.labels[0] = system;
    s.labels[1] = static_cast<int32_t>(first);
    s.labels[2] = static_cast<int32_t>(second);
    for (int k = 0; k < 3; k++) {
        s.labels[3 + k] = sign * shift[k];
        s.vector[k] = sign * vector[k];
    }
    block.push_back(s);
}

}  // namespace

mts_tensormap_t* Calculator::run_neighbor_list(const HostBatch& batch, const Gradients& grads) {
    if (grads.strain || grads.cell) {
        throw Error::invalid_parameter(std::string("the neighbors list calculator does not support gradients with respect to ") +
                                       (grads.cell ? "cell" : "strain"));
    }
    std::vector<std::vector<featomic_pair_t>> gpu_pairs;
    if (!batch.has_pairs) gpu_pairs = gpu_half_pairs(batch, nl_.cutoff, ws_);
    const auto& pairs = batch.has_pairs ? batch.pairs : gpu_pairs;
    const size_t n_systems = batch.n_systems();

    // keys: the (first, second) atom types of every pair, sorted and unique
    // (HalfNeighborList::keys, neighbor_list.rs:172-200; FullNeighborList::keys,
    // neighbor_list.rs:388-414); with self pairs, (t, t) for every type present
    std::map<std::pair<int32_t, int32_t>, std::vector<PairSample>> blocks;
    for (size_t s = 0; s < n_systems; s++) {
        const int32_t* types = batch.types.data() + batch.offsets[s];
        for (const featomic_pair_t& p : pairs[s]) {
            const int32_t ti = types[p.first], tj = types[p.second];
            if (nl_.full_neighbor_list) {
                blocks[{ti, tj}];
                blocks[{tj, ti}];
            } else {
                blocks[{std::min(ti, tj), std::max(ti, tj)}];
            }
        }
        if (nl_.self_pairs) {
            for (int64_t a = 0; a < batch.offsets[s + 1] - batch.offsets[s]; a++) blocks[{types[a], types[a]}];
        }
    }

    // samples in the reference's order: per system, its pairs in list order, then
    // its self pairs (HalfNeighborList::samples, neighbor_list.rs:202-270;
    // FullNeighborList::samples, neighbor_list.rs:416-489). Values: the pair vector,
    // negated where the pair is stored as second -> first (neighbor_list.rs:272-365,
    // 493-620).
    const int32_t zero_shift[3] = {0, 0, 0};
    const double zero_vector[3] = {0.0, 0.0, 0.0};
    for (size_t s = 0; s < n_systems; s++) {
        const int system = static_cast<int>(s);
        const int32_t* types = batch.types.data() + batch.offsets[s];
        for (const featomic_pair_t& p : pairs[s]) {
            const int32_t ti = types[p.first], tj = types[p.second];
            if (nl_.full_neighbor_list) {
                add_sample(blocks[{ti, tj}], system, p.first, p.second, p.cell_shift_indices, 1, p.vector);
                add_sample(blocks[{tj, ti}], system, p.second, p.first, p.cell_shift_indices, -1, p.vector);
            } else if (ti <= tj) {
                add_sample(blocks[{ti, tj}], system, p.first, p.second, p.cell_shift_indices, 1, p.vector);
            } else {
                add_sample(blocks[{tj, ti}], system, p.second, p.first, p.cell_shift_indices, -1, p.vector);
            }
        }
        if (nl_.self_pairs) {
            for (int64_t a = 0; a < batch.offsets[s + 1] - batch.offsets[s]; a++) {
                add_sample(blocks[{types[a], types[a]}], system, a, a, zero_shift, 1, zero_vector);
            }
        }
    }

    auto tm = std::make_unique<mts_tensormap_t>();
    std::vector<int32_t> keys;
    const LabelsData pair_xyz = LabelsData::range("pair_xyz", 3);
    const LabelsData xyz = LabelsData::range("xyz", 3);
    const LabelsData properties({"distance"}, {1});
    for (auto& kv : blocks) {
        keys.push_back(kv.first.first);
        keys.push_back(kv.first.second);
        const std::vector<PairSample>& samples = kv.second;
        const size_t ns = samples.size();
        std::vector<int32_t> sl;
        sl.reserve(6 * ns);
        for (const PairSample& p : samples) sl.insert(sl.end(), p.labels, p.labels + 6);
        auto block = make_block(LabelsData({"system", "first_atom", "second_atom", "cell_shift_a", "cell_shift_b", "cell_shift_c"}, sl),
                                {pair_xyz}, properties, {ns, 3, 1});
        for (size_t r = 0; r < ns; r++) {
            for (int k = 0; k < 3; k++) block->array.data[3 * r + k] = samples[r].vector[k];
        }
        if (grads.positions) {
            // d(r_second - r_first + shift)/d r_first = -I, d/d r_second = +I; true self
            // pairs have no gradient rows (positions_gradient_samples,
            // neighbor_list.rs:108-127). A pair between an atom and its own periodic
            // image has one row, with the derivative 0: the reference lists that atom
            // twice in the same sample, which metatensor's labels do not allow.
            std::vector<int32_t> gl;
            std::vector<double> gv;
            for (size_t r = 0; r < ns; r++) {
                const int32_t* l = samples[r].labels;
                const bool same_atom = l[1] == l[2];
                if (same_atom && l[3] == 0 && l[4] == 0 && l[5] == 0) continue;
                const int n_rows = same_atom ? 1 : 2;
                for (int k = 0; k < n_rows; k++) {
                    gl.insert(gl.end(), {static_cast<int32_t>(r), l[0], l[1 + k]});
                    const double sign = same_atom ? 0.0 : (k == 0 ? -1.0 : 1.0);
                    for (int x = 0; x < 3; x++) {
                        for (int c = 0; c < 3; c++) gv.push_back(x == c ? sign : 0.0);
                    }
                }
            }
            const size_t nr = gl.size() / 3;
            auto g = make_block(LabelsData({"sample", "system", "atom"}, gl), {xyz, pair_xyz}, properties, {nr, 3, 3, 1});
            g->array.data = std::move(gv);
            block->add_gradient("positions", std::move(g));
        }
        tm->blocks.push_back(std::move(block));
    }
    tm->keys = LabelsData({"first_atom_type", "second_atom_type"}, keys);
    return tm.release();
}

}  // namespace rascaline_amd
