// Neighbour list and per-center metadata kernels (gfx950).
//
// The neighbour list is built as a *full* list: every center i gets all (j, S)
// with |r_j - r_i + S.H| < rc, (j, S) != (i, 0), sorted by (j, S). The
// reference's half list (featomic/src/systems/neighbors.rs:349-403) is the
// subset j > i, plus j == i with S in the positive half-space, and the pair
// vectors are computed with the same floating-point expression
// (p_j - p_i) + (H_0k s_0 + H_1k s_1 + H_2k s_2), so they are bit-identical.
// Compile with -ffp-contract=off for this file.

#include <hip/hip_runtime.h>

#include <cstdint>

#include "device.hpp"
#include "kernels.hpp"

namespace rascaline_amd {

namespace {

__device__ inline void divmod(int a, int b, int* q, int* r) {
    int quotient = a / b, remainder = a % b;
    if (remainder < 0) {
        remainder += b;
        quotient -= 1;
    }
    *q = quotient;
    *r = remainder;
}

__device__ inline int floor_to_int(double x) {
    double f = floor(x);
    // f64 -> i32 casts saturate in the reference
    if (f > 2147483647.0) return 2147483647;
    if (f < -2147483648.0) return (-2147483647 - 1);
    return (int)f;
}

constexpr int SHIFT_BIAS = 512;

__device__ inline unsigned long long pack_key(int j, int s0, int s1, int s2) {
    return ((unsigned long long)(unsigned)j << 32) | ((unsigned long long)(unsigned)(s0 + SHIFT_BIAS) << 20) |
           ((unsigned long long)(unsigned)(s1 + SHIFT_BIAS) << 10) | (unsigned long long)(unsigned)(s2 + SHIFT_BIAS);
}

__device__ inline void unpack_key(unsigned long long key, int* j, int* s) {
    *j = (int)(key >> 32);
    s[0] = (int)((key >> 20) & 1023u) - SHIFT_BIAS;
    s[1] = (int)((key >> 10) & 1023u) - SHIFT_BIAS;
    s[2] = (int)(key & 1023u) - SHIFT_BIAS;
}

// pair vector with the reference's operation order (neighbors.rs:52-57, 365-367)
__device__ inline void pair_vector(const double* pi, const double* pj, const double* h, const int* s, double* v) {
    double vx = pj[0] - pi[0];
    double vy = pj[1] - pi[1];
    double vz = pj[2] - pi[2];
    const double sx = h[0] * (double)s[0] + h[3] * (double)s[1] + h[6] * (double)s[2];
    const double sy = h[1] * (double)s[0] + h[4] * (double)s[1] + h[7] * (double)s[2];
    const double sz = h[2] * (double)s[0] + h[5] * (double)s[1] + h[8] * (double)s[2];
    v[0] = vx + sx;
    v[1] = vy + sy;
    v[2] = vz + sz;
}

__global__ void nl_assign_bins(int n_atoms, const double* __restrict__ positions, const int* __restrict__ atom_system,
                               const NLSystem* __restrict__ systems, long long* __restrict__ atom_bin,
                               int* __restrict__ atom_wrap, int* __restrict__ bin_count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_atoms) return;
    const NLSystem& sys = systems[atom_system[i]];
    const double* p = positions + 3 * (long long)i;
    double frac[3];
    if (sys.periodic) {
        // Vector3D * Matrix3 (featomic/src/types/matrix.rs:400-407)
        frac[0] = p[0] * sys.inv[0] + p[1] * sys.inv[3] + p[2] * sys.inv[6];
        frac[1] = p[0] * sys.inv[1] + p[1] * sys.inv[4] + p[2] * sys.inv[7];
        frac[2] = p[0] * sys.inv[2] + p[1] * sys.inv[5] + p[2] * sys.inv[8];
    } else {
        frac[0] = p[0];
        frac[1] = p[1];
        frac[2] = p[2];
    }
    int b[3], w[3];
    for (int k = 0; k < 3; k++) {
        const int idx = floor_to_int(frac[k] * (double)sys.n_cells[k]);
        if (sys.periodic) {
            divmod(idx, sys.n_cells[k], &w[k], &b[k]);
        } else {
            // `i32 as usize` wraps negative values, the clamp then picks the last bin
            b[k] = (idx < 0 || idx > sys.n_cells[k] - 1) ? sys.n_cells[k] - 1 : idx;
            w[k] = 0;
        }
    }
    const long long bin = sys.bin_offset + ((long long)b[0] * sys.n_cells[1] + b[1]) * sys.n_cells[2] + b[2];
    atom_bin[i] = bin;
    atom_wrap[3 * i + 0] = w[0];
    atom_wrap[3 * i + 1] = w[1];
    atom_wrap[3 * i + 2] = w[2];
    atomicAdd(&bin_count[bin], 1);
}

__global__ void nl_fill_bins(int n_atoms, const long long* __restrict__ atom_bin, const long long* __restrict__ bin_start,
                             int* __restrict__ bin_fill, int* __restrict__ bin_atoms) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_atoms) return;
    const long long bin = atom_bin[i];
    const int slot = atomicAdd(&bin_fill[bin], 1);
    bin_atoms[bin_start[bin] + slot] = i;
}

// One wave per center. The lanes first take one searched bin each (its shift, first
// atom and atom count), and a prefix sum over the wave lays the bins' atoms end to
// end; the lanes then take these candidates 64 at a time, each finding its bin by a
// binary search over the prefix sums. So the candidate loads of a pass are issued
// together, and no lane idles on a bin shorter than the wave. `emit` = false counts,
// true writes packed keys: each hit goes to the center's running offset plus its rank
// among the wave's hits (the per-center order is fixed later by the segment sort).
// keys per center that the counting pass keeps for the emitting pass (centers with
// more are searched again)
constexpr int NL_STASH = 128;
constexpr int NL_WAVES = 4;

template <bool EMIT>
__global__ void __launch_bounds__(64 * NL_WAVES)
nl_candidates(int n_atoms, const double* __restrict__ positions, const int* __restrict__ atom_system,
              const NLSystem* __restrict__ systems, const long long* __restrict__ atom_bin,
              const int* __restrict__ atom_wrap, const long long* __restrict__ bin_start,
              const int* __restrict__ bin_len, const int* __restrict__ bin_atoms, double cutoff2,
              long long* __restrict__ counts, const long long* __restrict__ offsets,
              unsigned long long* __restrict__ keys, unsigned long long* __restrict__ stash,
              int* __restrict__ error_flag) {
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * NL_WAVES + (threadIdx.x >> 6);
    if (i >= n_atoms) return;  // uniform over the wave
    if (EMIT && stash) {
        // the counting pass kept this center's keys: copy them
        const long long first = offsets[i];
        const long long n = offsets[i + 1] - first;
        if (n <= NL_STASH) {
            for (int k = lane; k < n; k += 64) keys[first + k] = stash[(long long)i * NL_STASH + k];
            return;
        }
    }
    const NLSystem& sys = systems[atom_system[i]];
    const long long local_bin = atom_bin[i] - sys.bin_offset;
    const int nc0 = sys.n_cells[0], nc1 = sys.n_cells[1], nc2 = sys.n_cells[2];
    const int b0 = (int)(local_bin / ((long long)nc1 * nc2));
    const int b1 = (int)((local_bin / nc2) % nc1);
    const int b2 = (int)(local_bin % nc2);
    const int wi0 = atom_wrap[3 * i], wi1 = atom_wrap[3 * i + 1], wi2 = atom_wrap[3 * i + 2];
    const double* pi = positions + 3 * (long long)i;
    const unsigned long long below = (1ull << lane) - 1ull;
    const int w0 = 2 * sys.n_search[0] + 1, w1 = 2 * sys.n_search[1] + 1, w2 = 2 * sys.n_search[2] + 1;
    const int n_slots = w0 * w1 * w2;
    long long count = 0;
    long long out = EMIT ? offsets[i] : 0;
    for (int slot_base = 0; slot_base < n_slots; slot_base += 64) {
        // this lane's bin: offsets (d0, d1, d2) in the order of three nested loops
        int q0 = 0, q1 = 0, q2 = 0, len = 0;
        long long start = 0;
        const int slot = slot_base + lane;
        if (slot < n_slots) {
            const int d2 = slot % w2 - sys.n_search[2];
            const int d1 = (slot / w2) % w1 - sys.n_search[1];
            const int d0 = slot / (w1 * w2) - sys.n_search[0];
            int r0, r1, r2;
            divmod(b0 + d0, nc0, &q0, &r0);
            divmod(b1 + d1, nc1, &q1, &r1);
            divmod(b2 + d2, nc2, &q2, &r2);
            if (sys.periodic || (q0 == 0 && q1 == 0 && q2 == 0)) {
                const long long nb = sys.bin_offset + ((long long)r0 * nc1 + r1) * nc2 + r2;
                start = bin_start[nb];
                len = bin_len[nb];
            }
        }
        // exclusive prefix sum of the bins' lengths over the wave
        int incl = len;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int v = __shfl_up(incl, d);
            if (lane >= d) incl += v;
        }
        const int pre = incl - len;
        const int total = __shfl(incl, 63);
        for (int base = 0; base < total; base += 64) {
            const int a = base + lane;
            // the lane's bin: the last lane whose prefix is at most a
            int t = 0;
#pragma unroll
            for (int step = 32; step > 0; step >>= 1) {
                const int p = __shfl(pre, t + step);
                if (p <= a) t += step;
            }
            const int off = a - __shfl(pre, t);
            const long long st = ((long long)__shfl((int)(start >> 32), t) << 32) |
                                 (unsigned int)__shfl((int)(start & 0xffffffff), t);
            const int sq0 = __shfl(q0, t), sq1 = __shfl(q1, t), sq2 = __shfl(q2, t);
            bool hit = false, overflow = false;
            unsigned long long key = 0;
            if (a < total) {
                const int j = bin_atoms[st + off];
                int s[3];
                if (sys.periodic) {
                    s[0] = sq0 + wi0 - atom_wrap[3 * j];
                    s[1] = sq1 + wi1 - atom_wrap[3 * j + 1];
                    s[2] = sq2 + wi2 - atom_wrap[3 * j + 2];
                } else {
                    s[0] = s[1] = s[2] = 0;
                }
                if (!(j == i && s[0] == 0 && s[1] == 0 && s[2] == 0)) {
                    double v[3];
                    pair_vector(pi, positions + 3 * (long long)j, sys.cell, s, v);
                    const double d2v = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
                    hit = d2v < cutoff2;
                    if ((EMIT || stash) && hit) {
                        overflow = s[0] < -SHIFT_BIAS || s[0] >= SHIFT_BIAS || s[1] < -SHIFT_BIAS ||
                                   s[1] >= SHIFT_BIAS || s[2] < -SHIFT_BIAS || s[2] >= SHIFT_BIAS;
                        key = pack_key(j, s[0], s[1], s[2]);
                    }
                }
            }
            const unsigned long long mask = __ballot(hit);
            if (EMIT && hit) {
                keys[out + __popcll(mask & below)] = key;
                if (overflow) atomicOr(error_flag, 1);
            }
            if (!EMIT && stash && hit) {
                // the first NL_STASH keys, for the emitting pass to copy
                const long long rank = count + __popcll(mask & below);
                if (rank < NL_STASH) stash[(long long)i * NL_STASH + rank] = key;
                if (overflow) atomicOr(error_flag, 1);
            }
            const int hits = __popcll(mask);
            out += hits;
            count += hits;
        }
    }
    if (!EMIT && lane == 0) counts[i] = count;
}

// Bitonic sort of each center's key segment of at most 64 keys in registers, one wave
// per center (key k in lane k, compare-exchange partners by lane shuffles); longer
// segments are left to nl_sort_segments.
constexpr int SORT_WAVES = 4;

__global__ void __launch_bounds__(64 * SORT_WAVES) nl_sort_short_segments(int n_atoms, const long long* __restrict__ offsets,
                                                                         unsigned long long* __restrict__ keys) {
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * SORT_WAVES + (threadIdx.x >> 6);
    if (i >= n_atoms) return;  // uniform over the wave
    const long long start = offsets[i];
    const int len = (int)(offsets[i + 1] - start);
    if (len <= 1 || len > 64) return;
    unsigned long long key = lane < len ? keys[start + lane] : ~0ull;
    for (int k = 2; k <= 64; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            const unsigned int lo = __shfl_xor((unsigned int)(key & 0xffffffffull), j);
            const unsigned int hi = __shfl_xor((unsigned int)(key >> 32), j);
            const unsigned long long other = ((unsigned long long)hi << 32) | lo;
            const bool ascending = (lane & k) == 0;
            const bool lower = (lane & j) == 0;
            const unsigned long long small = key < other ? key : other;
            const unsigned long long large = key < other ? other : key;
            key = (lower == ascending) ? small : large;
        }
    }
    if (lane < len) keys[start + lane] = key;
}

// Bitonic sort of each center's key segment of more than 64 keys in LDS (one workgroup
// per center).
constexpr int SORT_CAPACITY = 2048;

__global__ void __launch_bounds__(256) nl_sort_segments(int n_atoms, const long long* __restrict__ offsets,
                                                        unsigned long long* __restrict__ keys, int* __restrict__ error_flag) {
    __shared__ unsigned long long buf[SORT_CAPACITY];
    const int i = blockIdx.x;
    if (i >= n_atoms) return;
    const long long start = offsets[i];
    const int len = (int)(offsets[i + 1] - start);
    if (len <= 64) return;  // nl_sort_short_segments
    if (len > SORT_CAPACITY) {
        if (threadIdx.x == 0) atomicOr(error_flag, 2);
        return;
    }
    int size = 1;
    while (size < len) size <<= 1;
    for (int k = threadIdx.x; k < size; k += blockDim.x) buf[k] = k < len ? keys[start + k] : ~0ull;
    __syncthreads();
    for (int width = 2; width <= size; width <<= 1) {
        for (int stride = width >> 1; stride > 0; stride >>= 1) {
            for (int k = threadIdx.x; k < size; k += blockDim.x) {
                const int partner = k ^ stride;
                if (partner > k) {
                    const bool ascending = (k & width) == 0;
                    const unsigned long long a = buf[k], b = buf[partner];
                    if ((a > b) == ascending) {
                        buf[k] = b;
                        buf[partner] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
    for (int k = threadIdx.x; k < len; k += blockDim.x) keys[start + k] = buf[k];
}

// Decode keys into neighbour index / shift, and compute vectors and distances.
__global__ void nl_finalize(long long n_entries, int n_atoms, const long long* __restrict__ offsets,
                            const unsigned long long* __restrict__ keys, const double* __restrict__ positions,
                            const int* __restrict__ atom_system, const NLSystem* __restrict__ systems,
                            int* __restrict__ entry_center, int* __restrict__ entry_j, int* __restrict__ entry_shift,
                            double* __restrict__ entry_vec, double* __restrict__ entry_dist) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n_entries) return;
    // binary search the owning center
    int lo = 0, hi = n_atoms;
    while (hi - lo > 1) {
        const int mid = (lo + hi) / 2;
        if (offsets[mid] <= e) lo = mid; else hi = mid;
    }
    const int i = lo;
    int j, s[3];
    unpack_key(keys[e], &j, s);
    const NLSystem& sys = systems[atom_system[i]];
    double v[3];
    pair_vector(positions + 3 * (long long)i, positions + 3 * (long long)j, sys.cell, s, v);
    const double d2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    entry_center[e] = i;
    entry_j[e] = j;
    entry_shift[3 * e + 0] = s[0];
    entry_shift[3 * e + 1] = s[1];
    entry_shift[3 * e + 2] = s[2];
    entry_vec[3 * e + 0] = v[0];
    entry_vec[3 * e + 1] = v[1];
    entry_vec[3 * e + 2] = v[2];
    entry_dist[e] = sqrt(d2);
}

// Per-center neighbour-type mask (incl. the center's own type) and the number
// of distinct neighbour atoms j != i of each species.
__global__ void center_meta(int n_atoms, int n_species, const long long* __restrict__ offsets,
                            const int* __restrict__ entry_j, const int* __restrict__ species,
                            unsigned* __restrict__ type_mask, int* __restrict__ unique_count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_atoms) return;
    unsigned mask = 1u << species[i];
    for (int t = 0; t < n_species; t++) unique_count[(long long)i * n_species + t] = 0;
    int prev = -1;
    for (long long e = offsets[i]; e < offsets[i + 1]; e++) {
        const int j = entry_j[e];
        const int t = species[j];
        mask |= 1u << t;
        if (j != prev && j != i) unique_count[(long long)i * n_species + t] += 1;
        prev = j;
    }
    type_mask[i] = mask;
}

// Block membership per (block, atom): PS blocks are (ct, t1 <= t2), SE blocks (ct, nt).
__global__ void block_membership(int n_atoms, int n_species, int ps_mode, int n_blocks, const int* __restrict__ species,
                                 const unsigned* __restrict__ type_mask, const int* __restrict__ unique_count,
                                 const int* __restrict__ block_types, long long* __restrict__ sample_flag,
                                 long long* __restrict__ row_count) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)n_blocks * n_atoms) return;
    const int b = (int)(idx / n_atoms);
    const int i = (int)(idx % n_atoms);
    const int ct = block_types[3 * b], t1 = block_types[3 * b + 1], t2 = block_types[3 * b + 2];
    const unsigned mask = type_mask[i];
    long long flag = 0, rows = 0;
    if (species[i] == ct) {
        if (ps_mode) {
            if (((mask >> t1) & 1u) && ((mask >> t2) & 1u)) {
                flag = 1;
                rows = 1 + unique_count[(long long)i * n_species + t1] + (t2 != t1 ? unique_count[(long long)i * n_species + t2] : 0);
            }
        } else {
            if ((mask >> t1) & 1u) {
                flag = 1;
                rows = 1 + unique_count[(long long)i * n_species + t1];
            }
        }
    }
    sample_flag[idx] = flag;
    row_count[idx] = rows;
}

__global__ void gather_block_totals(int n_blocks, long long n_atoms, const long long* __restrict__ sample_scan,
                                    const long long* __restrict__ row_scan, long long* __restrict__ totals) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b > n_blocks) return;
    totals[2 * b] = sample_scan[(long long)b * n_atoms];
    totals[2 * b + 1] = row_scan[(long long)b * n_atoms];
}

// Sample labels (system, atom) for every block sample.
__global__ void sample_labels(int n_blocks, int n_atoms, const long long* __restrict__ sample_flag,
                              const long long* __restrict__ sample_scan, const int* __restrict__ atom_system,
                              const int* __restrict__ atom_local, int* __restrict__ labels) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)n_blocks * n_atoms) return;
    if (!sample_flag[idx]) return;
    const int i = (int)(idx % n_atoms);
    const long long s = sample_scan[idx];
    labels[2 * s] = atom_system[i];
    labels[2 * s + 1] = atom_local[i];
}

}  // namespace

// ---------------------------------------------------------------------------- launchers

void launch_nl_assign_bins(hipStream_t stream, int n_atoms, const double* positions, const int* atom_system,
                           const NLSystem* systems, long long* atom_bin, int* atom_wrap, int* bin_count) {
    if (n_atoms == 0) return;
    nl_assign_bins<<<(n_atoms + 255) / 256, 256, 0, stream>>>(n_atoms, positions, atom_system, systems, atom_bin,
                                                              atom_wrap, bin_count);
}

void launch_nl_fill_bins(hipStream_t stream, int n_atoms, const long long* atom_bin, const long long* bin_start,
                         int* bin_fill, int* bin_atoms) {
    if (n_atoms == 0) return;
    nl_fill_bins<<<(n_atoms + 255) / 256, 256, 0, stream>>>(n_atoms, atom_bin, bin_start, bin_fill, bin_atoms);
}

size_t nl_stash_keys(int n_atoms) { return (size_t)n_atoms * NL_STASH; }

void launch_nl_candidates(hipStream_t stream, bool emit, int n_atoms, const double* positions, const int* atom_system,
                          const NLSystem* systems, const long long* atom_bin, const int* atom_wrap,
                          const long long* bin_start, const int* bin_len, const int* bin_atoms, double cutoff2,
                          long long* counts, const long long* offsets, unsigned long long* keys, int* error_flag,
                          unsigned long long* stash) {
    if (n_atoms == 0) return;
    const int threads = 64 * NL_WAVES;
    const int blocks = (n_atoms + NL_WAVES - 1) / NL_WAVES;
    if (emit) {
        nl_candidates<true><<<blocks, threads, 0, stream>>>(n_atoms, positions, atom_system, systems, atom_bin, atom_wrap,
                                                            bin_start, bin_len, bin_atoms, cutoff2, counts, offsets,
                                                            keys, stash, error_flag);
    } else {
        nl_candidates<false><<<blocks, threads, 0, stream>>>(n_atoms, positions, atom_system, systems, atom_bin, atom_wrap,
                                                             bin_start, bin_len, bin_atoms, cutoff2, counts, offsets,
                                                             keys, stash, error_flag);
    }
}

void launch_nl_sort_segments(hipStream_t stream, int n_atoms, const long long* offsets, unsigned long long* keys,
                             int* error_flag) {
    if (n_atoms == 0) return;
    nl_sort_short_segments<<<(n_atoms + SORT_WAVES - 1) / SORT_WAVES, 64 * SORT_WAVES, 0, stream>>>(n_atoms, offsets,
                                                                                                   keys);
    nl_sort_segments<<<n_atoms, 256, 0, stream>>>(n_atoms, offsets, keys, error_flag);
}

void launch_nl_finalize(hipStream_t stream, long long n_entries, int n_atoms, const long long* offsets,
                        const unsigned long long* keys, const double* positions, const int* atom_system,
                        const NLSystem* systems, int* entry_center, int* entry_j, int* entry_shift, double* entry_vec,
                        double* entry_dist) {
    if (n_entries == 0) return;
    nl_finalize<<<(unsigned)((n_entries + 255) / 256), 256, 0, stream>>>(n_entries, n_atoms, offsets, keys, positions,
                                                                        atom_system, systems, entry_center, entry_j,
                                                                        entry_shift, entry_vec, entry_dist);
}

void launch_center_meta(hipStream_t stream, int n_atoms, int n_species, const long long* offsets, const int* entry_j,
                        const int* species, unsigned* type_mask, int* unique_count) {
    if (n_atoms == 0) return;
    center_meta<<<(n_atoms + 255) / 256, 256, 0, stream>>>(n_atoms, n_species, offsets, entry_j, species, type_mask,
                                                           unique_count);
}

void launch_block_membership(hipStream_t stream, int n_atoms, int n_species, int ps_mode, int n_blocks,
                             const int* species, const unsigned* type_mask, const int* unique_count,
                             const int* block_types, long long* sample_flag, long long* row_count) {
    const long long total = (long long)n_blocks * n_atoms;
    if (total == 0) return;
    block_membership<<<(unsigned)((total + 255) / 256), 256, 0, stream>>>(n_atoms, n_species, ps_mode, n_blocks, species,
                                                                         type_mask, unique_count, block_types,
                                                                         sample_flag, row_count);
}

void launch_gather_block_totals(hipStream_t stream, int n_blocks, long long n_atoms, const long long* sample_scan,
                                const long long* row_scan, long long* totals) {
    gather_block_totals<<<(n_blocks + 1 + 127) / 128, 128, 0, stream>>>(n_blocks, n_atoms, sample_scan, row_scan, totals);
}

void launch_sample_labels(hipStream_t stream, int n_blocks, int n_atoms, const long long* sample_flag,
                          const long long* sample_scan, const int* atom_system, const int* atom_local, int* labels) {
    const long long total = (long long)n_blocks * n_atoms;
    if (total == 0) return;
    sample_labels<<<(unsigned)((total + 255) / 256), 256, 0, stream>>>(n_blocks, n_atoms, sample_flag, sample_scan,
                                                                      atom_system, atom_local, labels);
}

}  // namespace rascaline_amd
