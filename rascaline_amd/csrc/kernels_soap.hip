// SOAP spherical expansion and power spectrum kernels (gfx950, wave64).
//
// Layouts (T = float or double):
//   coef  [atom][species][L2][n]         c_nlm of every center for every neighbour species
//   dself [atom][species][3][L2][n]      d c / d r_center (self gradient)
//   dstrain / dcell [atom][species][9][L2][n]
// One 64-lane workgroup handles one center. Per-pair quantities (Y_lm, r dY_lm,
// spline values) are computed lane-parallel for a chunk of neighbour entries
// into LDS, then every lane accumulates its own (lm, n) slots.

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>

#include "device.hpp"
#include "kernels.hpp"
#include "soap_common.hpp"

namespace rascaline_amd {

namespace {

constexpr int WAVE = 64;
constexpr int MAX_SEGMENT = 2048;
constexpr int MAX_LOCAL_BLOCKS = 528;  // 32 * 33 / 2

// Per-chunk LDS layout of the per-entry quantities.
struct ChunkLayout {
    int L, n, L2, leg, chunk;
    int y, dy, rv, ra, rb, dir, vec, shift, flag, p, dp, pt;  // offsets in elements (per entry)
    int per_entry;

    __host__ __device__ ChunkLayout(int L_, int n_, int chunk_) : L(L_), n(n_), L2(L_ * L_), leg(L_ * (L_ + 1) / 2), chunk(chunk_) {
        int o = 0;
        y = o; o += L2;
        dy = o; o += 3 * L2;
        rv = o; o += L * n;
        ra = o; o += L * n;
        rb = o; o += L * n;
        dir = o; o += 3;
        vec = o; o += 3;
        shift = o; o += 3;
        flag = o; o += 1;
        p = o; o += leg;
        dp = o; o += leg;
        pt = o; o += leg;
        per_entry = o;
    }
};

template <typename T>
struct SoapArgs {
    int n_atoms, n_species, L, n;
    const long long* offsets;
    const int* entry_j;
    const int* entry_shift;
    const double* entry_vec;
    const double* entry_dist;
    const int* species;
    DeviceRadial<T> radial;
    DeviceSoapScalars scalars;
    const T* sh_a;
    const T* sh_b;
    const T* sh_c;  // spherical-harmonics coefficient table (ShTable layout)
    const T* self_contribution;
    int chunk;
    int* error_flag;
};

// Fill the per-entry LDS slot `base` for neighbour entry e of center i.
template <typename T>
__device__ void entry_data(const SoapArgs<T>& a, const ChunkLayout& lay, long long e, int i, T* base) {
    const double r = a.entry_dist[e];
    double d[3];
    if (r < 1e-6) {
        d[0] = 0.0;
        d[1] = 0.0;
        d[2] = 1.0;
    } else {
        d[0] = a.entry_vec[3 * e] / r;
        d[1] = a.entry_vec[3 * e + 1] / r;
        d[2] = a.entry_vec[3 * e + 2] / r;
    }
    T dt[3] = {(T)d[0], (T)d[1], (T)d[2]};
    spherical_harmonics<T>(lay.L, dt, base + lay.y, base + lay.dy, base + lay.p, base + lay.dp, base + lay.pt, a.sh_a, a.sh_b);
    double f, df;
    scaling_functions(a.scalars, r, &f, &df);
    const int n = lay.n;
    for (int l = 0; l < lay.L; l++) {
        // R into the rv slots, dR/dr into the ra slots, then combine in place
        radial_integral<T>(a.radial, l, e, r, base + lay.rv + l * n, base + lay.ra + l * n);
        for (int k = 0; k < n; k++) {
            const T R = base[lay.rv + l * n + k], dR = base[lay.ra + l * n + k];
            base[lay.rv + l * n + k] = (T)f * R;
            base[lay.ra + l * n + k] = (T)df * R + (T)f * dR;
            base[lay.rb + l * n + k] = (T)f * R / (T)r;
        }
    }
    for (int x = 0; x < 3; x++) {
        base[lay.dir + x] = dt[x];
        base[lay.vec + x] = (T)a.entry_vec[3 * e + x];
        base[lay.shift + x] = (T)a.entry_shift[3 * e + x];
    }
    base[lay.flag] = (a.entry_j[e] == i) ? (T)1 : (T)0;
}

__device__ __forceinline__ int l_of_lm(int lm) {
    int l = (int)sqrtf((float)lm);
    while ((l + 1) * (l + 1) <= lm) l++;
    while (l * l > lm) l--;
    return l;
}

// MODE 0: coef + dself; MODE 1: dstrain; MODE 2: dcell.
// The center's entries are taken in neighbour-species order (stable partition),
// in chunks of up to `chunk` entries whose per-entry data all lanes compute in
// parallel; each lane then accumulates its own (lm, n) slots, flushing its
// accumulators to global memory whenever the species changes.
template <typename T, int MS, int MODE>
__global__ void __launch_bounds__(WAVE) soap_center_kernel(SoapArgs<T> a, T* __restrict__ out_coef, T* __restrict__ out_grad) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int list[MAX_SEGMENT];
    __shared__ int chunk_t[64];
    const int i = blockIdx.x;
    const int lane = threadIdx.x;
    const long long e0 = a.offsets[i];
    const int len = (int)(a.offsets[i + 1] - e0);
    const int S = a.n_species;
    if (len > MAX_SEGMENT) {
        if (lane == 0) atomicOr(a.error_flag, 2);
        return;
    }
    // stable partition of the entries by neighbour species
    int cursor = 0;
    for (int t = 0; t < S; t++) {
        for (int base = 0; base < len; base += WAVE) {
            const int e = base + lane;
            const bool pred = e < len && a.species[a.entry_j[e0 + e]] == t;
            const unsigned long long mask = __ballot(pred);
            const int pos = __popcll(mask & ((1ull << lane) - 1ull));
            if (pred) list[cursor + pos] = e;
            cursor += __popcll(mask);
        }
    }
    __syncthreads();

    const ChunkLayout lay(a.L, a.n, a.chunk);
    T* chunk = reinterpret_cast<T*>(smem);
    const int n = a.n;
    const int slots = lay.L2 * n;
    const int ct = a.species[i];
    const int groups = (slots + WAVE * MS - 1) / (WAVE * MS);
    constexpr int NACC = MODE == 0 ? 3 : 9;

    for (int g = 0; g < groups; g++) {
        const int sb = g * WAVE * MS;
        T acc[NACC][MS];
        T accv[MS];
        unsigned written = 0;
        int cur_t = -1;

        auto reset = [&]() {
#pragma unroll
            for (int k = 0; k < MS; k++) {
                accv[k] = 0;
#pragma unroll
                for (int x = 0; x < NACC; x++) acc[x][k] = 0;
            }
        };
        auto flush = [&](int t) {
#pragma unroll
            for (int k = 0; k < MS; k++) {
                const int s = sb + lane + WAVE * k;
                if (s < slots) {
                    if (MODE == 0) {
                        T v = accv[k];
                        if (t == ct && s < n) v += a.self_contribution[s];
                        out_coef[((long long)i * S + t) * slots + s] = v;
                        if (out_grad) {
#pragma unroll
                            for (int x = 0; x < 3; x++) out_grad[(((long long)i * S + t) * 3 + x) * slots + s] = acc[x][k];
                        }
                    } else {
#pragma unroll
                        for (int x = 0; x < 9; x++) out_grad[(((long long)i * S + t) * 9 + x) * slots + s] = acc[x][k];
                    }
                }
            }
        };

        reset();
        for (int cb = 0; cb < len; cb += lay.chunk) {
            const int cnt = min(lay.chunk, len - cb);
            if (lane < cnt) {
                const int e = list[cb + lane];
                entry_data<T>(a, lay, e0 + e, i, chunk + lane * lay.per_entry);
                chunk_t[lane] = a.species[a.entry_j[e0 + e]];
            }
            __syncthreads();
            for (int c = 0; c < cnt; c++) {
                const int t = chunk_t[c];
                if (t != cur_t) {
                    if (cur_t >= 0) {
                        flush(cur_t);
                        written |= 1u << cur_t;
                    }
                    reset();
                    cur_t = t;
                }
                const T* en = chunk + c * lay.per_entry;
                const bool self_image = en[lay.flag] != (T)0;
#pragma unroll
                for (int k = 0; k < MS; k++) {
                    const int s = sb + lane + WAVE * k;
                    if (s < slots) {
                        const int lm = s / n, nn = s - lm * n;
                        const int l = l_of_lm(lm);
                        const T y = en[lay.y + lm];
                        if (MODE == 0) {
                            accv[k] += en[lay.rv + l * n + nn] * y;
                            if (out_grad && !self_image) {
                                const T ra = en[lay.ra + l * n + nn] * y, rb = en[lay.rb + l * n + nn];
#pragma unroll
                                for (int x = 0; x < 3; x++) acc[x][k] -= ra * en[lay.dir + x] + rb * en[lay.dy + x * lay.L2 + lm];
                            }
                        } else {
                            const T ra = en[lay.ra + l * n + nn] * y, rb = en[lay.rb + l * n + nn];
                            T gx[3];
#pragma unroll
                            for (int x = 0; x < 3; x++) gx[x] = ra * en[lay.dir + x] + rb * en[lay.dy + x * lay.L2 + lm];
#pragma unroll
                            for (int p = 0; p < 3; p++) {
                                const T w = MODE == 1 ? en[lay.vec + p] : en[lay.shift + p];
#pragma unroll
                                for (int x = 0; x < 3; x++) acc[p * 3 + x][k] += w * gx[x];
                            }
                        }
                    }
                }
            }
            __syncthreads();
        }
        if (cur_t >= 0) {
            flush(cur_t);
            written |= 1u << cur_t;
        }
        // species without neighbours: zeros (plus the self contribution)
        reset();
        for (int t = 0; t < S; t++) {
            if (!((written >> t) & 1u)) flush(t);
        }
    }
}

// Local PS block index of (t1 <= t2) among S species.
__device__ __forceinline__ int pair_index(int t1, int t2, int S) { return t1 * S - t1 * (t1 - 1) / 2 + (t2 - t1); }

// Σ_m C2[m][n2] D1[x][m][n1] (if use1) + Σ_m C1[m][n1] D2[x][m][n2] (if use2) for all
// (x, l, n1, n2), scaled by kappa / ((-1)^l sqrt(2l+1)) (soap/power_spectrum.rs:340-348,604-675).
template <typename T>
__device__ void ps_contract(int ncomp, int L, int n, const T* C1, const T* C2, const T* D1, const T* D2, bool use1,
                            bool use2, T kappa, int slots, T* out) {
    const int P = L * n * n;
    for (int o = threadIdx.x; o < ncomp * P; o += WAVE) {
        const int x = o / P, p = o - x * P;
        const int l = p / (n * n), n1 = (p / n) % n, n2 = p % n;
        T sum = 0;
        for (int lm = l * l; lm < (l + 1) * (l + 1); lm++) {
            if (use1) sum += C2[lm * n + n2] * D1[x * slots + lm * n + n1];
            if (use2) sum += C1[lm * n + n1] * D2[x * slots + lm * n + n2];
        }
        const T norm = (l % 2 == 0 ? (T)1 : (T)-1) * (T)sqrt((double)(2 * l + 1));
        out[o] = (sum * kappa) / norm;
    }
}

template <typename T>
__global__ void __launch_bounds__(WAVE) ps_values_kernel(int n_atoms, int S, int L, int n, const T* __restrict__ coef,
                                                         const int* __restrict__ species, const unsigned* __restrict__ type_mask,
                                                         const long long* __restrict__ sample_flag,
                                                         const long long* __restrict__ sample_scan, T* __restrict__ out) {
    extern __shared__ __align__(16) unsigned char smem[];
    T* C = reinterpret_cast<T*>(smem);
    const int i = blockIdx.x;
    const int slots = L * L * n;
    for (int k = threadIdx.x; k < S * slots; k += WAVE) C[k] = coef[(long long)i * S * slots + k];
    __syncthreads();
    const int ct = species[i];
    const unsigned mask = type_mask[i];
    const int P = L * n * n;
    const int npairs = S * (S + 1) / 2;
    for (int t1 = 0; t1 < S; t1++) {
        if (!((mask >> t1) & 1u)) continue;
        for (int t2 = t1; t2 < S; t2++) {
            if (!((mask >> t2) & 1u)) continue;
            const long long idx = (long long)(ct * npairs + pair_index(t1, t2, S)) * n_atoms + i;
            if (!sample_flag[idx]) continue;
            const long long s = sample_scan[idx];
            const T kappa = t1 != t2 ? (T)1.4142135623730951 : (T)1;
            ps_contract<T>(1, L, n, C + t1 * slots, C + t2 * slots, C + t1 * slots, C + t2 * slots, true, false, kappa,
                           slots, out + s * P);
        }
    }
}

// Gradients w.r.t. positions for the power spectrum (PS = true) or the spherical
// expansion (PS = false). Iterates over the center's sorted neighbour entries,
// sums the images of each neighbour atom into DC = d c^{type_j} / d r_j, and
// emits one gradient row per block containing type_j; the center's own rows use
// the self gradients.
template <typename T, bool PS>
__global__ void __launch_bounds__(WAVE) gradient_rows_kernel(SoapArgs<T> a, const T* __restrict__ coef,
                                                             const T* __restrict__ dself, const unsigned* __restrict__ type_mask,
                                                             const long long* __restrict__ sample_flag,
                                                             const long long* __restrict__ sample_scan,
                                                             const long long* __restrict__ row_scan,
                                                             const long long* __restrict__ block_sample_base,
                                                             const long long* __restrict__ se_grad_base,
                                                             const int* __restrict__ atom_system, const int* __restrict__ atom_local,
                                                             T* __restrict__ out, int* __restrict__ labels) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int counters[MAX_LOCAL_BLOCKS];
    __shared__ int snapshot[MAX_LOCAL_BLOCKS];
    const int i = blockIdx.x;
    const int lane = threadIdx.x;
    const int S = a.n_species, L = a.L, n = a.n;
    const int L2 = L * L, slots = L2 * n;
    const int P = L * n * n;
    const int npairs = S * (S + 1) / 2;
    const int nloc = PS ? npairs : S;
    const ChunkLayout lay(L, n, a.chunk);
    T* C = reinterpret_cast<T*>(smem);          // S * slots (PS only)
    T* DC = C + (PS ? S * slots : 0);            // 3 * slots
    T* D2 = DC + 3 * slots;                      // 3 * slots (center rows)
    T* chunk = D2 + 3 * slots;
    const long long e0 = a.offsets[i];
    const int len = (int)(a.offsets[i + 1] - e0);
    const int ct = a.species[i];
    const unsigned mask = type_mask[i];
    const int n_atoms = a.n_atoms;

    if (PS) {
        for (int k = lane; k < S * slots; k += WAVE) C[k] = coef[(long long)i * S * slots + k];
    }
    for (int k = lane; k < 3 * slots; k += WAVE) DC[k] = 0;
    for (int k = lane; k < nloc; k += WAVE) {
        counters[k] = 0;
        snapshot[k] = -1;
    }
    __syncthreads();

    auto global_block = [&](int local) -> int { return PS ? ct * npairs + local : ct * S + local; };

    // emit all rows of neighbour j (or of the center when j == i, with D = dself)
    auto emit = [&](int j, int tj, bool center) {
        if (PS) {
            for (int t1 = 0; t1 < S; t1++) {
                if (!((mask >> t1) & 1u)) continue;
                for (int t2 = t1; t2 < S; t2++) {
                    if (!((mask >> t2) & 1u)) continue;
                    if (!center && t1 != tj && t2 != tj) continue;
                    const int local = pair_index(t1, t2, S);
                    const long long idx = (long long)global_block(local) * n_atoms + i;
                    const long long row = row_scan[idx] + (center ? snapshot[local] : counters[local] + (i < j ? 1 : 0));
                    const T kappa = t1 != t2 ? (T)1.4142135623730951 : (T)1;
                    const T *D1v, *D2v;
                    bool use1, use2;
                    if (center) {
                        // self gradients of both species, loaded into DC (t1) and D2 (t2)
                        for (int k = lane; k < 3 * slots; k += WAVE) {
                            DC[k] = dself[((long long)i * S + t1) * 3 * slots + k];
                            D2[k] = dself[((long long)i * S + t2) * 3 * slots + k];
                        }
                        __syncthreads();
                        D1v = DC;
                        D2v = D2;
                        use1 = use2 = true;
                    } else {
                        D1v = DC;
                        D2v = DC;
                        use1 = t1 == tj;
                        use2 = t2 == tj;
                    }
                    ps_contract<T>(3, L, n, C + t1 * slots, C + t2 * slots, D1v, D2v, use1, use2, kappa, slots,
                                   out + row * 3 * P);
                    if (lane == 0) {
                        labels[3 * row] = (int)(sample_scan[idx] - block_sample_base[global_block(local)]);
                        labels[3 * row + 1] = atom_system[i];
                        labels[3 * row + 2] = atom_local[j];
                    }
                    __syncthreads();
                }
            }
        } else {
            for (int nt = 0; nt < S; nt++) {
                if (!((mask >> nt) & 1u)) continue;
                if (!center && nt != tj) continue;
                const long long idx = (long long)global_block(nt) * n_atoms + i;
                const long long row = row_scan[idx] + (center ? snapshot[nt] : counters[nt] + (i < j ? 1 : 0));
                const T* D = center ? dself + ((long long)i * S + nt) * 3 * slots : DC;
                const long long tb = global_block(nt);
                for (int l = 0; l < L; l++) {
                    const int w = (2 * l + 1) * n;
                    T* o = out + se_grad_base[tb * L + l] + row * 3 * w;
                    for (int k = lane; k < 3 * w; k += WAVE) {
                        const int x = k / w, r = k - x * w;
                        o[k] = D[x * slots + l * l * n + r];
                    }
                }
                if (lane == 0) {
                    labels[3 * row] = (int)(sample_scan[idx] - block_sample_base[tb]);
                    labels[3 * row + 1] = atom_system[i];
                    labels[3 * row + 2] = atom_local[j];
                }
            }
            __syncthreads();
        }
    };

    auto bump = [&](int tj) {
        // count neighbour tj in every block that contains it
        if (lane == 0) {
            if (PS) {
                for (int t = 0; t < S; t++) {
                    if (!((mask >> t) & 1u)) continue;
                    const int lo = t < tj ? t : tj, hi = t < tj ? tj : t;
                    counters[pair_index(lo, hi, S)] += 1;
                }
            } else {
                counters[tj] += 1;
            }
        }
        __syncthreads();
    };

    auto take_snapshot = [&]() {
        for (int k = lane; k < nloc; k += WAVE) snapshot[k] = counters[k];
        __syncthreads();
    };

    bool snapped = false;
    for (int cb = 0; cb < len; cb += lay.chunk) {
        const int cnt = min(lay.chunk, len - cb);
        if (lane < cnt) entry_data<T>(a, lay, e0 + cb + lane, i, chunk + lane * lay.per_entry);
        __syncthreads();
        for (int c = 0; c < cnt; c++) {
            const long long e = e0 + cb + c;
            const int j = a.entry_j[e];
            if (j == i) continue;  // periodic self image: no gradient row of its own
            if (!snapped && j > i) {
                take_snapshot();
                snapped = true;
            }
            const T* en = chunk + c * lay.per_entry;
            for (int s = lane; s < slots; s += WAVE) {
                const int lm = s / n, nn = s - lm * n;
                const int l = l_of_lm(lm);
                const T y = en[lay.y + lm];
                const T ra = en[lay.ra + l * n + nn] * y, rb = en[lay.rb + l * n + nn];
                for (int x = 0; x < 3; x++) DC[x * slots + s] += ra * en[lay.dir + x] + rb * en[lay.dy + x * L2 + lm];
            }
            const bool last_image = (cb + c + 1 >= len) || a.entry_j[e + 1] != j;
            if (last_image) {
                __syncthreads();
                const int tj = a.species[j];
                emit(j, tj, false);
                bump(tj);
                for (int k = lane; k < 3 * slots; k += WAVE) DC[k] = 0;
                __syncthreads();
            }
        }
        __syncthreads();
    }
    if (!snapped) take_snapshot();
    emit(i, ct, true);
}

// ---------------------------------------------------------------------------
// Power-spectrum position gradients, register-blocked (n <= NMAX).
//
// One 256-thread workgroup per center. Neighbour entries are processed in
// chunks: one thread per entry computes Y_lm, r dY_lm and the spline values into
// LDS, then all threads sum the images of neighbour j into DC[x][lm][n] =
// d c_i^{type_j} / d r_j. Each gradient row [3][L][n][n] of a block containing
// type_j is produced by items (x, l, n1..n1+R-1) that keep R x n accumulators
// in registers: per m they read one row of C (or DC) and R scalars, so that a
// thread writes R x n contiguous outputs and consecutive threads write
// consecutive addresses.

struct GradLayout {
    int L, n, L2, leg;
    int y, dy, ra, rb, dir, p, dp, pt, per_entry;
    __host__ __device__ GradLayout(int L_, int n_) : L(L_), n(n_), L2(L_ * L_), leg(L_ * (L_ + 1) / 2) {
        int o = 0;
        y = o; o += L2;
        dy = o; o += 3 * L2;
        ra = o; o += L * n;
        rb = o; o += L * n;
        dir = o; o += 4;
        p = o; o += leg;
        dp = o; o += leg;
        pt = o; o += leg;
        per_entry = o;
    }
};

template <typename T>
__device__ void entry_grad_data(const SoapArgs<T>& a, const GradLayout& lay, long long e, T* base) {
    const double r = a.entry_dist[e];
    double d[3];
    if (r < 1e-6) {
        d[0] = 0.0;
        d[1] = 0.0;
        d[2] = 1.0;
    } else {
        d[0] = a.entry_vec[3 * e] / r;
        d[1] = a.entry_vec[3 * e + 1] / r;
        d[2] = a.entry_vec[3 * e + 2] / r;
    }
    T dt[3] = {(T)d[0], (T)d[1], (T)d[2]};
    spherical_harmonics<T>(lay.L, dt, base + lay.y, base + lay.dy, base + lay.p, base + lay.dp, base + lay.pt, a.sh_a, a.sh_b);
    double f, df;
    scaling_functions(a.scalars, r, &f, &df);
    const int n = lay.n;
    for (int l = 0; l < lay.L; l++) {
        radial_integral<T>(a.radial, l, e, r, base + lay.ra + l * n, base + lay.rb + l * n);
        for (int k = 0; k < n; k++) {
            const T R = base[lay.ra + l * n + k], dR = base[lay.rb + l * n + k];
            base[lay.ra + l * n + k] = (T)df * R + (T)f * dR;
            base[lay.rb + l * n + k] = (T)f * R / (T)r;
        }
    }
    base[lay.dir + 0] = dt[0];
    base[lay.dir + 1] = dt[1];
    base[lay.dir + 2] = dt[2];
}

constexpr int GRAD_THREADS = 256;
constexpr int MAX_ROW_BLOCKS = 32;

template <typename T, int NMAX, int R>
__device__ __forceinline__ void emit_row_items(int n_blocks, const long long* jb_row, const int* jb_t1, const int* jb_t2,
                                               const unsigned char* jb_use, const T* C, const T* D1base, const T* D2base,
                                               long long d_stride_t, int L, int n, int slots, T* __restrict__ out) {
    const int nR = (n + R - 1) / R;
    const int per_block = 3 * L * nR;
    const int P = L * n * n;
    const int total = n_blocks * per_block;
    for (int item = threadIdx.x; item < total; item += GRAD_THREADS) {
        const int k = item / per_block;
        int rem = item - k * per_block;
        const int x = rem / (L * nR);
        rem -= x * (L * nR);
        const int l = rem / nR;
        const int n1 = (rem - l * nR) * R;
        const int t1 = jb_t1[k], t2 = jb_t2[k];
        const bool use1 = jb_use[k] & 1, use2 = jb_use[k] & 2;
        // D1 / D2: gradient arrays [3][L2][n] of the terms d c^{t1} and d c^{t2}
        const T* D1 = D1base + (d_stride_t ? d_stride_t * t1 : 0);
        const T* D2 = D2base + (d_stride_t ? d_stride_t * t2 : 0);
        T acc[R][NMAX];
#pragma unroll
        for (int r = 0; r < R; r++)
#pragma unroll
            for (int q = 0; q < NMAX; q++) acc[r][q] = 0;
        for (int m = l * l; m < (l + 1) * (l + 1); m++) {
            if (use1) {
                T d[R];
#pragma unroll
                for (int r = 0; r < R; r++) d[r] = (n1 + r < n) ? D1[x * slots + m * n + n1 + r] : (T)0;
                const T* crow = C + t2 * slots + m * n;
#pragma unroll
                for (int q = 0; q < NMAX; q++) {
                    if (q < n) {
                        const T cv = crow[q];
#pragma unroll
                        for (int r = 0; r < R; r++) acc[r][q] += cv * d[r];
                    }
                }
            }
            if (use2) {
                T c[R];
#pragma unroll
                for (int r = 0; r < R; r++) c[r] = (n1 + r < n) ? C[t1 * slots + m * n + n1 + r] : (T)0;
                const T* drow = D2 + x * slots + m * n;
#pragma unroll
                for (int q = 0; q < NMAX; q++) {
                    if (q < n) {
                        const T dv = drow[q];
#pragma unroll
                        for (int r = 0; r < R; r++) acc[r][q] += c[r] * dv;
                    }
                }
            }
        }
        const T kappa = t1 != t2 ? (T)1.4142135623730951 : (T)1;
        const T norm = (l % 2 == 0 ? (T)1 : (T)-1) * (T)sqrt((double)(2 * l + 1));
        T* o = out + jb_row[k] * 3 * P + (long long)x * P + l * n * n + n1 * n;
#pragma unroll
        for (int r = 0; r < R; r++) {
            if (n1 + r < n) {
#pragma unroll
                for (int q = 0; q < NMAX; q++) {
                    if (q < n) o[r * n + q] = (acc[r][q] * kappa) / norm;
                }
            }
        }
    }
}

// Per-entry data for the gradient kernel, split over threads: threads
// [0, cnt) compute the spherical harmonics of entry tid, threads
// [64, 64 + cnt * L) the radial integral of (entry, l).
template <typename T>
__device__ void chunk_grad_data(const SoapArgs<T>& a, const GradLayout& lay, long long e_first, int cnt, T* chunk_data) {
    const int tid = threadIdx.x;
    if (tid < cnt) {
        const long long e = e_first + tid;
        T* base = chunk_data + tid * lay.per_entry;
        const double r = a.entry_dist[e];
        double d[3];
        if (r < 1e-6) {
            d[0] = 0.0;
            d[1] = 0.0;
            d[2] = 1.0;
        } else {
            d[0] = a.entry_vec[3 * e] / r;
            d[1] = a.entry_vec[3 * e + 1] / r;
            d[2] = a.entry_vec[3 * e + 2] / r;
        }
        T dt[3] = {(T)d[0], (T)d[1], (T)d[2]};
        spherical_harmonics<T>(lay.L, dt, base + lay.y, base + lay.dy, base + lay.p, base + lay.dp, base + lay.pt, a.sh_a, a.sh_b);
        base[lay.dir + 0] = dt[0];
        base[lay.dir + 1] = dt[1];
        base[lay.dir + 2] = dt[2];
    }
    const int w = tid - WAVE;
    if (w >= 0 && w < cnt * lay.L) {
        const int c = w / lay.L, l = w - c * lay.L;
        const long long e = e_first + c;
        T* base = chunk_data + c * lay.per_entry;
        const double r = a.entry_dist[e];
        double f, df;
        scaling_functions(a.scalars, r, &f, &df);
        const int n = lay.n;
        radial_integral<T>(a.radial, l, e, r, base + lay.ra + l * n, base + lay.rb + l * n);
        for (int k = 0; k < n; k++) {
            const T Rv = base[lay.ra + l * n + k], dR = base[lay.rb + l * n + k];
            base[lay.ra + l * n + k] = (T)df * Rv + (T)f * dR;
            base[lay.rb + l * n + k] = (T)f * Rv / (T)r;
        }
    }
}

template <typename T, int NMAX, int R>
__global__ void __launch_bounds__(GRAD_THREADS) ps_gradient_kernel(
    SoapArgs<T> a, int chunk, const T* __restrict__ coef, const T* __restrict__ dself, const unsigned* __restrict__ type_mask,
    const long long* __restrict__ sample_scan, const long long* __restrict__ row_scan,
    const long long* __restrict__ block_sample_base, const int* __restrict__ atom_system, const int* __restrict__ atom_local,
    T* __restrict__ out, int* __restrict__ labels) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int chunk_j[64];
    __shared__ long long jb_row[MAX_ROW_BLOCKS];
    __shared__ int jb_t1[MAX_ROW_BLOCKS], jb_t2[MAX_ROW_BLOCKS], jb_local[MAX_ROW_BLOCKS];
    __shared__ unsigned char jb_use[MAX_ROW_BLOCKS];
    __shared__ int jb_count;

    const int i = blockIdx.x;
    const int tid = threadIdx.x;
    const int S = a.n_species, L = a.L, n = a.n;
    const int L2 = L * L, slots = L2 * n;
    const int npairs = S * (S + 1) / 2;
    const int n_atoms = a.n_atoms;
    const GradLayout lay(L, n);
    // dynamic LDS: per-block row bases and sample indices, then C, DC, chunk data
    long long* rowbase = reinterpret_cast<long long*>(smem);
    int* sample_in_block = reinterpret_cast<int*>(rowbase + npairs);
    int* counters = sample_in_block + npairs;
    int* snapshot = counters + npairs;
    const size_t meta_bytes = ((size_t)npairs * (8 + 12) + 15) / 16 * 16;
    T* C = reinterpret_cast<T*>(smem + meta_bytes);
    T* DC = C + S * slots;
    T* chunk_data = DC + 3 * slots;
    const long long e0 = a.offsets[i];
    const int len = (int)(a.offsets[i + 1] - e0);
    const int ct = a.species[i];
    const unsigned mask = type_mask[i];

    for (int k = tid; k < S * slots; k += GRAD_THREADS) C[k] = coef[(long long)i * S * slots + k];
    for (int k = tid; k < 3 * slots; k += GRAD_THREADS) DC[k] = 0;
    for (int k = tid; k < npairs; k += GRAD_THREADS) {
        const long long idx = (long long)(ct * npairs + k) * n_atoms + i;
        rowbase[k] = row_scan[idx];
        sample_in_block[k] = (int)(sample_scan[idx] - block_sample_base[ct * npairs + k]);
        counters[k] = 0;
        snapshot[k] = -1;
    }
    const int sys = atom_system[i];
    __syncthreads();

    bool snapped = false;
    for (int cb = 0; cb < len; cb += chunk) {
        const int cnt = min(chunk, len - cb);
        chunk_grad_data<T>(a, lay, e0 + cb, cnt, chunk_data);
        if (tid < cnt) chunk_j[tid] = a.entry_j[e0 + cb + tid];
        __syncthreads();
        for (int c = 0; c < cnt; c++) {
            const int j = chunk_j[c];
            if (j == i) continue;  // periodic self image: contributes to no gradient row
            if (!snapped && j > i) {
                for (int k = tid; k < npairs; k += GRAD_THREADS) snapshot[k] = counters[k];
                snapped = true;
            }
            const T* en = chunk_data + c * lay.per_entry;
            for (int s = tid; s < 3 * slots; s += GRAD_THREADS) {
                const int x = s / slots;
                const int r = s - x * slots;
                const int lm = r / n, nn = r - lm * n;
                const int l = l_of_lm(lm);
                DC[s] += en[lay.ra + l * n + nn] * en[lay.y + lm] * en[lay.dir + x] + en[lay.rb + l * n + nn] * en[lay.dy + x * L2 + lm];
            }
            const bool last = (cb + c + 1 >= len) || (c + 1 < cnt ? chunk_j[c + 1] != j : a.entry_j[e0 + cb + c + 1] != j);
            if (last) {
                const int tj = a.species[j];
                const int local_j = atom_local[j];
                // one thread per species t present around the center: block (min, max)(t, tj)
                if (tid < S && ((mask >> tid) & 1u)) {
                    const int t = tid;
                    const int k = __popc(mask & ((1u << t) - 1u));
                    const int lo = t < tj ? t : tj, hi = t < tj ? tj : t;
                    const int local = pair_index(lo, hi, S);
                    const long long row = rowbase[local] + counters[local] + (i < j ? 1 : 0);
                    jb_row[k] = row;
                    jb_t1[k] = lo;
                    jb_t2[k] = hi;
                    jb_local[k] = local;
                    jb_use[k] = (unsigned char)((lo == tj ? 1 : 0) | (hi == tj ? 2 : 0));
                    labels[3 * row] = sample_in_block[local];
                    labels[3 * row + 1] = sys;
                    labels[3 * row + 2] = local_j;
                }
                __syncthreads();
                emit_row_items<T, NMAX, R>(__popc(mask), jb_row, jb_t1, jb_t2, jb_use, C, DC, DC, 0, L, n, slots, out);
                __syncthreads();
                if (tid < S && ((mask >> tid) & 1u)) counters[jb_local[__popc(mask & ((1u << tid) - 1u))]] += 1;
                for (int k = tid; k < 3 * slots; k += GRAD_THREADS) DC[k] = 0;
                __syncthreads();
            }
        }
        __syncthreads();
    }
    if (!snapped) {
        for (int k = tid; k < npairs; k += GRAD_THREADS) snapshot[k] = counters[k];
    }
    __syncthreads();

    // center rows of every block: d c^{t} / d r_i are the self gradients
    const T* self_base = dself + (long long)i * S * 3 * slots;
    int t1 = 0, t2 = 0;
    for (;;) {
        // next group of up to MAX_ROW_BLOCKS blocks in (t1, t2) order
        if (tid == 0) {
            int count = 0;
            int a1 = t1, a2 = t2;
            while (a1 < S && count < MAX_ROW_BLOCKS) {
                if (a2 < a1) a2 = a1;
                if (a2 >= S) {
                    a1++;
                    a2 = a1;
                    continue;
                }
                if (((mask >> a1) & 1u) && ((mask >> a2) & 1u)) {
                    const int local = pair_index(a1, a2, S);
                    const long long row = rowbase[local] + snapshot[local];
                    jb_row[count] = row;
                    jb_t1[count] = a1;
                    jb_t2[count] = a2;
                    jb_use[count] = 3;
                    labels[3 * row] = sample_in_block[local];
                    labels[3 * row + 1] = sys;
                    labels[3 * row + 2] = atom_local[i];
                    count++;
                }
                a2++;
            }
            jb_count = count;
            jb_local[0] = a1;
            jb_local[1] = a2;
        }
        __syncthreads();
        const int count = jb_count;
        const int next1 = jb_local[0], next2 = jb_local[1];
        if (count > 0) {
            emit_row_items<T, NMAX, R>(count, jb_row, jb_t1, jb_t2, jb_use, C, self_base, self_base,
                                       (long long)3 * slots, L, n, slots, out);
        }
        __syncthreads();
        if (next1 >= S || count == 0) break;
        t1 = next1;
        t2 = next2;
    }
}

// ---------------------------------------------------------------------------
// Power-spectrum position gradients, planned, windowed and tiled (n padded to NC,
// a multiple of 4).
//
// A planning pass (one wave per center) splits every center's sorted neighbour
// entries into runs (all periodic images of one neighbour atom j), assigns every
// output row in the reference's order (rows sorted by atom within each block,
// the center's own row before its first neighbour j > i), writes all gradient
// sample labels, and groups the runs into windows of at most `max_runs` runs and
// `max_entries` entries (a longer run is split over several windows whose sums
// are carried). Runs of periodic images of the center itself produce no row and
// are left out. The gradient kernel (one 256-thread workgroup per center) then
// processes the windows:
//   1. one lane per (entry, m) runs the Legendre recursion along l (plus the
//      direction, trigonometric and cutoff factors on the m = 0 lane), one lane
//      per (entry, l) evaluates the radial spline;
//   2. one lane per (x, run, l, m >= 0) forms Y_l(±m) and r dY_l(±m)/dx and sums
//      DC_r[x][lm][:] = d c_i^{type j} / d r_j over the run's images;
//   3. every gradient row is produced by (l, x, quad of n1) tiles of 4 x NC
//      register accumulators: for a neighbour of type tj and a present type t,
//      block (min, max)(t, tj) gets Σ_m DC[m][n1] C_t[m][n2] (t > tj),
//      Σ_m C_t[m][n1] DC[m][n2] (t < tj) or both (t == tj). The tiles of one wave
//      leave through its LDS staging area so that consecutive lanes store
//      consecutive 16-byte pieces.
// The center's own rows (self gradients in both factors) follow the windows.
constexpr int TILE_THREADS = 256;
constexpr int TILE_MAX_RUNS = 8;
constexpr int TILE_MAX_TYPES = 32;
constexpr int TILE_MAX_L = 16;

// Per-entry LDS layout: radial values R and dR/dr padded to NC (4-element aligned),
// direction, trigonometric factors, cutoff factors, Legendre functions.
struct TileEntryLayout {
    int L, n, NC, L2, leg;
    int rv, rd, dir, trig, scal, p, per_entry;
    __host__ __device__ TileEntryLayout(int L_, int n_, int NC_)
        : L(L_), n(n_), NC(NC_), L2(L_ * L_), leg(L_ * (L_ + 1) / 2) {
        int o = 0;
        rv = o; o += L * NC;
        rd = o; o += L * NC;
        dir = o; o += 4;    // direction, neighbour type
        trig = o; o += 4;   // cos theta, sin theta, cos phi, sin phi
        scal = o; o += 4;   // f, df/dr, f / r, neighbour atom (record_index)
        p = o; o += leg;
        per_entry = (o + 3) / 4 * 4;
    }
};

// Workgroup barrier that orders LDS only: the kernels using it exchange data
// between waves through LDS alone, so global stores keep draining across it.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Offsets in the per-l / per-(l, m) coefficient table of the spherical harmonics
// (built on the host, every value rounded to T as the serial evaluation rounds it).
struct ShTable {
    int dpa, dpb, sq2l1, diag, gl, il, size;
    __host__ __device__ explicit ShTable(int L) {
        const int leg = L * (L + 1) / 2;
        dpa = 0;
        dpb = leg;
        sq2l1 = 2 * leg;
        diag = sq2l1 + L;
        gl = diag + L;
        il = gl + L;
        size = il + L;
    }
};

template <typename T>
struct Vec4Of;
template <>
struct Vec4Of<float> {
    typedef float type __attribute__((ext_vector_type(4)));
};
template <>
struct Vec4Of<double> {
    typedef double type __attribute__((ext_vector_type(2)));
};

// Load / store 4 consecutive values (16-byte aligned for float, 16-byte pairs for double).
template <typename T>
__device__ __forceinline__ void load4(const T* p, T* v) {
    typedef typename Vec4Of<T>::type V;
    if constexpr (sizeof(T) == 4) {
        const V x = *reinterpret_cast<const V*>(p);
        v[0] = x[0];
        v[1] = x[1];
        v[2] = x[2];
        v[3] = x[3];
    } else {
        const V x = *reinterpret_cast<const V*>(p);
        const V y = *reinterpret_cast<const V*>(p + 2);
        v[0] = x[0];
        v[1] = x[1];
        v[2] = y[0];
        v[3] = y[1];
    }
}

template <typename T>
__device__ __forceinline__ void store4(T* p, T a0, T a1, T a2, T a3) {
    typedef typename Vec4Of<T>::type V;
    if constexpr (sizeof(T) == 4) {
        V x;
        x[0] = a0;
        x[1] = a1;
        x[2] = a2;
        x[3] = a3;
        *reinterpret_cast<V*>(p) = x;
    } else {
        V x, y;
        x[0] = a0;
        x[1] = a1;
        y[0] = a2;
        y[1] = a3;
        *reinterpret_cast<V*>(p) = x;
        *reinterpret_cast<V*>(p + 2) = y;
    }
}

// q = u / d, r = u % d for 0 <= u < 2^22, 1 <= d < 2^12, with inv = 1.0f / d.
__device__ __forceinline__ void small_divmod(int u, int d, float inv, int& q, int& r) {
    q = (int)((float)u * inv);
    r = u - q * d;
    // the float quotient is off by at most one: correct with selects, not branches
    const int up = r >= d ? 1 : 0, down = r < 0 ? 1 : 0;
    q += up - down;
    r += (down - up) * d;
}

// Rows of the 4 x NC register tile of lane-quad q, for RP rows staged per pass:
// accumulator row a = h RP + rr (h = a / RP, rr = a % RP) holds tile row
// RP q + rr + h RP nq, so that in every pass the nq lanes of one tile hold
// RP nq adjacent rows.
template <int NC, int RP>
__device__ __forceinline__ int tile_row(int q, int a) {
    constexpr int nq = NC / 4;
    return RP * q + (a % RP) + (a / RP) * RP * nq;
}

// acc[a][c] += Σ_m R[m][tile_row(q, a)] * Cm[m][c] over m in [l², (l+1)²); R and
// Cm are [L2][NC].
template <typename T, int NC, int RP>
__device__ __forceinline__ void tile_fma(int l, int q, const T* __restrict__ R, const T* __restrict__ Cm, T (&acc)[4][NC]) {
    for (int m = l * l; m < (l + 1) * (l + 1); m++) {
        T rv[4];
        const T* rrow = R + m * NC;
        if constexpr (RP == 4) {
            load4<T>(rrow + 4 * q, rv);
        } else {
#pragma unroll
            for (int a = 0; a < 4; a++) rv[a] = rrow[tile_row<NC, RP>(q, a)];
        }
        T cv[NC];
#pragma unroll
        for (int c = 0; c < NC; c += 4) load4<T>(Cm + m * NC + c, cv + c);
#pragma unroll
        for (int r = 0; r < 4; r++)
#pragma unroll
            for (int c = 0; c < NC; c++) acc[r][c] += rv[r] * cv[c];
    }
}

// acc[a][c] = Σ_m R[m][q + 2a] * Cm[m][c] over m in [l², (l+1)²) for NC = 8, RP = 1
// (tile_fma with the first term assigned rather than added to zeros).
__device__ __forceinline__ void tile_product_f8(int l, int q, const float* __restrict__ R,
                                                const float* __restrict__ Cm, float (&acc)[4][8]) {
    const int m0 = l * l;
    {
        float rv[4], cv[8];
#pragma unroll
        for (int a = 0; a < 4; a++) rv[a] = R[m0 * 8 + q + 2 * a];
        load4<float>(Cm + m0 * 8, cv);
        load4<float>(Cm + m0 * 8 + 4, cv + 4);
#pragma unroll
        for (int a = 0; a < 4; a++)
#pragma unroll
            for (int c = 0; c < 8; c++) acc[a][c] = rv[a] * cv[c];
    }
    for (int m = m0 + 1; m < (l + 1) * (l + 1); m++) {
        float rv[4], cv[8];
#pragma unroll
        for (int a = 0; a < 4; a++) rv[a] = R[m * 8 + q + 2 * a];
        load4<float>(Cm + m * 8, cv);
        load4<float>(Cm + m * 8 + 4, cv + 4);
#pragma unroll
        for (int a = 0; a < 4; a++)
#pragma unroll
            for (int c = 0; c < 8; c++) acc[a][c] += rv[a] * cv[c];
    }
}

// Write the tile rows of lane-quad q (those < n) of an [n][n] tile at `o`, scaled by s.
template <typename T, int NC, int RP>
__device__ __forceinline__ void tile_store(T* __restrict__ o, int q, int n, T s, const T (&acc)[4][NC]) {
#pragma unroll
    for (int a = 0; a < 4; a++) {
        const int row = tile_row<NC, RP>(q, a);
        if (row < n) {
#pragma unroll
            for (int c = 0; c < NC; c++)
                if (c < n) o[row * n + c] = acc[a][c] * s;
        }
    }
}

// Per-wave staged store: every lane holds the tile rows of lane-quad q of an
// [n][n] tile (n == NC) bound for `tile` (or nothing, tile == nullptr). The rows go
// through the wave's LDS staging area, RP per pass (the nq lanes of a tile then
// hold RP nq adjacent rows), and are written back so that consecutive lanes store
// consecutive 16-byte pieces: a store instruction covers whole contiguous runs.
template <typename T, int NC, int RP>
__device__ __forceinline__ void staged_tile_store(T* __restrict__ out, long long tile, int q, T s, const T (&acc)[4][NC],
                                                  T* stage, long long* stage_off) {
    // `tile` is an element offset into `out` (a kernel argument, so the stores are
    // global stores), or -1 for no tile
    constexpr int VEC = 16 / sizeof(T);           // values per 16-byte piece
    constexpr int PER_OWNER = RP * NC / VEC;      // pieces per lane and pass
    constexpr int nq = NC / 4;
    const int lane = threadIdx.x & (WAVE - 1);
    typedef typename Vec4Of<T>::type V;
#pragma unroll
    for (int h = 0; h < 4 / RP; h++) {
        if (tile >= 0) {
#pragma unroll
            for (int r = 0; r < RP; r++)
#pragma unroll
                for (int c = 0; c < NC; c += 4)
                    store4<T>(stage + lane * RP * NC + r * NC + c, acc[h * RP + r][c] * s, acc[h * RP + r][c + 1] * s,
                              acc[h * RP + r][c + 2] * s, acc[h * RP + r][c + 3] * s);
            stage_off[lane] = tile + (RP * q + h * RP * nq) * NC;
        } else {
            stage_off[lane] = -1;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
        long long off[PER_OWNER];
        V val[PER_OWNER];
#pragma unroll
        for (int s2 = 0; s2 < PER_OWNER; s2++) {
            const int p = s2 * WAVE + lane;
            const int owner = p / PER_OWNER, sub = p - owner * PER_OWNER;
            off[s2] = stage_off[owner];
            if (off[s2] >= 0) off[s2] += sub * VEC;
            val[s2] = *reinterpret_cast<const V*>(stage + owner * RP * NC + sub * VEC);
        }
#pragma unroll
        for (int s2 = 0; s2 < PER_OWNER; s2++) {
            if (off[s2] >= 0) *reinterpret_cast<V*>(out + off[s2]) = val[s2];
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
    }
}

// Whole-tile staged store for 8 x 8 FP32 tiles (n = 8, one staged row per pass): every
// Floats per staged 8 x 8 tile in the oriented store: 64 plus a pad of 8, so that the
// four tiles one read instruction covers start 8 banks apart (the transposed reads
// and the row writes then meet at most two-way bank conflicts).
constexpr int ORIENTED_SLOT = 72;
// LDS staging floats per wave of the gradient-row kernel: the oriented store's 8
// padded slots, or one RP x NC row block per lane.
template <typename T, int NC, int RP>
__host__ __device__ constexpr int stage_floats_per_wave() {
    return (sizeof(T) == 4 && NC == 8 && RP == 1) ? 8 * ORIENTED_SLOT : WAVE * RP * NC;
}

// lane holds the rows tile_row(q, 0..3) of X (already scaled) for the output block at
// element offset `tile` (-1: none), and `flags` says which orientations the block
// sums: bit 0 X, bit 1 X^T. In each of four passes 16 lanes (8 tiles) stage their
// whole tiles, then the wave writes those 8 blocks as 16-byte pieces from consecutive
// lanes, each piece read from X, from X^T or from both, so that a product is formed
// once whatever the block's orientation.
__device__ __forceinline__ void staged_tile_store_oriented(float* __restrict__ out, long long tile, int flags, int q,
                                                           const float (&acc)[4][8], float* stage, long long* stage_off) {
    const int lane = threadIdx.x & (WAVE - 1);
    const int slot = (lane >> 1) & 7;
#pragma unroll
    for (int pass = 0; pass < 4; pass++) {
        if ((lane >> 4) == pass && tile >= 0) {
            float* X = stage + slot * ORIENTED_SLOT;
#pragma unroll
            for (int a = 0; a < 4; a++) {
                const int row = q + 2 * a;  // tile_row<8, 1>(q, a)
                store4<float>(X + row * 8, acc[a][0], acc[a][1], acc[a][2], acc[a][3]);
                store4<float>(X + row * 8 + 4, acc[a][4], acc[a][5], acc[a][6], acc[a][7]);
            }
        }
        if ((lane >> 4) == pass && q == 0) stage_off[slot] = tile >= 0 ? (tile | ((long long)flags << 60)) : -1;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
        // both pieces of the lane at once: the offsets and every LDS read are issued
        // before the first use (one wait), then the orientations are selected
        long long off[2];
        float xd[2][4], xt[2][4];
#pragma unroll
        for (int s2 = 0; s2 < 2; s2++) {
            const int p = s2 * WAVE + lane;
            const int u = p >> 4, n1 = (p & 15) >> 1, h = p & 1;
            const float* X = stage + u * ORIENTED_SLOT;
            off[s2] = stage_off[u];
            load4<float>(X + n1 * 8 + 4 * h, xd[s2]);
#pragma unroll
            for (int c = 0; c < 4; c++) xt[s2][c] = X[(4 * h + c) * 8 + n1];
        }
#pragma unroll
        for (int s2 = 0; s2 < 2; s2++) {
            const int p = s2 * WAVE + lane;
            const int n1 = (p & 15) >> 1, h = p & 1;
            if (off[s2] >= 0) {
                const int fl = (int)(off[s2] >> 60);
                float v[4];
#pragma unroll
                for (int c = 0; c < 4; c++) v[c] = ((fl & 1) ? xd[s2][c] : 0.f) + ((fl & 2) ? xt[s2][c] : 0.f);
                store4<float>(out + (off[s2] & ((1ll << 60) - 1)) + n1 * 8 + 4 * h, v[0], v[1], v[2], v[3]);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
    }
}

// Planning pass: one 64-lane workgroup per center. Runs go to runs[e0 + r] =
// {first entry, entry count, neighbour type, -}, their rows (one per present type,
// in present-type order) to run_rows[(e0 + r) * S + k], windows to windows[e0 + w]
// = {first entry, entry count, first run, run count}, the center's own rows to
// self_rows[i * npairs + block], and every row's (sample, system, atom) label to
// labels. Row indices are global over the batch (from row_scan).
template <bool PS>
__global__ void __launch_bounds__(WAVE) ps_grad_plan_kernel(
    int n_atoms, int S, const long long* __restrict__ offsets, const int* __restrict__ entry_j,
    const int* __restrict__ species, const unsigned* __restrict__ type_mask, const long long* __restrict__ sample_scan,
    const long long* __restrict__ row_scan, const long long* __restrict__ block_sample_base,
    const int* __restrict__ atom_system, const int* __restrict__ atom_local, int max_runs, int max_entries,
    int4* __restrict__ runs, int* __restrict__ run_rows, int4* __restrict__ windows, int* __restrict__ n_windows,
    int* __restrict__ self_rows, int* __restrict__ labels, int* __restrict__ split) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int present[TILE_MAX_TYPES];
    // local blocks of a center: (t1 <= t2) type pairs for the power spectrum,
    // neighbour types for the spherical expansion
    const int npairs = S * (S + 1) / 2;
    const int nloc = PS ? npairs : S;
    long long* rowbase = reinterpret_cast<long long*>(smem);
    int* sample_in_block = reinterpret_cast<int*>(rowbase + nloc);
    int* counters = sample_in_block + nloc;
    int* snapshot = counters + nloc;

    const int i = blockIdx.x;
    const int lane = threadIdx.x;
    const long long e0 = offsets[i];
    const int len = (int)(offsets[i + 1] - e0);
    const int ct = species[i];
    const unsigned mask = type_mask[i];
    const int n_present = __popc(mask);
    const int sys = atom_system[i];
    for (int k = lane; k < nloc; k += WAVE) {
        const long long idx = (long long)(ct * nloc + k) * n_atoms + i;
        rowbase[k] = row_scan[idx];
        sample_in_block[k] = (int)(sample_scan[idx] - block_sample_base[ct * nloc + k]);
        counters[k] = 0;
    }
    if (lane == 0) {
        int k = 0;
        for (int t = 0; t < S; t++)
            if ((mask >> t) & 1u) present[k++] = t;
    }
    __syncthreads();

    // uniform planning state
    bool snapped = false, split_run = false;
    int n_runs = 0, n_win = 0;
    int cur_start = 0, cur_entries = 0, cur_run_start = 0, cur_runs = 0;
    auto emit = [&](int start, int entries, int run_start, int count) {
        if (lane == 0) windows[e0 + n_win] = make_int4(start, entries, run_start, count);
        n_win++;
    };
    // tj, aj: the neighbour's type and its index in its system, loaded with the entries
    auto process_run = [&](int first, int count, int j, int tj, int aj) {
        if (j == i) {
            // periodic images of the center: no row; close the current window around them
            if (cur_runs > 0) emit(cur_start, cur_entries, cur_run_start, cur_runs);
            cur_runs = 0;
            cur_entries = 0;
            return;
        }
        const int r = n_runs++;
        if (lane == 0) runs[e0 + r] = make_int4(first, count, tj, 0);
        if (!snapped && j > i) {
            for (int k = lane; k < nloc; k += WAVE) snapshot[k] = counters[k];
            snapped = true;
        }
        __builtin_amdgcn_wave_barrier();
        if (!PS) {
            // one row, in block (center type, tj)
            if (lane == 0) {
                const long long row = rowbase[tj] + counters[tj] + (i < j ? 1 : 0);
                counters[tj] += 1;
                run_rows[(e0 + r) * S] = (int)row;
                labels[3 * row] = sample_in_block[tj];
                labels[3 * row + 1] = sys;
                labels[3 * row + 2] = aj;
            }
        } else if (lane < n_present) {
            const int t = present[lane];
            const int lo = t < tj ? t : tj, hi = t < tj ? tj : t;
            const int local = pair_index(lo, hi, S);
            const long long row = rowbase[local] + counters[local] + (i < j ? 1 : 0);
            counters[local] += 1;
            run_rows[(e0 + r) * S + lane] = (int)row;
            labels[3 * row] = sample_in_block[local];
            labels[3 * row + 1] = sys;
            labels[3 * row + 2] = aj;
        }
        __builtin_amdgcn_wave_barrier();
        // windows: up to max_runs runs and max_entries entries; a longer run is split
        int pos = first, rem = count;
        for (;;) {
            if (cur_runs > 0 && (cur_runs == max_runs || cur_entries + rem > max_entries)) {
                emit(cur_start, cur_entries, cur_run_start, cur_runs);
                cur_runs = 0;
                cur_entries = 0;
            }
            if (cur_runs == 0) {
                cur_start = pos;
                cur_run_start = r;
                if (rem > max_entries) {
                    split_run = true;
                    emit(pos, max_entries, r, 1);
                    pos += max_entries;
                    rem -= max_entries;
                    continue;
                }
            }
            cur_runs++;
            cur_entries += rem;
            break;
        }
    };

    int open_first = -1, open_j = -1, open_tj = 0, open_aj = 0;
    int prev_last = -1;
    for (int base = 0; base < len; base += WAVE) {
        const int e = base + lane;
        const int j = e < len ? entry_j[e0 + e] : -1;
        const int tjl = e < len ? species[j] : 0;
        const int ajl = e < len ? atom_local[j] : 0;
        int prev = __shfl_up(j, 1);
        if (lane == 0) prev = prev_last;
        const bool start = e < len && (e == 0 || j != prev);
        unsigned long long bits = __ballot(start);
        prev_last = __shfl(j, WAVE - 1);
        while (bits != 0ull) {
            const int b = __ffsll((long long)bits) - 1;
            bits &= bits - 1ull;
            const int jb = __shfl(j, b), tjb = __shfl(tjl, b), ajb = __shfl(ajl, b);
            if (open_first >= 0) process_run(open_first, base + b - open_first, open_j, open_tj, open_aj);
            open_first = base + b;
            open_j = jb;
            open_tj = tjb;
            open_aj = ajb;
        }
    }
    if (open_first >= 0) process_run(open_first, len - open_first, open_j, open_tj, open_aj);
    if (cur_runs > 0) emit(cur_start, cur_entries, cur_run_start, cur_runs);
    if (lane == 0) {
        n_windows[i] = n_win;
        if (split) split[i] = split_run ? 1 : 0;
    }

    // the center's own rows: before its first neighbour j > i in every block
    if (!snapped) {
        for (int k = lane; k < nloc; k += WAVE) snapshot[k] = counters[k];
    }
    __builtin_amdgcn_wave_barrier();
    if (!PS) {
        for (int p = lane; p < n_present; p += WAVE) {
            const int nt = present[p];
            const long long row = rowbase[nt] + snapshot[nt];
            self_rows[(long long)i * S + nt] = (int)row;
            labels[3 * row] = sample_in_block[nt];
            labels[3 * row + 1] = sys;
            labels[3 * row + 2] = atom_local[i];
        }
        return;
    }
    for (int p = lane; p < n_present * n_present; p += WAVE) {
        const int k1 = p / n_present, k2 = p - k1 * n_present;
        if (k2 < k1) continue;
        const int local = pair_index(present[k1], present[k2], S);
        const long long row = rowbase[local] + snapshot[local];
        self_rows[(long long)i * npairs + local] = (int)row;
        labels[3 * row] = sample_in_block[local];
        labels[3 * row + 1] = sys;
        labels[3 * row + 2] = atom_local[i];
    }
}

// radial_integral for n = 8, with the interval's knots and every table value loaded
// together (16-byte pieces) at the interval of the uniform-grid estimate (from the
// per-channel grid table), and loaded again only when the stored knots move the
// interval. The interval is that of radial_integral; the derivative's 1 / delta is one
// reciprocal (within an ulp of radial_integral's divisions).
template <typename T>
__device__ void radial_integral_n8(const DeviceRadial<T>& rad, int l, long long e, double r, T* values, T* grads) {
    T a0[8], b0[8], a1[8], b1[8];
    if (rad.use_spline) {
        const int K = rad.n_knots[l] - 1;
        const int off = rad.knot_offset[l];
        const T* pos = rad.knot_pos + off;
        int k = (int)((r - rad.knot_grid[2 * l]) * rad.knot_grid[2 * l + 1]);
        if (k < 0) k = 0;
        if (k > K - 1) k = K - 1;
        T x0, x1;
        auto load = [&]() {
            x0 = pos[k];
            x1 = pos[k + 1];
            const T* v0 = rad.knot_val + (long long)(off + k) * 8;
            const T* m0 = rad.knot_der + (long long)(off + k) * 8;
            load4<T>(v0, a0);
            load4<T>(v0 + 4, a0 + 4);
            load4<T>(v0 + 8, a1);
            load4<T>(v0 + 12, a1 + 4);
            load4<T>(m0, b0);
            load4<T>(m0 + 4, b0 + 4);
            load4<T>(m0 + 8, b1);
            load4<T>(m0 + 12, b1 + 4);
        };
        load();
        if ((k > 0 && (T)r < x0) || (k < K - 1 && (T)r >= x1)) {
            while (k > 0 && (T)r < pos[k]) k--;
            while (k < K - 1 && (T)r >= pos[k + 1]) k++;
            load();
        }
        const T delta = x1 - x0;
        const T t = ((T)r - x0) / delta, t2 = t * t, t3 = t2 * t;
        const T h00 = 2 * t3 - 3 * t2 + 1, h10 = t3 - 2 * t2 + t, h01 = -2 * t3 + 3 * t2, h11 = t3 - t2;
        const T d00 = 6 * (t2 - t), d10 = 3 * t2 - 4 * t + 1, d01 = -d00, d11 = 3 * t2 - 2 * t;
        const T inv_delta = (T)1 / delta;
        T v[8], g[8];
#pragma unroll
        for (int i = 0; i < 8; i++) {
            v[i] = h00 * a0[i] + h10 * delta * b0[i] + h01 * a1[i] + h11 * delta * b1[i];
            g[i] = d00 * a0[i] * inv_delta + d10 * b0[i] + d01 * a1[i] * inv_delta + d11 * b1[i];
        }
        store4<T>(values, v[0], v[1], v[2], v[3]);
        store4<T>(values + 4, v[4], v[5], v[6], v[7]);
        store4<T>(grads, g[0], g[1], g[2], g[3]);
        store4<T>(grads + 4, g[4], g[5], g[6], g[7]);
    } else {
        const long long base = (e * (rad.max_angular + 1) + l) * 8;
        load4<T>(rad.exact_val + base, a0);
        load4<T>(rad.exact_val + base + 4, a0 + 4);
        load4<T>(rad.exact_der + base, b0);
        load4<T>(rad.exact_der + base + 4, b0 + 4);
        store4<T>(values, a0[0], a0[1], a0[2], a0[3]);
        store4<T>(values + 4, a0[4], a0[5], a0[6], a0[7]);
        store4<T>(grads, b0[0], b0[1], b0[2], b0[3]);
        store4<T>(grads + 4, b0[4], b0[5], b0[6], b0[7]);
    }
}

// an atom index in a record slot: its bit pattern in FP32, its value in FP64
template <typename T>
__device__ __forceinline__ T record_index(int j);
template <>
__device__ __forceinline__ float record_index<float>(int j) {
    return __int_as_float(j);
}
template <>
__device__ __forceinline__ double record_index<double>(int j) {
    return (double)j;
}

// Per-entry data of the tiled gradient kernel, one (entry, slot) per thread over
// the whole batch. Slots [0, L): m = slot, the associated Legendre functions P_lm
// for l >= m (the recursion of featomic/src/math/spherical_harmonics.rs:228-251
// run along l for this m, the diagonal chain recomputed per slot); slot 0 also
// stores the direction, trigonometric factors and the cutoff / scaling factors.
// Slots [L, 2L): l = slot - L, the radial spline R_nl and dR_nl/dr, zero-padded
// to NC.
template <typename T>
__device__ void entry_prep_slot(const SoapArgs<T>& a, const TileEntryLayout& lay, long long e, int slot, int NC, T* base,
                                const T* sha, const T* shb, const T* shc);

// A workgroup fills the records of `per_block` consecutive entries in LDS, then
// writes them out as one contiguous block of 16-byte stores (the slots' own
// writes are 4- or 8-byte pieces spread over each record).
template <typename T>
__global__ void __launch_bounds__(256) entry_prep_kernel(SoapArgs<T> a, long long n_entries, int NC, int per_block,
                                                         T* __restrict__ data) {
    extern __shared__ __align__(16) unsigned char smem[];
    T* records = reinterpret_cast<T*>(smem);
    const TileEntryLayout lay(a.L, a.n, NC);
    const int L = lay.L;
    const long long e_first = (long long)blockIdx.x * per_block;
    const int count = (int)min((long long)per_block, n_entries - e_first);
    const int tid = threadIdx.x;
    // the Legendre recursion's coefficients, in LDS for the recursion's loops
    __shared__ T sha[TILE_MAX_L * (TILE_MAX_L + 1) / 2 + 1];
    __shared__ T shb[TILE_MAX_L * (TILE_MAX_L + 1) / 2 + 1];
    __shared__ T shc[TILE_MAX_L * (TILE_MAX_L + 1) + 4 * TILE_MAX_L];
    {
        const ShTable tb(L);
        for (int k = tid; k <= lay.leg; k += blockDim.x) {
            sha[k] = a.sh_a[k];
            shb[k] = a.sh_b[k];
        }
        for (int k = tid; k < tb.size; k += blockDim.x) shc[k] = a.sh_c[k];
    }
    __syncthreads();
    // threads 0-127 take the Legendre slots (entry, m), threads 128-255 the radial slots
    // (entry, l): each wave runs one of the two paths (per_block * L <= 128)
    const int part = tid < 128 ? 0 : 1;
    const int u = tid - 128 * part;
    if (u < count * L) {
        const int local = u / L;
        const int slot = u - local * L + part * L;
        const long long e = e_first + local;
        T* base = records + (size_t)local * lay.per_entry;
        entry_prep_slot<T>(a, lay, e, slot, NC, base, sha, shb, shc);
    }
    __syncthreads();
    typedef typename Vec4Of<T>::type V;
    constexpr int VEC = 16 / sizeof(T);
    const int nvec = count * lay.per_entry / VEC;
    const V* src = reinterpret_cast<const V*>(records);
    V* dst = reinterpret_cast<V*>(data + e_first * lay.per_entry);
    for (int k = tid; k < nvec; k += blockDim.x) dst[k] = src[k];
}

// One slot of an entry's record (see entry_prep_kernel), written at `base`.
template <typename T>
__device__ void entry_prep_slot(const SoapArgs<T>& a, const TileEntryLayout& lay, long long e, int slot, int NC, T* base,
                                const T* sha, const T* shb, const T* shc) {
    const int L = lay.L;
    if (slot < L) {
        const ShTable tb(L);
        const int m = slot;
        const double r = a.entry_dist[e];
        double d[3];
        if (r < 1e-6) {
            d[0] = 0.0;
            d[1] = 0.0;
            d[2] = 1.0;
        } else {
            d[0] = a.entry_vec[3 * e] / r;
            d[1] = a.entry_vec[3 * e + 1] / r;
            d[2] = a.entry_vec[3 * e + 2] / r;
        }
        const T x = (T)d[0], y = (T)d[1], z = (T)d[2];
        const T sqrt_xy = (T)hypot((double)x, (double)y);
        const T cos_theta = z, sin_theta = sqrt_xy;
        if (m == 0) {
            T cos_phi = 1, sin_phi = 0;
            if ((double)sqrt_xy > (sizeof(T) == 8 ? 2.220446049250313e-16 : 1.1920929e-07)) {
                cos_phi = x / sqrt_xy;
                sin_phi = y / sqrt_xy;
            }
            base[lay.dir + 0] = x;
            base[lay.dir + 1] = y;
            base[lay.dir + 2] = z;
            // the neighbour's type, and below its atom index (bit pattern in FP32), for
            // the kernels that group entries by type and skip a center's own images
            const int j = a.entry_j[e];
            base[lay.dir + 3] = (T)a.species[j];
            base[lay.trig + 0] = cos_theta;
            base[lay.trig + 1] = sin_theta;
            base[lay.trig + 2] = cos_phi;
            base[lay.trig + 3] = sin_phi;
            double f, df;
            scaling_functions(a.scalars, r, &f, &df);
            base[lay.scal + 0] = (T)f;
            base[lay.scal + 1] = (T)df;
            base[lay.scal + 2] = (T)f / (T)r;
            base[lay.scal + 3] = record_index<T>(j);
            for (int k = lay.p + lay.leg; k < lay.per_entry; k++) base[k] = 0;
        }
        T value = (T)0.3989422804014327;
        for (int l = 1; l <= m; l++) value *= -shc[tb.diag + l] * sin_theta;
        T* P = base + lay.p;
        P[leg_index(m, m)] = value;
        if (m + 1 < L) {
            T p2 = value;
            T p1 = cos_theta * shc[tb.sq2l1 + m + 1] * value;
            P[leg_index(m + 1, m)] = p1;
            for (int l = m + 2; l < L; l++) {
                const int k = leg_index(l, m);
                const T p0 = sha[k] * (cos_theta * p1 + shb[k] * p2);
                P[k] = p0;
                p2 = p1;
                p1 = p0;
            }
        }
    } else {
        const int l = slot - L;
        const int n = lay.n;
        T* va = base + lay.rv + l * NC;
        T* vb = base + lay.rd + l * NC;
        if (n == 8) {
            radial_integral_n8<T>(a.radial, l, e, a.entry_dist[e], va, vb);
        } else {
            radial_integral<T>(a.radial, l, e, a.entry_dist[e], va, vb);
        }
        for (int k = n; k < NC; k++) {
            va[k] = 0;
            vb[k] = 0;
        }
    }
}

// Stage 2: one item per (x, run, l, m >= 0) forms Y_l(±m) and r dY_l(±m)/dx of
// every image of the run (spherical_harmonics.rs:256-298, :304-424) and sums
// DC_run[x][l² + l ± m][:] += (df R + f dR) Y d_x + (f / r) R r dY/dx
// (soap/spherical_expansion_pair.rs:382-419). A run carried from the previous
// window (slot 0) continues its sums.
// `shc` is the spherical-harmonics coefficient table (ShTable layout), in global
// memory or an LDS copy.
template <typename T, int NC>
__device__ void plan_entry_stage2(const T* shc, const TileEntryLayout& lay, int nr, const int* run_lo,
                                  const int* run_hi, bool carry_in, const int* leg_l, const int* leg_m, const T* ent,
                                  T* DC, int n_threads = TILE_THREADS) {
    const int leg = lay.leg, L2 = lay.L2;
    const int CS = L2 * NC;
    const ShTable tb(lay.L);
    const T SQRT_2 = (T)1.4142135623730951;
    const int per_x = nr * leg;
    for (int item = threadIdx.x; item < 3 * per_x; item += n_threads) {
        const int x = item / per_x;
        const int rest = item - x * per_x;
        const int r = rest / leg, k = rest - r * leg;
        const int l = leg_l[k], m = leg_m[k];
        const int lm0 = l * l + l;
        T* dp_ = DC + ((size_t)r * 3 + x) * CS + (lm0 + m) * NC;
        T* dm_ = DC + ((size_t)r * 3 + x) * CS + (lm0 - m) * NC;
        T accp[NC], accm[NC];
        if (r == 0 && carry_in) {
#pragma unroll
            for (int c = 0; c < NC; c += 4) {
                load4<T>(dp_ + c, accp + c);
                load4<T>(dm_ + c, accm + c);
            }
        } else {
#pragma unroll
            for (int c = 0; c < NC; c++) {
                accp[c] = 0;
                accm[c] = 0;
            }
        }
        const T c_gl = shc[tb.gl + l], c_dpa = shc[tb.dpa + k], c_dpb = shc[tb.dpb + k];
        for (int e = run_lo[r]; e < run_hi[r]; e++) {
            const T* base = ent + e * lay.per_entry;
            const T* P = base + lay.p;
            const T cos_theta = base[lay.trig + 0], sin_theta = base[lay.trig + 1];
            const T cos_phi = base[lay.trig + 2], sin_phi = base[lay.trig + 3];
            const T dx = base[lay.dir + x];
            T yp, ym = 0, gp, gm = 0;
            if (m == 0) {
                yp = P[k] / SQRT_2;
                if (l == 0) {
                    gp = 0;
                } else {
                    const T lf = c_gl * P[leg_index(l, 1)];
                    gp = x == 0 ? cos_phi * cos_theta * lf : (x == 1 ? sin_phi * cos_theta * lf : -sin_theta * lf);
                }
            } else {
                T cos_1 = 1, sin_1 = 0, cos_2 = -cos_phi, sin_2 = sin_phi;
                const T minus_two_cos = -2 * cos_phi;
                T sin_m = 0, cos_m = 1;
                for (int mm = 1; mm <= m; mm++) {
                    sin_m = minus_two_cos * sin_1 - sin_2;
                    cos_m = minus_two_cos * cos_1 - cos_2;
                    sin_2 = sin_1;
                    sin_1 = sin_m;
                    cos_2 = cos_1;
                    cos_1 = cos_m;
                }
                const T p = P[k];
                yp = p * cos_m;
                ym = p * sin_m;
                const T pm1 = P[leg_index(l, m - 1)];
                const T pp1 = m < l ? P[leg_index(l, m + 1)] : (T)0;
                const T dp = c_dpa * pm1 - c_dpb * pp1;
                const T pt = (sin_theta > (T)0.1) ? (T)m / sin_theta * p : -(T)0.5 / cos_theta * dp;
                const T s_dp = sin_m * dp, c_dp = cos_m * dp;
                if (x == 0) {
                    gp = sin_phi * pt * sin_m - (T)0.5 * cos_theta * cos_phi * c_dp;
                    gm = -sin_phi * pt * cos_m - (T)0.5 * cos_theta * cos_phi * s_dp;
                } else if (x == 1) {
                    gp = -cos_phi * pt * sin_m - (T)0.5 * cos_theta * sin_phi * c_dp;
                    gm = cos_phi * pt * cos_m - (T)0.5 * cos_theta * sin_phi * s_dp;
                } else {
                    gp = (T)0.5 * sin_theta * c_dp;
                    gm = (T)0.5 * sin_theta * s_dp;
                }
            }
            const T f = base[lay.scal + 0], df = base[lay.scal + 1], f_r = base[lay.scal + 2];
            const T ypd = yp * dx, ymd = ym * dx;
            T Rv[NC], dR[NC];
#pragma unroll
            for (int c = 0; c < NC; c += 4) {
                load4<T>(base + lay.rv + l * NC + c, Rv + c);
                load4<T>(base + lay.rd + l * NC + c, dR + c);
            }
#pragma unroll
            for (int c = 0; c < NC; c++) {
                const T ra = df * Rv[c] + f * dR[c];
                const T rb = f_r * Rv[c];
                accp[c] += ra * ypd + rb * gp;
                accm[c] += ra * ymd + rb * gm;
            }
        }
#pragma unroll
        for (int c = 0; c < NC; c += 4) store4<T>(dp_ + c, accp[c], accp[c + 1], accp[c + 2], accp[c + 3]);
        if (m > 0) {
#pragma unroll
            for (int c = 0; c < NC; c += 4) store4<T>(dm_ + c, accm[c], accm[c + 1], accm[c + 2], accm[c + 3]);
        }
    }
}

// Split x into bfloat16 hi + lo (round to nearest even) so that hi + lo carries
// ~16 significant bits of x; products are then formed as hi*hi + hi*lo + lo*hi.
__device__ __forceinline__ void bf16_split(float x, short& hi, short& lo) {
    unsigned u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    hi = (short)(u >> 16);
    const float r = x - __uint_as_float(((unsigned)(unsigned short)hi) << 16);
    unsigned v = __float_as_uint(r);
    v += 0x7FFFu + ((v >> 16) & 1u);
    lo = (short)(v >> 16);
}

typedef short short4_t __attribute__((ext_vector_type(4)));
typedef short short8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float float4_t __attribute__((ext_vector_type(4)));

// Gradient rows of the complete runs of one window (the contraction phase of the
// tiled kernels), run by the workgroup's TILE_THREADS computing threads with no
// workgroup barrier: per run r (neighbour type run_tj[r], rows run_row[r][k] per
// present type k) and x, l, the rows s_l (DC_r^T C_t, C_t^T DC_r or both) of
// soap/power_spectrum.rs:621-678, written through the wave's staging area `stage`.
template <typename T, int NC, int RP, bool FAST>
__device__ __forceinline__ void ps_window_rows(int L, int n, int n_present, const int* present, int nc_runs,
                                               const int* run_tj, const int (*run_row)[TILE_MAX_TYPES],
                                               const T* C, const T* DC, const T* scale_l, T* stage,
                                               long long* stage_offs, T* __restrict__ out) {
    const int tid = threadIdx.x;
    const int wave_id = tid / WAVE;
    const int L2 = L * L;
    const int CS = L2 * NC;
    const int P = L * n * n;
    constexpr int nq = NC / 4;
    const bool staged = n == NC;
    if constexpr (FAST && sizeof(T) == 4 && NC == 8) {
        // mixed precision: X[r][g][n1][n2] = Σ_m DC_r[x][m][n1] C_{t_g}[m][n2] for up to 4
        // runs x 2 types per pair of 16 x 16 tiles, bfloat16 hi + lo products on MFMA with
        // FP32 accumulation (3 MFMAs per k block of 16); the wave stages X in LDS and
        // writes each 8 x 8 block in the orientation of its output block
        const int lane = tid & (WAVE - 1);
        const int ngroups = (n_present + 1) / 2;
        const int items = 3 * L * ngroups;
        T* X = stage;  // [4 runs][2 types][8][8]
        for (int it = wave_id; it < items; it += TILE_THREADS / WAVE) {
            const int g = it % ngroups, lx = it / ngroups;
            const int l = L - 1 - lx / 3, x = lx % 3;
            const int g0 = 2 * g, ng = min(2, n_present - g0);
            const int kcount = 2 * l + 1;
            const int i16 = lane & 15, kq = 4 * (lane >> 4);
            for (int tr = 0; tr < (nc_runs + 1) / 2; tr++) {
                float4_t acc = {0.f, 0.f, 0.f, 0.f};
                const int r = 2 * tr + (i16 >> 3), n1 = i16 & 7;
                const int tg = i16 >> 3, n2 = i16 & 7;
                for (int kb = 0; kb < kcount; kb += 16) {
                    short4_t ah, al, bh, bl;
#pragma unroll
                    for (int e = 0; e < 4; e++) {
                        const int mi = kb + kq + e;
                        float av = 0.f, bv = 0.f;
                        if (mi < kcount) {
                            if (r < nc_runs) av = DC[((size_t)r * 3 + x) * CS + (l * l + mi) * NC + n1];
                            if (tg < ng) bv = C[(size_t)present[g0 + tg] * CS + (l * l + mi) * NC + n2];
                        }
                        short h, lo;
                        bf16_split(av, h, lo);
                        ah[e] = h;
                        al[e] = lo;
                        bf16_split(bv, h, lo);
                        bh[e] = h;
                        bl[e] = lo;
                    }
                    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bh, acc, 0, 0, 0);
                    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bl, acc, 0, 0, 0);
                    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(al, bh, acc, 0, 0, 0);
                }
                // D[row][col]: row = 4 (lane / 16) + v within the tile, col = lane % 16
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const int rho = 16 * tr + kq + v;
                    X[(((rho >> 3) * 2 + tg) * 8 + (rho & 7)) * 8 + n2] = acc[v];
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
            const int npieces = nc_runs * ng * 16;
            for (int p = lane; p < npieces; p += WAVE) {
                const int r = p / (ng * 16), rem = p - r * ng * 16;
                const int gg = rem >> 4, n1 = (rem & 15) >> 1, h = rem & 1;
                const int t = present[g0 + gg], tj = run_tj[r];
                const T* Xb = X + (r * 2 + gg) * 64;
                T v[4];
#pragma unroll
                for (int c = 0; c < 4; c++) {
                    const int n2 = 4 * h + c;
                    T sum = 0;
                    if (t >= tj) sum += Xb[n1 * 8 + n2];
                    if (t <= tj) sum += Xb[n2 * 8 + n1];
                    v[c] = sum;
                }
                const T sc = t != tj ? scale_l[l] * (T)1.4142135623730951 : scale_l[l];
                T* dst = out + (long long)run_row[r][g0 + gg] * 3 * P + (long long)x * P + l * 64 + n1 * 8 + 4 * h;
                store4<T>(dst, v[0] * sc, v[1] * sc, v[2] * sc, v[3] * sc);
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
        }
        return;
    } else if constexpr (FAST && sizeof(T) == 4 && (NC == 12 || NC == 16)) {
        // mixed precision for n = 12 or 16: one 16 x 16 tile per (run, type),
        // X[n1][n2] = Σ_m DC_r[x][m][n1] C_t[m][n2] in bfloat16 hi + lo on MFMA. The
        // type's operand (B) is split once per (l, x, type) item; the wave stages up
        // to RP * NC / 4 tiles, then writes their n x n blocks as 16-byte pieces
        constexpr int STAGE_TILES = RP * NC / 4;
        constexpr int KB_MAX = (2 * TILE_MAX_L - 1 + 15) / 16;
        constexpr int nq4 = NC / 4;
        const int lane = tid & (WAVE - 1);
        const int i16 = lane & 15, kq = 4 * (lane >> 4);
        const int items = 3 * L * n_present;
        T* X = stage;  // [STAGE_TILES][16][16]
        for (int it = wave_id; it < items; it += TILE_THREADS / WAVE) {
            const int g = it % n_present, lx = it / n_present;
            const int l = L - 1 - lx / 3, x = lx % 3;
            const int kcount = 2 * l + 1;
            const int t = present[g];
            const T* Cb = C + (size_t)t * CS + l * l * NC;
            short4_t bh[KB_MAX], bl[KB_MAX];
#pragma unroll
            for (int kb = 0; kb < KB_MAX; kb++) {
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    const int mi = 16 * kb + kq + e;
                    const float bv = (mi < kcount && i16 < NC) ? Cb[mi * NC + i16] : 0.f;
                    short h, lo;
                    bf16_split(bv, h, lo);
                    bh[kb][e] = h;
                    bl[kb][e] = lo;
                }
            }
            const T sc_same = scale_l[l], sc_diff = scale_l[l] * (T)1.4142135623730951;
            for (int r0 = 0; r0 < nc_runs; r0 += STAGE_TILES) {
                const int nr = min(STAGE_TILES, nc_runs - r0);
                for (int rr = 0; rr < nr; rr++) {
                    const T* D = DC + ((size_t)(r0 + rr) * 3 + x) * CS + l * l * NC;
                    float4_t acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                    for (int kb = 0; kb < KB_MAX; kb++) {
                        if (16 * kb >= kcount) break;
                        short4_t ah, al;
#pragma unroll
                        for (int e = 0; e < 4; e++) {
                            const int mi = 16 * kb + kq + e;
                            const float av = (mi < kcount && i16 < NC) ? D[mi * NC + i16] : 0.f;
                            short h, lo;
                            bf16_split(av, h, lo);
                            ah[e] = h;
                            al[e] = lo;
                        }
                        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bh[kb], acc, 0, 0, 0);
                        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bl[kb], acc, 0, 0, 0);
                        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(al, bh[kb], acc, 0, 0, 0);
                    }
                    // D[row][col]: row = 4 (lane / 16) + v, col = lane % 16
#pragma unroll
                    for (int v = 0; v < 4; v++) X[rr * 256 + (kq + v) * 16 + i16] = acc[v];
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
                const int npieces = nr * NC * nq4;
                for (int p = lane; p < npieces; p += WAVE) {
                    const int rr = p / (NC * nq4), rem = p - rr * (NC * nq4);
                    const int n1 = rem / nq4, h = rem - n1 * nq4;
                    const int r = r0 + rr, tj = run_tj[r];
                    const T* Xb = X + rr * 256;
                    T v[4];
#pragma unroll
                    for (int c = 0; c < 4; c++) {
                        const int n2 = 4 * h + c;
                        T sum = 0;
                        if (t >= tj) sum += Xb[n1 * 16 + n2];
                        if (t <= tj) sum += Xb[n2 * 16 + n1];
                        v[c] = sum;
                    }
                    const T sc = t != tj ? sc_diff : sc_same;
                    T* dst = out + (long long)run_row[r][g] * 3 * P + (long long)x * P + l * NC * NC + n1 * NC + 4 * h;
                    store4<T>(dst, v[0] * sc, v[1] * sc, v[2] * sc, v[3] * sc);
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
            }
        }
        return;
    }
    if constexpr (!FAST && sizeof(T) == 4 && NC == 8 && RP == 1) {
        if (n == NC) {
            // FP32, n = 8: one product X = DC_r^T C_t per (run, type, x, l) tile, for every
            // lane of the wave alike; the block's orientation is chosen when it is written
            const int NRK = nc_runs * n_present;
            const int per_lx = NRK * nq;
            const int total = per_lx * 3 * L;
            const float inv_lx = 1.0f / (float)(per_lx > 0 ? per_lx : 1);
            const float inv_np = 1.0f / (float)(n_present > 0 ? n_present : 1);
            for (int ubase = wave_id * WAVE; ubase < total; ubase += TILE_THREADS) {
                const int u = ubase + (tid & (WAVE - 1));
                long long tile = -1;
                int q = 0, flags = 0;
                float acc[4][8];  // written only for u < total; staged only then
                if (u < total) {
                    int lx, wq, r, k;
                    small_divmod(u, per_lx, inv_lx, lx, wq);
                    const int rk = wq / nq;
                    q = wq - rk * nq;
                    small_divmod(rk, n_present, inv_np, r, k);
                    const int lrev = lx / 3;
                    const int x = lx - 3 * lrev;
                    const int l = L - 1 - lrev;
                    const int t = present[k], tj = run_tj[r];
                    tile_product_f8(l, q, reinterpret_cast<const float*>(DC) + ((size_t)r * 3 + x) * CS,
                                    reinterpret_cast<const float*>(C) + (size_t)t * CS, acc);
                    const float sc = t != tj ? (float)scale_l[l] * 1.4142135623730951f : (float)scale_l[l];
#pragma unroll
                    for (int rr = 0; rr < 4; rr++)
#pragma unroll
                        for (int c = 0; c < 8; c++) acc[rr][c] *= sc;
                    flags = (t >= tj ? 1 : 0) | (t <= tj ? 2 : 0);
                    tile = (long long)run_row[r][k] * 3 * P + (long long)x * P + l * 64;
                }
                staged_tile_store_oriented(reinterpret_cast<float*>(out), tile, flags, q, acc,
                                           reinterpret_cast<float*>(stage), stage_offs);
            }
            return;
        }
    }
    // gradient rows of the complete runs: units (l descending, x) x (run, type) x quad,
    // so that the nq lanes of one (run, type, x, l) tile are adjacent
    const int NRK = nc_runs * n_present;
    const int per_lx = NRK * nq;
    const int total = per_lx * 3 * L;
    const float inv_lx = 1.0f / (float)(per_lx > 0 ? per_lx : 1);
    const float inv_np = 1.0f / (float)(n_present > 0 ? n_present : 1);
    for (int ubase = wave_id * WAVE; ubase < total; ubase += TILE_THREADS) {
        const int u = ubase + (tid & (WAVE - 1));
        long long tile = -1;
        int q = 0;
        T s = 0;
        T acc[4][NC];
#pragma unroll
        for (int rr = 0; rr < 4; rr++)
#pragma unroll
            for (int c = 0; c < NC; c++) acc[rr][c] = 0;
        if (u < total) {
            int lx, wq, rk, r, k;
            small_divmod(u, per_lx, inv_lx, lx, wq);
            rk = wq / nq;
            q = wq - rk * nq;
            small_divmod(rk, n_present, inv_np, r, k);
            const int lrev = lx / 3;
            const int x = lx - 3 * lrev;
            const int l = L - 1 - lrev;
            const int t = present[k], tj = run_tj[r];
            const T* D = DC + ((size_t)r * 3 + x) * CS;
            const T* Ct = C + (size_t)t * CS;
            if (t >= tj) tile_fma<T, NC, RP>(l, q, D, Ct, acc);
            if (t <= tj) tile_fma<T, NC, RP>(l, q, Ct, D, acc);
            s = t != tj ? scale_l[l] * (T)1.4142135623730951 : scale_l[l];
            tile = (long long)run_row[r][k] * 3 * P + (long long)x * P + l * n * n;
        }
        if (staged) {
            staged_tile_store<T, NC, RP>(out, tile, q, s, acc, stage, stage_offs);
        } else if (tile >= 0) {
            tile_store<T, NC, RP>(out + tile, q, n, s, acc);
        }
    }
}

// The center's own gradient rows (atom == center) of every block (t1 <= t2) present
// around center i, from its self gradients DS [n_present][3][L2][NC] (LDS) and its
// coefficients C: s_l (DS_{t1}^T C_{t2} + C_{t1}^T DS_{t2}) (soap/power_spectrum.rs:621-678),
// run by the TILE_THREADS computing threads with no workgroup barrier.
template <typename T, int NC, int RP>
__device__ __forceinline__ void ps_self_rows(int i, int L, int n, int S, int n_present, const int* present,
                                             const int* __restrict__ self_rows, const T* C, const T* DS,
                                             const T* scale_l, T* stage, long long* stage_offs, T* __restrict__ out) {
    const int tid = threadIdx.x;
    const int wave_id = tid / WAVE;
    const int CS = L * L * NC;
    const int P = L * n * n;
    const int npairs = S * (S + 1) / 2;
    constexpr int nq = NC / 4;
    const bool staged = n == NC;
    const int npp = n_present * (n_present + 1) / 2;
    const int per_lx = npp * nq;
    const int total = per_lx * 3 * L;
    const float inv_lx = 1.0f / (float)(per_lx > 0 ? per_lx : 1);
    for (int ubase = wave_id * WAVE; ubase < total; ubase += TILE_THREADS) {
        const int u = ubase + (tid & (WAVE - 1));
        long long tile = -1;
        int q = 0;
        T s = 0;
        T acc[4][NC];
#pragma unroll
        for (int rr = 0; rr < 4; rr++)
#pragma unroll
            for (int c = 0; c < NC; c++) acc[rr][c] = 0;
        if (u < total) {
            int lx, wq;
            small_divmod(u, per_lx, inv_lx, lx, wq);
            const int p = wq / nq;
            q = wq - p * nq;
            const int lrev = lx / 3;
            const int x = lx - 3 * lrev;
            const int l = L - 1 - lrev;
            // p -> (k1 <= k2) in row-major order of the upper triangle
            int k1 = 0, rem = p;
            while (rem >= n_present - k1) {
                rem -= n_present - k1;
                k1++;
            }
            const int k2 = k1 + rem;
            const int t1 = present[k1], t2 = present[k2];
            const int local = pair_index(t1, t2, S);
            const long long row = self_rows[(long long)i * npairs + local];
            tile_fma<T, NC, RP>(l, q, DS + ((size_t)k1 * 3 + x) * CS, C + (size_t)t2 * CS, acc);
            tile_fma<T, NC, RP>(l, q, C + (size_t)t1 * CS, DS + ((size_t)k2 * 3 + x) * CS, acc);
            s = t1 != t2 ? scale_l[l] * (T)1.4142135623730951 : scale_l[l];
            tile = row * 3 * P + (long long)x * P + l * n * n;
        }
        if (staged) {
            staged_tile_store<T, NC, RP>(out, tile, q, s, acc, stage, stage_offs);
        } else if (tile >= 0) {
            tile_store<T, NC, RP>(out + tile, q, n, s, acc);
        }
    }
}

template <typename T, int NC, int RP, bool FAST>
__global__ void __launch_bounds__(TILE_THREADS) ps_gradient_tiled_kernel(
    SoapArgs<T> a, int max_runs, int max_entries, const T* __restrict__ coef, const T* __restrict__ dself,
    const unsigned* __restrict__ type_mask, const int4* __restrict__ runs, const int* __restrict__ run_rows,
    const int4* __restrict__ windows, const int* __restrict__ n_windows, const int* __restrict__ self_rows,
    const T* __restrict__ entry_data, T* __restrict__ out, const int* __restrict__ only) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int run_lo[TILE_MAX_RUNS], run_hi[TILE_MAX_RUNS], run_tj[TILE_MAX_RUNS], run_complete[TILE_MAX_RUNS];
    __shared__ int run_row[TILE_MAX_RUNS][TILE_MAX_TYPES];
    __shared__ int present[TILE_MAX_TYPES];
    __shared__ int carry_flag;
    __shared__ T scale_l[TILE_MAX_L];
    __shared__ int leg_l[TILE_MAX_L * (TILE_MAX_L + 1) / 2];
    __shared__ int leg_m[TILE_MAX_L * (TILE_MAX_L + 1) / 2];

    const int i = blockIdx.x;
    // with `only`, the centers it flags (the others go to the rank-2 kernel)
    if (only && only[i] == 0) return;
    const int tid = threadIdx.x;
    const int S = a.n_species, L = a.L, n = a.n;
    const int L2 = L * L;
    const int CS = L2 * NC;
    const int npairs = S * (S + 1) / 2;
    const TileEntryLayout lay(L, n, NC);
    T* C = reinterpret_cast<T*>(smem);
    T* region = C + (size_t)S * CS;
    T* DC = region;
    const size_t dc_size = (size_t)(max_runs > S ? max_runs : S) * 3 * CS;
    T* ent = region + dc_size;
    // staging of the stores overlays the entry data (free while rows are emitted)
    const int wave_id = tid / WAVE;
    constexpr int SPW = stage_floats_per_wave<T, NC, RP>();
    T* stage = ent + (size_t)wave_id * SPW;
    long long* stage_offs = reinterpret_cast<long long*>(ent + (size_t)(TILE_THREADS / WAVE) * SPW) + wave_id * WAVE;
    const bool staged = n == NC;

    const long long e0 = a.offsets[i];
    const unsigned mask = type_mask[i];
    const int n_present = __popc(mask);
    constexpr int nq = NC / 4;
    const int P = L * n * n;

    for (int k = tid; k < S * CS; k += TILE_THREADS) {
        const int t = k / CS;
        const int r = k - t * CS;
        const int lm = r / NC, c = r - lm * NC;
        C[k] = c < n ? coef[((long long)i * S + t) * L2 * n + lm * n + c] : (T)0;
    }
    if (tid < L) scale_l[tid] = (T)((tid % 2 == 0 ? 1.0 : -1.0) / sqrt(2.0 * tid + 1.0));
    for (int k = tid; k < lay.leg; k += TILE_THREADS) {
        int l = 0;
        while ((l + 1) * (l + 2) / 2 <= k) l++;
        leg_l[k] = l;
        leg_m[k] = k - l * (l + 1) / 2;
    }
    if (tid == 0) {
        int k = 0;
        for (int t = 0; t < S; t++)
            if ((mask >> t) & 1u) present[k++] = t;
    }
    // the tables above are read by every thread, also by the self rows of a center
    // without any window
    lds_barrier();

    const int nw = n_windows[i];
    for (int w = 0; w < nw; w++) {
        const int4 wd = windows[e0 + w];
        const int es = wd.x, ec = wd.y, rs = wd.z, rc = wd.w;
        if (tid < rc) {
            const int4 ri = runs[e0 + rs + tid];
            run_lo[tid] = max(ri.x, es) - es;
            run_hi[tid] = min(ri.x + ri.y, es + ec) - es;
            run_tj[tid] = ri.z;
            run_complete[tid] = ri.x + ri.y <= es + ec ? 1 : 0;
            if (tid == 0) carry_flag = ri.x < es ? 1 : 0;
        }
        for (int it = tid; it < rc * n_present; it += TILE_THREADS) {
            const int r = it / n_present, k = it - r * n_present;
            run_row[r][k] = run_rows[(e0 + rs + r) * S + k];
        }
        {
            // the window's entries are contiguous: one coalesced copy into LDS
            typedef typename Vec4Of<T>::type V;
            constexpr int VEC = 16 / sizeof(T);
            const V* src = reinterpret_cast<const V*>(entry_data + (e0 + es) * lay.per_entry);
            V* dst = reinterpret_cast<V*>(ent);
            const int count = ec * lay.per_entry / VEC;
            for (int k = tid; k < count; k += TILE_THREADS) dst[k] = src[k];
        }
        lds_barrier();
        plan_entry_stage2<T, NC>(a.sh_c, lay, rc, run_lo, run_hi, carry_flag != 0, leg_l, leg_m, ent, DC);
        lds_barrier();

        const int nc_runs = run_complete[rc - 1] ? rc : rc - 1;
        ps_window_rows<T, NC, RP, FAST>(L, n, n_present, present, nc_runs, run_tj, run_row, C, DC, scale_l, stage,
                                        stage_offs, out);
        lds_barrier();
    }

    // the center's own rows: d c^{t} / d r_i are the self gradients
    T* DS = region;
    for (int k = tid; k < n_present * 3 * CS; k += TILE_THREADS) {
        const int kt = k / (3 * CS);
        const int r = k - kt * 3 * CS;
        const int x = r / CS;
        const int r2 = r - x * CS;
        const int lm = r2 / NC, c = r2 - lm * NC;
        DS[k] = c < n ? dself[(((long long)i * S + present[kt]) * 3 + x) * L2 * n + lm * n + c] : (T)0;
    }
    lds_barrier();
    ps_self_rows<T, NC, RP>(i, L, n, S, n_present, present, self_rows, C, DS, scale_l, stage, stage_offs, out);
}

// ---------------------------------------------------------------------------
// Rank-2 gradient rows on MFMA (FP32, n = 8)
//
// One neighbour image e of a run contributes to the expansion gradient
//     DC_e[x][l^2+l+m][c] = alpha_x[lm] ra_l[c] + beta_x[lm] rb_l[c],
// alpha_x = Y_lm d_x, beta_x = r dY_lm/dx, ra = df R + f dR, rb = (f / r) R
// (soap/spherical_expansion_pair.rs:382-419): rank two in (m, c) for each (x, l).
// A row block of the power-spectrum gradient (soap/power_spectrum.rs:621-678),
// X = DC^T C_t summed over the run's images, is therefore
//     X[n1][n2] = sum_e ra_e[n1] u_e[n2] + rb_e[n1] v_e[n2],  u = alpha^T C_t, v = beta^T C_t:
// the contraction over m shrinks to the 8-vectors u and v, and every 8 x 8 block is
// a short sum of outer products.
//
// One workgroup per center, windows of whole runs (at most E images, 8 runs):
//  * the images' tables: (Y, T1, T2) per (l, m) with alpha_x = Y d_x and
//    beta_x = ga_x T1 + gb_x T2 (ga = (sin phi, -cos phi, 0), gb = (-cos theta cos phi,
//    -cos theta sin phi, sin theta) / 2; spherical_harmonics.rs:256-424), and
//    (ra, rb) per (l, c);
//  * per unit (x, l, four present types), one wave forms
//    W[(t, n2)][(e, a)] = sum_m A_a[e][m] C_t[m][n2] with v_mfma_f32_16x16x4f32
//    (A = alpha or beta, formed in the lanes from the tables), keeps W in its own
//    LDS slice, then writes the 8 x 8 blocks with v_mfma_f32_4x4x1f32: 16
//    independent 4 x 4 products, four blocks per instruction. Block b = (tile, fast
//    quad fa, slow quad sb): lane 4b + c holds out[4 sb + c][4 fa .. 4 fa + 3], one
//    16-byte piece of the row, and the 16 lanes of a tile cover its 256 contiguous
//    bytes, so the accumulators are stored as they are. The orientation of a block
//    (X for t >= tj, X^T for t <= tj, both when t == tj) is the lanes' choice of
//    operands: X takes (u, v) on the fast side and (ra, rb) on the slow side, X^T
//    the reverse.
// The center's own rows s_l (DS_t1^T C_t2 + C_t1^T DS_t2) come first, with the same
// 4 x 4 products over m, from the self gradients DS in LDS.
constexpr int R2_THREADS = 512;
constexpr int R2_WAVES = R2_THREADS / WAVE;
constexpr int R2_MAX_L = 16;

// Loader and compute waves of the rank-2 kernel. One loader wave fetches each window's
// plan and image records (a whole window per burst, one window ahead, into a staging
// area) and builds its tables into one of two LDS buffers, while seven compute waves
// write the previous window's rows. Compute waves issue no global load after their
// first store, so they never wait for their own stores to drain (loads and stores
// share one in-order counter per wave), and the loader, which never stores, waits
// only for its own loads.
constexpr int R2_LOADERS = 1;
constexpr int R2_COMPUTE = R2_WAVES - R2_LOADERS;

struct Rank2Layout {
    // float offsets in the dynamic LDS
    int L, L2, csp, E, wp, S, per_entry;
    int cp, st, st_runs, st_rows, st_raw, cms, einv, buf0, buf_size, w, ds, pairs, total;
    // offsets within a window buffer
    int b_hdr, b_img, b_row, b_ytab, b_rr, b_geo;
    __host__ __device__ Rank2Layout(int L_, int S_, int E_) : L(L_), L2(L_ * L_), E(E_), S(S_) {
        csp = L2 * 8 + 8;  // padded per present type: two types' column reads meet no bank conflict
        wp = 2 * E + 2;    // (e, a) pairs of one (t, n2) column, padded for conflict-free 8-byte reads
        per_entry = TileEntryLayout(L, 8, 8).per_entry;
        b_hdr = 0;         // images, runs
        b_img = 4;              // per image {neighbour type, first of its run, last of its run, run}
        b_row = b_img + 4 * 16;  // per (run, present type): 64-bit element offset of the row
        b_ytab = (b_row + 2 * TILE_MAX_RUNS * S + 3) / 4 * 4;  // (Y, T1, T2, -) per (image, lm)
        b_rr = b_ytab + E * L2 * 4;                        // (ra, rb) per (image, l, c)
        b_geo = b_rr + E * L * 16;                         // d, ga, gb per image
        buf_size = (b_geo + E * 12 + 3) / 4 * 4;
        cp = 0;
        // the loader's staging area: window descriptor and runs (int4 each), run rows,
        // image records, cos / sin (m phi) and reciprocals
        st = S * csp;
        st_runs = st + 4;
        st_rows = st_runs + 4 * TILE_MAX_RUNS;
        st_raw = (st_rows + TILE_MAX_RUNS * S + 3) / 4 * 4;
        cms = st_raw + E * per_entry;
        einv = cms + E * L * 2;
        buf0 = einv + E * 4;
        w = buf0 + 2 * buf_size;  // per compute wave: W [32][wp]
        const int window_end = w + R2_COMPUTE * 32 * wp;
        // the center's own rows, before the windows: self gradients DS [present][3][csp]
        // over the second buffer and the W slices, row of each present pair after them
        ds = buf0 + buf_size;
        const int ds_end = ds + S * 3 * csp;
        pairs = window_end > ds_end ? window_end : ds_end;
        total = pairs + S * S;
    }
};

__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
}

__device__ __forceinline__ float4_t mfma4x4(float a, float b, float4_t c) {
    return __builtin_amdgcn_mfma_f32_4x4x1f32(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ float4_t mfma16x16(float a, float b, float4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// wave-uniform copy of a value every lane holds alike (a scalar register)
__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ float uniform(float v) {
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

// Loader: window w's descriptor, runs, run rows and image records into the staging
// area, every load issued before the first wait (records of up to E images at L <= 7
// in registers; larger bases in several passes).
template <int E>
__device__ __attribute__((noinline)) void rank2_fetch(int L, int S, int lane, int n_present, long long e0, int w,
                                                      const int4* __restrict__ windows, const int4* __restrict__ runs,
                                                      const int* __restrict__ run_rows,
                                                      const float* __restrict__ entry_data, float* lds) {
    // arguments by value: a layout passed by reference would live in scratch memory,
    // written by every wave at the kernel's start
    const Rank2Layout ly(L, S, E);
    const int4 wd = windows[e0 + w];
    const int es = wd.x, ec = wd.y, rs = wd.z, rc = wd.w;
    const float4_t* src = reinterpret_cast<const float4_t*>(entry_data + (e0 + es) * ly.per_entry);
    const int count = ec * ly.per_entry / 4;
    constexpr int NV = (E * 152 / 4 + WAVE - 1) / WAVE;  // float4 pieces per lane and pass
    float4_t* dst = reinterpret_cast<float4_t*>(lds + ly.st_raw);
    float4_t v[NV];
#pragma unroll
    for (int k = 0; k < NV; k++) {
        const int idx = k * WAVE + lane;
        if (idx < count) v[k] = src[idx];
    }
    int4 run = {0, 0, 0, 0};
    if (lane < rc) run = runs[e0 + rs + lane];
    constexpr int NR = TILE_MAX_RUNS * TILE_MAX_TYPES / WAVE;
    int rowv[NR];
#pragma unroll
    for (int k = 0; k < NR; k++) {
        const int idx = k * WAVE + lane;
        if (idx < rc * n_present) {
            const int r = idx / n_present;
            rowv[k] = run_rows[(e0 + rs + r) * S + idx - r * n_present];
        }
    }
    int* ist = reinterpret_cast<int*>(lds + ly.st);
    if (lane == 0) {
        ist[0] = es;
        ist[1] = ec;
        ist[2] = rs;
        ist[3] = rc;
    }
    if (lane < rc) {
        int* ir = reinterpret_cast<int*>(lds + ly.st_runs) + 4 * lane;
        ir[0] = run.x;
        ir[1] = run.y;
        ir[2] = run.z;
    }
#pragma unroll
    for (int k = 0; k < NR; k++) {
        const int idx = k * WAVE + lane;
        if (idx < rc * n_present) {
            const int r = idx / n_present;
            reinterpret_cast<int*>(lds + ly.st_rows)[r * S + idx - r * n_present] = rowv[k];
        }
    }
#pragma unroll
    for (int k = 0; k < NV; k++) {
        const int idx = k * WAVE + lane;
        if (idx < count) dst[idx] = v[k];
    }
#pragma unroll 1
    for (int idx = NV * WAVE + lane; idx < count; idx += WAVE) dst[idx] = src[idx];
}

// Loader: the staged window's tables into `buf`: the runs and their rows, and per image
// (Y, T1, T2) per (l, m) with alpha_x = Y d_x and beta_x = ga_x T1 + gb_x T2
// (ga = (sin phi, -cos phi, 0), gb = (-cos theta cos phi, -cos theta sin phi, sin theta) / 2;
// the m = 0 derivative as T2 = -2 c_l P_l1; spherical_harmonics.rs:256-424),
// (ra, rb) = (df R + f dR, (f / r) R) per (l, c), and (d, ga, gb).
__device__ __forceinline__ void rank2_tables(const Rank2Layout& ly, const TileEntryLayout& lay, const ShTable& tb,
                                             const float* shc, const unsigned char* lm_l, int lane, int n_present,
                                             float* lds, float* buf) {
    const int L = ly.L, L2 = ly.L2, S = ly.S;
    const int* ist = reinterpret_cast<const int*>(lds + ly.st);
    const int es = ist[0], ec = ist[1], rc = ist[3];
    const float* raw = lds + ly.st_raw;
    float* cms = lds + ly.cms;
    float* einv = lds + ly.einv;
    int* ib = reinterpret_cast<int*>(buf);
    if (lane == 0) {
        ib[ly.b_hdr] = ec;
        ib[ly.b_hdr + 1] = rc;
    }
    if (lane < ec) {
        // the run of image `lane` (runs are consecutive image ranges covering the window)
        const int* ir = reinterpret_cast<const int*>(lds + ly.st_runs);
        int r = 0;
        while (r + 1 < rc && ir[4 * (r + 1)] - es <= lane) r++;
        const int lo = ir[4 * r] - es, cnt = ir[4 * r + 1];
        ib[ly.b_img + 4 * lane + 0] = ir[4 * r + 2];
        ib[ly.b_img + 4 * lane + 1] = lane == lo ? 1 : 0;
        ib[ly.b_img + 4 * lane + 2] = lane == lo + cnt - 1 ? 1 : 0;
        ib[ly.b_img + 4 * lane + 3] = r;
    }
    long long* rowoff = reinterpret_cast<long long*>(buf + ly.b_row);
#pragma unroll 1
    for (int it = lane; it < rc * n_present; it += WAVE) {
        const int r = it / n_present, k = it - r * n_present;
        rowoff[r * S + k] = (long long)reinterpret_cast<const int*>(lds + ly.st_rows)[r * S + k] * (3 * 64 * L);
    }
    // cos(m phi), sin(m phi) by the recurrence of spherical_harmonics.rs (same signs), and
    // the reciprocals of the Legendre derivative terms, per image
#pragma unroll 1
    for (int it = lane; it < ec * L; it += WAVE) {
        const int e = it / L, am = it - e * L;
        const float* g = raw + e * lay.per_entry + lay.trig;
        const float cos_phi = g[2], sin_phi = g[3];
        float cos_1 = 1, sin_1 = 0, cos_2 = -cos_phi, sin_2 = sin_phi;
        const float minus_two_cos = -2 * cos_phi;
        float sin_m = 0, cos_m = 1;
        for (int mm = 1; mm <= am; mm++) {
            sin_m = minus_two_cos * sin_1 - sin_2;
            cos_m = minus_two_cos * cos_1 - cos_2;
            sin_2 = sin_1;
            sin_1 = sin_m;
            cos_2 = cos_1;
            cos_1 = cos_m;
        }
        cms[it * 2 + 0] = cos_m;
        cms[it * 2 + 1] = sin_m;
        if (am == 0) {
            const float cos_theta = g[0], sin_theta = g[1];
            einv[e * 4 + 0] = sin_theta > 0.1f ? 1.0f / sin_theta : 0.f;
            einv[e * 4 + 1] = -0.5f / cos_theta;
        }
    }
    wave_lds_sync();
    float* ytab = buf + ly.b_ytab;
#pragma unroll 1
    for (int it = lane; it < ec * L2; it += WAVE) {
        const int e = it / L2, lmp = it - e * L2;
        const int l = lm_l[lmp];
        const int m = lmp - l * l - l;
        const float* base = raw + e * lay.per_entry;
        const float* Pl = base + lay.p;
        const float sin_theta = base[lay.trig + 1];
        const int am = m < 0 ? -m : m;
        const int k = leg_index(l, am);
        const float p = Pl[k];
        float Y, T1, T2;
        if (m == 0) {
            Y = p / 1.4142135623730951f;
            T1 = 0.f;
            T2 = l == 0 ? 0.f : -2.f * (shc[tb.gl + l] * Pl[leg_index(l, 1)]);
        } else {
            const float cos_m = cms[(e * L + am) * 2], sin_m = cms[(e * L + am) * 2 + 1];
            const float pm1 = Pl[leg_index(l, am - 1)];
            const float pp1 = am < l ? Pl[leg_index(l, am + 1)] : 0.f;
            const float dp = shc[tb.dpa + k] * pm1 - shc[tb.dpb + k] * pp1;
            const float pt = (sin_theta > 0.1f) ? ((float)am * einv[e * 4]) * p : einv[e * 4 + 1] * dp;
            const float c1 = m > 0 ? cos_m : sin_m;
            const float s1 = m > 0 ? sin_m : -cos_m;
            Y = p * c1;
            T1 = pt * s1;
            T2 = c1 * dp;
        }
        store4<float>(ytab + (size_t)(e * L2 + lmp) * 4, Y, T1, T2, 0.f);
    }
    float* rr = buf + ly.b_rr;
#pragma unroll 1
    for (int it = lane; it < ec * L * 8; it += WAVE) {
        const int e = it / (L * 8), lc = it - e * L * 8;
        const float* base = raw + e * lay.per_entry;
        const float f = base[lay.scal + 0], df = base[lay.scal + 1], f_r = base[lay.scal + 2];
        const float R = base[lay.rv + lc], dR = base[lay.rd + lc];
        rr[(e * L * 8 + lc) * 2 + 0] = df * R + f * dR;
        rr[(e * L * 8 + lc) * 2 + 1] = f_r * R;
    }
    float* geo = buf + ly.b_geo;
#pragma unroll 1
    for (int e = lane; e < ec; e += WAVE) {
        const float* base = raw + e * lay.per_entry;
        const float cos_theta = base[lay.trig + 0], sin_theta = base[lay.trig + 1];
        const float cos_phi = base[lay.trig + 2], sin_phi = base[lay.trig + 3];
        float* g = geo + e * 12;
        g[0] = base[lay.dir + 0];
        g[1] = base[lay.dir + 1];
        g[2] = base[lay.dir + 2];
        g[3] = sin_phi;
        g[4] = -cos_phi;
        g[5] = 0.f;
        g[6] = -0.5f * cos_theta * cos_phi;
        g[7] = -0.5f * cos_theta * sin_phi;
        g[8] = 0.5f * sin_theta;
    }
    wave_lds_sync();
}

// The center's own rows s_l (DS_t1^T C_t2 + C_t1^T DS_t2) for every present pair
// t1 <= t2 (soap/power_spectrum.rs:621-678), from the self gradients DS and the
// coefficients C in LDS, on v_mfma_f32_16x16x4f32 by the compute waves: one product
// per (x, l, 2 x 2 block of present-type pairs), rows (type t2, n2) and columns
// (type t1, n1), so that lane j holds out[n1][4 h .. 4 h + 3] of the pair (t1, t2) — a
// 16-byte piece of its row — with both terms accumulated in place:
//   D[(t2, n2)][(t1, n1)] = sum_m C_t2[m][n2] DS_t1[m][n1] + DS_t2[m][n2] C_t1[m][n1].
__device__ __forceinline__ void rank2_self_rows(int cw, int lane, int L, int n_present, int CSP, int P,
                                                const float* Cp, const float* DS, const int* pair_row,
                                                const float* scale_l, float* __restrict__ out) {
    // blocks (b1 <= b2) of 2 x 2 type pairs: types 2 b1 + (0, 1) on the column side,
    // 2 b2 + (0, 1) on the row side
    const int nb = (n_present + 1) / 2;
    const int nbp = nb * (nb + 1) / 2;
    const int units = 3 * L * nbp;
    const int j16 = lane & 15, jk = lane >> 4;
    for (int u = cw; u < units; u += R2_COMPUTE) {
        const int bp = u % nbp, lx = u / nbp;
        const int l = L - 1 - lx / 3, x = lx % 3;
        int b1 = 0, rem = bp;
        while (rem >= nb - b1) {
            rem -= nb - b1;
            b1++;
        }
        const int b2 = b1 + rem;
        const int k2 = 2 * b2 + (j16 >> 3), k1 = 2 * b1 + (j16 >> 3);
        const bool kv2 = k2 < n_present, kv1 = k1 < n_present;
        const int k2s = kv2 ? k2 : 0, k1s = kv1 ? k1 : 0;
        const float* C2 = Cp + k2s * CSP + l * l * 8 + (lane & 7);
        const float* D2 = DS + (k2s * 3 + x) * CSP + l * l * 8 + (lane & 7);
        const float* C1 = Cp + k1s * CSP + l * l * 8 + (lane & 7);
        const float* D1 = DS + (k1s * 3 + x) * CSP + l * l * 8 + (lane & 7);
        float4_t acc = {0.f, 0.f, 0.f, 0.f};
        const int ks_count = (2 * l + 1 + 3) / 4;
        for (int ks = 0; ks < ks_count; ks++) {
            const int mp = 4 * ks + jk;
            const bool mv = mp <= 2 * l;
            const int off = (mv ? mp : 0) * 8;
            const float c2 = C2[off], d2 = D2[off], c1 = C1[off], d1 = D1[off];
            acc = mfma16x16((mv && kv2) ? c2 : 0.f, (mv && kv1) ? d1 : 0.f, acc);
            acc = mfma16x16((mv && kv2) ? d2 : 0.f, (mv && kv1) ? c1 : 0.f, acc);
        }
        // lane j: rows 4 (j / 16) + v = (t2 = 2 b2 + j / 32, n2 = 4 ((j / 16) % 2) + v),
        // column j % 16 = (t1 = 2 b1 + (j % 16) / 8, n1 = j % 8)
        const int t2k = 2 * b2 + (lane >> 5), t1k = 2 * b1 + (j16 >> 3);
        if (t2k < n_present && t1k < n_present && t1k <= t2k) {
            const float sc = t1k != t2k ? scale_l[l] * 1.4142135623730951f : scale_l[l];
            float* dst = out + (long long)pair_row[t1k * n_present + t2k] * 3 * P + (long long)x * P + l * 64 +
                         (lane & 7) * 8 + 4 * ((lane >> 4) & 1);
            store4<float>(dst, acc[0] * sc, acc[1] * sc, acc[2] * sc, acc[3] * sc);
        }
    }
}

// (six waves per SIMD: three workgroups per CU; the loader's hoisted addresses must
// not take the compute waves' registers)
template <int E>
__global__ void __launch_bounds__(R2_THREADS) __attribute__((amdgpu_waves_per_eu(6, 8))) ps_gradient_rank2_kernel(
    SoapArgs<float> a, const float* __restrict__ coef, const float* __restrict__ dself,
    const unsigned* __restrict__ type_mask, const int4* __restrict__ runs, const int* __restrict__ run_rows,
    const int4* __restrict__ windows, const int* __restrict__ n_windows, const int* __restrict__ self_rows,
    const int* __restrict__ split, const float* __restrict__ entry_data, float* __restrict__ out) {
    constexpr int RTM = E / 8;  // row tiles of 16 (image, alpha / beta) rows
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int present[TILE_MAX_TYPES];
    __shared__ float scale_l[R2_MAX_L];
    __shared__ float shc[2 * R2_MAX_L * (R2_MAX_L + 1) / 2 + 4 * R2_MAX_L];
    __shared__ unsigned char lm_l[R2_MAX_L * R2_MAX_L];

    const int i = blockIdx.x;
    if (split[i] != 0) return;  // a run longer than a window: the tiled kernel's center
    const int nw = n_windows[i];
    const unsigned mask0 = type_mask[i];
    if (mask0 == 0u) return;  // no neighbour at all: no rows
    const int tid = threadIdx.x;
    const int lane = tid & (WAVE - 1);
    // wave-uniform (a scalar register), so that every loop over units is uniform
    const int wave_id = uniform(tid / WAVE);
    const int S = a.n_species, L = a.L;
    const int L2 = L * L;
    const Rank2Layout ly(L, S, E);
    const TileEntryLayout lay(L, 8, 8);
    const ShTable tb(L);
    float* lds = reinterpret_cast<float*>(smem);
    float* Cp = lds + ly.cp;
    const int CSP = ly.csp;
    const int P = L * 64;
    const long long e0 = a.offsets[i];
    const unsigned mask = type_mask[i];
    const int n_present = __popc(mask);
    const bool loader = wave_id < R2_LOADERS;

    if (tid == 0) {
        int k = 0;
        for (int t = 0; t < S; t++)
            if ((mask >> t) & 1u) present[k++] = t;
    }
    if (tid < L) scale_l[tid] = (float)((tid % 2 == 0 ? 1.0 : -1.0) / sqrt(2.0 * tid + 1.0));
    for (int k = tid; k < tb.size; k += R2_THREADS) shc[k] = a.sh_c[k];
    for (int k = tid; k < L2; k += R2_THREADS) {
        int l = 0;
        while ((l + 1) * (l + 1) <= k) l++;
        lm_l[k] = (unsigned char)l;
    }
    if (loader && nw > 0) rank2_fetch<E>(L, S, lane, __popc(mask), e0, 0, windows, runs, run_rows, entry_data, lds);
    lds_barrier();
    // coefficients and self gradients per present type (padded rows of csp floats), and
    // the row of each present pair, all read before the first store
    float* DS = lds + ly.ds;
    int* pair_row = reinterpret_cast<int*>(lds + ly.pairs);
    const int npairs = S * (S + 1) / 2;
    for (int k = tid; k < n_present * L2 * 8; k += R2_THREADS) {
        const int kk = k / (L2 * 8), r = k - kk * L2 * 8;
        Cp[kk * CSP + r] = coef[((long long)i * S + present[kk]) * L2 * 8 + r];
    }
    for (int k = tid; k < n_present * 3 * L2 * 8; k += R2_THREADS) {
        const int kx = k / (L2 * 8), r = k - kx * L2 * 8;
        const int kk = kx / 3, x = kx - 3 * kk;
        DS[kx * CSP + r] = dself[(((long long)i * S + present[kk]) * 3 + x) * L2 * 8 + r];
    }
    for (int k = tid; k < n_present * n_present; k += R2_THREADS) {
        const int k1 = k / n_present, k2 = k - k1 * n_present;
        if (k1 <= k2) pair_row[k] = self_rows[(long long)i * npairs + pair_index(present[k1], present[k2], S)];
    }
    if (loader && nw > 0) rank2_tables(ly, lay, tb, shc, lm_l, lane, n_present, lds, lds + ly.buf0);
    lds_barrier();
    if (loader) {
        if (nw > 1) rank2_fetch<E>(L, S, lane, n_present, e0, 1, windows, runs, run_rows, entry_data, lds);
    } else {
        rank2_self_rows(wave_id - R2_LOADERS, lane, L, n_present, CSP, P, Cp, DS, pair_row, scale_l, out);
    }
    lds_barrier();

    // compute-wave lane roles: 4 x 4 blocks (tile lane >> 4, fast quad fa, slow quad sb,
    // column q); 16 x 16 product operands (row / column j16, k row jk)
    const int cw = wave_id - R2_LOADERS;
    const int fa = (lane >> 3) & 1, sb = (lane >> 2) & 1, q = lane & 3;
    const int tau = lane >> 4;
    const int j16 = lane & 15, jk = lane >> 4;
    const int piece = (4 * sb + q) * 8 + 4 * fa;  // the lane's 16-byte piece in an 8 x 8 block
    const bool alpha_row = (j16 & 1) == 0;
    const int WP = ly.wp;
    float* Wv = lds + ly.w + (loader ? 0 : cw) * 32 * WP;
    const int n_tc = (n_present + 3) / 4;

    if (loader) {
        // one barrier per window, as the compute waves below
        for (int w = 0; w < nw; w++) {
            if (w + 1 < nw) {
                rank2_tables(ly, lay, tb, shc, lm_l, lane, n_present, lds, lds + ly.buf0 + ((w + 1) & 1) * ly.buf_size);
                if (w + 2 < nw) rank2_fetch<E>(L, S, lane, n_present, e0, w + 2, windows, runs, run_rows, entry_data, lds);
            }
            lds_barrier();
        }
        return;
    }
    for (int w = 0; w < nw; w++) {
        float* buf = lds + ly.buf0 + (w & 1) * ly.buf_size;
        {
            const int* ib = reinterpret_cast<const int*>(buf);
            const int ec = uniform(ib[ly.b_hdr]), rc = uniform(ib[ly.b_hdr + 1]);
            const float* ytab = buf + ly.b_ytab;
            const float* rr = buf + ly.b_rr;
            const float* geo = buf + ly.b_geo;
            // the lane's rows of the 16 x 16 products: image e = rho / 2, alpha (rho even) or beta
            const int RT = (2 * ec + 15) / 16;
            bool ev[RTM];
            int ybase[RTM], gbase[RTM];
#pragma unroll
            for (int rt = 0; rt < RTM; rt++) {
                const int e = (16 * rt + j16) >> 1;
                ev[rt] = e < ec;
                const int es_ = ev[rt] ? e : 0;
                gbase[rt] = es_ * 12;
                ybase[rt] = (es_ * L2 + jk) * 4;
            }
            // units: (type chunk, l descending, x), dealt in snake order over the compute waves
            const int units = 3 * L * n_tc;
            for (int slot = 0;; slot++) {
                const int round = slot;
                const int pos = (round & 1) ? R2_COMPUTE - 1 - cw : cw;
                const int u = round * R2_COMPUTE + pos;
                if (u >= units) break;
                const int tc = u / (3 * L), lx = u - tc * 3 * L;
                const int l = L - 1 - lx / 3, x = lx % 3;
                const int lsq = l * l;
                // (B) W[(t, n2)][(e, a)] = sum_m A_a[e][m] C_t[m][n2] for the chunk's four types
                // the image's d_x, ga_x, gb_x (LDS reads: a select over x would become a
                // private array in scratch memory, whose loads wait for the wave's stores)
                float dx[RTM], gax[RTM], gbx[RTM];
#pragma unroll
                for (int rt = 0; rt < RTM; rt++) {
                    dx[rt] = geo[gbase[rt] + x];
                    gax[rt] = geo[gbase[rt] + 3 + x];
                    gbx[rt] = geo[gbase[rt] + 6 + x];
                }
                const int kk0 = 4 * tc + (j16 >> 3);
                const bool kv0 = kk0 < n_present, kv1 = kk0 + 2 < n_present;
                const float* cb0 = Cp + kk0 * CSP + (lane & 7);
                const float* cb1 = cb0 + 2 * CSP;
                float4_t accb[RTM][2];
#pragma unroll
                for (int rt = 0; rt < RTM; rt++) {
                    accb[rt][0] = float4_t{0.f, 0.f, 0.f, 0.f};
                    accb[rt][1] = float4_t{0.f, 0.f, 0.f, 0.f};
                }
                const int ks_count = (2 * l + 1 + 3) / 4;
                for (int ks = 0; ks < ks_count; ks++) {
                    const int mp = 4 * ks + jk;
                    const bool mv = mp <= 2 * l;
                    // reads at a valid slot always, zeros selected afterwards
                    const int lm = lsq + (mv ? mp : 0);
                    const float c0 = cb0[lm * 8], c1 = cb1[lm * 8];
                    const float b0 = (mv && kv0) ? c0 : 0.f;
                    const float b1 = (mv && kv1) ? c1 : 0.f;
#pragma unroll
                    for (int rt = 0; rt < RTM; rt++) {
                        if (rt < RT) {
                            const float4_t y =
                                *reinterpret_cast<const float4_t*>(ytab + ybase[rt] + (lsq + 4 * ks) * 4);
                            float av = alpha_row ? y[0] * dx[rt] : gax[rt] * y[1] + gbx[rt] * y[2];
                            av = (mv && ev[rt]) ? av : 0.f;
                            accb[rt][0] = mfma16x16(av, b0, accb[rt][0]);
                            accb[rt][1] = mfma16x16(av, b1, accb[rt][1]);
                        }
                    }
                }
                // W carries the power spectrum's s_l = (-1)^l / sqrt(2 l + 1)
                const float s_l = uniform(scale_l[l]);
#pragma unroll
                for (int rt = 0; rt < RTM; rt++) {
                    if (rt < RT) {
#pragma unroll
                        for (int ct = 0; ct < 2; ct++) {
                            float* wdst = Wv + (16 * ct + j16) * WP + 16 * rt + 4 * jk;
#pragma unroll
                            for (int v = 0; v < 4; v++) wdst[v] = accb[rt][ct][v] * s_l;
                        }
                    }
                }
                wave_lds_sync();
                // (C) per run, the 8 x 8 blocks of the chunk's (up to) four types, one per
                // 16 lanes: X = sum_e ra_e u_e^T + rb_e v_e^T in the block's orientation
                const int nk = min(4, n_present - 4 * tc);
                const bool lane_type = tau < nk;
                const int kslot = 4 * tc + (lane_type ? tau : 0);
                const int t_lane = present[kslot];
                const int offWA = (8 * tau + 4 * fa + q) * WP, offWB = (8 * tau + 4 * sb + q) * WP;
                const int offRA = (l * 8 + 4 * fa + q) * 2, offRB = (l * 8 + 4 * sb + q) * 2;
                const long long* rowoff = reinterpret_cast<const long long*>(buf + ly.b_row) + kslot;
                float* out_unit = out + (long long)x * P + l * 64 + piece;
                // image by image (unrolled, no inner loop), the sums of a run's images in the
                // accumulators, its block stored at its last image
                const float* wA = Wv + offWA;
                const float* wB = Wv + offWB;
                const float* rA = rr + offRA;
                const float* rB = rr + offRB;
                float4_t acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int e = 0; e < E; e++) {
                    if (e >= ec) break;
                    const int4 rd = *reinterpret_cast<const int4*>(ib + ly.b_img + 4 * e);
                    const int tj = uniform(rd.x), first = uniform(rd.y), last = uniform(rd.z), r = uniform(rd.w);
                    const bool f1 = t_lane >= tj, f2 = t_lane <= tj, fb = f1 && f2;
                    const float2 W_a = *reinterpret_cast<const float2*>(wA + 2 * e);
                    const float2 W_b = *reinterpret_cast<const float2*>(wB + 2 * e);
                    const float2 R_a = *reinterpret_cast<const float2*>(rA + e * (16 * L));
                    const float2 R_b = *reinterpret_cast<const float2*>(rB + e * (16 * L));
                    if (first) acc = float4_t{0.f, 0.f, 0.f, 0.f};
                    // X (t >= tj): (u, v) fast, (ra, rb) slow; X^T (t < tj) the reverse;
                    // both when t == tj
                    acc = mfma4x4(f1 ? W_a.x : R_a.x, f1 ? R_b.x : W_b.x, acc);
                    acc = mfma4x4(f1 ? W_a.y : R_a.y, f1 ? R_b.y : W_b.y, acc);
                    acc = mfma4x4(fb ? R_a.x : 0.f, W_b.x, acc);
                    acc = mfma4x4(fb ? R_a.y : 0.f, W_b.y, acc);
                    if (last && lane_type) {
                        // sqrt(2) for the blocks of two different types
                        const float sc = fb ? 1.0f : 1.4142135623730951f;
                        store4<float>(out_unit + rowoff[r * S], acc[0] * sc, acc[1] * sc, acc[2] * sc, acc[3] * sc);
                    }
                }
                wave_lds_sync();
            }
        }
        lds_barrier();
    }
}

// Spherical-expansion position gradients from the planned windows: the same
// per-center workgroup, window staging and stage-2 sums as the power-spectrum
// kernel, DC_run = d c_i / d r_j for each neighbour atom j (images summed), whose
// rows are the output itself: in block (l, center type, type of j), row (i, j)
// holds [3][2l+1][n] (soap/spherical_expansion.rs:426-525). A run's rows for one
// l are 3 (2l+1) n contiguous values, written in 4-value pieces by consecutive
// lanes. The center's own rows copy its self gradients.
template <typename T, int NC>
__global__ void __launch_bounds__(TILE_THREADS) se_gradient_tiled_kernel(
    SoapArgs<T> a, int max_runs, int max_entries, const T* __restrict__ dself, const unsigned* __restrict__ type_mask,
    const int4* __restrict__ runs, const int* __restrict__ run_rows, const int4* __restrict__ windows,
    const int* __restrict__ n_windows, const int* __restrict__ self_rows, const long long* __restrict__ se_grad_base,
    const T* __restrict__ entry_data, T* __restrict__ out) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int run_lo[TILE_MAX_RUNS], run_hi[TILE_MAX_RUNS], run_complete[TILE_MAX_RUNS];
    __shared__ long long run_base[TILE_MAX_RUNS];
    __shared__ int run_type[TILE_MAX_RUNS];
    __shared__ int carry_flag;
    __shared__ int leg_l[TILE_MAX_L * (TILE_MAX_L + 1) / 2];
    __shared__ int leg_m[TILE_MAX_L * (TILE_MAX_L + 1) / 2];

    const int i = blockIdx.x;
    const int tid = threadIdx.x;
    const int S = a.n_species, L = a.L, n = a.n;
    const int L2 = L * L;
    const int CS = L2 * NC;
    const TileEntryLayout lay(L, n, NC);
    T* DC = reinterpret_cast<T*>(smem);
    T* ent = DC + (size_t)max_runs * 3 * CS;
    const long long e0 = a.offsets[i];
    const int ct = a.species[i];
    const unsigned mask = type_mask[i];
    // values per run: 3 L² n, as 4-value pieces when n fills the padded width
    const bool quads = n == NC;
    const int per_run = quads ? 3 * CS / 4 : 3 * L2 * n;
    const int unit = quads ? 3 * n / 4 : 3 * n;  // pieces per (l, m) over the 3 directions

    for (int k = tid; k < lay.leg; k += TILE_THREADS) {
        int l = 0;
        while ((l + 1) * (l + 2) / 2 <= k) l++;
        leg_l[k] = l;
        leg_m[k] = k - l * (l + 1) / 2;
    }
    lds_barrier();

    // piece k of a run's (or the center's) rows -> (l, x, offset within the (2l+1) n values)
    auto decode = [&](int k, int& l, int& x, int& off) {
        const int u = k / unit;  // in [l², (l + 1)²)
        l = (int)sqrtf((float)u);
        while (l * l > u) l--;
        while ((l + 1) * (l + 1) <= u) l++;
        const int q = (2 * l + 1) * (quads ? n / 4 : n);  // pieces per direction
        const int rem = k - l * l * unit;
        x = rem / q;
        off = (rem - x * q) * (quads ? 4 : 1);
    };

    const int nw = n_windows[i];
    for (int w = 0; w < nw; w++) {
        const int4 wd = windows[e0 + w];
        const int es = wd.x, ec = wd.y, rs = wd.z, rc = wd.w;
        if (tid < rc) {
            const int4 ri = runs[e0 + rs + tid];
            run_lo[tid] = max(ri.x, es) - es;
            run_hi[tid] = min(ri.x + ri.y, es + ec) - es;
            run_complete[tid] = ri.x + ri.y <= es + ec ? 1 : 0;
            run_type[tid] = ri.z;
            run_base[tid] = run_rows[(e0 + rs + tid) * S];
            if (tid == 0) carry_flag = ri.x < es ? 1 : 0;
        }
        {
            typedef typename Vec4Of<T>::type V;
            constexpr int VEC = 16 / sizeof(T);
            const V* src = reinterpret_cast<const V*>(entry_data + (e0 + es) * lay.per_entry);
            V* dst = reinterpret_cast<V*>(ent);
            const int count = ec * lay.per_entry / VEC;
            for (int k = tid; k < count; k += TILE_THREADS) dst[k] = src[k];
        }
        lds_barrier();
        plan_entry_stage2<T, NC>(a.sh_c, lay, rc, run_lo, run_hi, carry_flag != 0, leg_l, leg_m, ent, DC);
        lds_barrier();

        const int nc_runs = run_complete[rc - 1] ? rc : rc - 1;
        for (int t = tid; t < nc_runs * per_run; t += TILE_THREADS) {
            const int r = t / per_run, k = t - r * per_run;
            int l, x, off;
            decode(k, l, x, off);
            const int wl = (2 * l + 1) * n;
            const long long tb = (long long)ct * S + run_type[r];
            T* dst = out + se_grad_base[tb * L + l] + run_base[r] * 3 * wl + (long long)x * wl + off;
            const T* src = DC + ((size_t)r * 3 + x) * CS + l * l * NC;
            if (quads) {
                T v[4];
                load4<T>(src + off, v);
                store4<T>(dst, v[0], v[1], v[2], v[3]);
            } else {
                const int lm = off / n, c = off - lm * n;
                *dst = src[lm * NC + c];
            }
        }
        lds_barrier();
    }

    // the center's own rows: its self gradients d c_i^{nt} / d r_i
    const int n_present = __popc(mask);
    const int per_self = quads ? 3 * L2 * n / 4 : 3 * L2 * n;
    for (int t = tid; t < n_present * per_self; t += TILE_THREADS) {
        const int p = t / per_self, k = t - p * per_self;
        int nt = -1;
        for (int s = 0, c = 0; s < S; s++) {
            if ((mask >> s) & 1u) {
                if (c == p) {
                    nt = s;
                    break;
                }
                c++;
            }
        }
        int l, x, off;
        decode(k, l, x, off);
        const int wl = (2 * l + 1) * n;
        const long long tb = (long long)ct * S + nt;
        const long long row = self_rows[(long long)i * S + nt];
        T* dst = out + se_grad_base[tb * L + l] + row * 3 * wl + (long long)x * wl + off;
        const T* src = dself + (((long long)i * S + nt) * 3 + x) * L2 * n + l * l * n + off;
        if (quads) {
            T v[4];
            load4<T>(src, v);
            store4<T>(dst, v[0], v[1], v[2], v[3]);
        } else {
            *dst = *src;
        }
    }
}

// Power-spectrum values with the m-contraction on MFMA (FP32 in and out, bfloat16
// hi + lo products, FP32 accumulation), one center per workgroup
// (soap/power_spectrum.rs:529-618). Per l, the Gram matrix over all (type, n)
// columns, G_l[(t1, n1)][(t2, n2)] = Σ_m c^{t1}[l, m, n1] c^{t2}[l, m, n2], is
// formed as the 16 x 16 tiles of its upper triangle; every block (t1 <= t2) with
// a sample at this center then takes its n x n piece, times κ / ((-1)^l √(2l+1)).
// Columns of types absent around the center are zero.
constexpr int PSV_THREADS = 256;
constexpr int PSV_MAX_TYPES = 16;

__global__ void __launch_bounds__(PSV_THREADS) ps_values_mfma_kernel(
    int n_atoms, int S, int L, int n, int W16, const float* __restrict__ coef, const int* __restrict__ species,
    const unsigned* __restrict__ type_mask, const long long* __restrict__ sample_flag,
    const long long* __restrict__ sample_scan, float* __restrict__ out) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ long long pair_row[PSV_MAX_TYPES * (PSV_MAX_TYPES + 1) / 2];
    __shared__ int pair_t1[PSV_MAX_TYPES * (PSV_MAX_TYPES + 1) / 2], pair_t2[PSV_MAX_TYPES * (PSV_MAX_TYPES + 1) / 2];
    const int Wp = 16 * W16;
    const int L2 = L * L;
    float* Cm = reinterpret_cast<float*>(smem);  // [l m][t n], zero-padded to Wp columns
    float* G = Cm + (size_t)L2 * Wp;             // [Wp][Wp], upper tiles
    const int i = blockIdx.x, tid = threadIdx.x;
    const int lane = tid & (WAVE - 1), wave = tid / WAVE;
    const unsigned mask = type_mask[i];
    const int ct = species[i];
    const int npairs = S * (S + 1) / 2;
    for (int k = tid; k < L2 * Wp; k += PSV_THREADS) {
        const int lm = k / Wp, col = k - lm * Wp;
        const int t = col / n, c = col - t * n;
        float v = 0.f;
        if (t < S && ((mask >> t) & 1u)) v = coef[((long long)i * S + t) * L2 * n + lm * n + c];
        Cm[k] = v;
    }
    for (int p = tid; p < npairs; p += PSV_THREADS) {
        int t1 = 0, rem = p;
        while (rem >= S - t1) {
            rem -= S - t1;
            t1++;
        }
        const int t2 = t1 + rem;
        const long long idx = (long long)(ct * npairs + pair_index(t1, t2, S)) * n_atoms + i;
        const bool on = ((mask >> t1) & 1u) && ((mask >> t2) & 1u) && sample_flag[idx] != 0;
        pair_t1[p] = t1;
        pair_t2[p] = t2;
        pair_row[p] = on ? sample_scan[idx] : -1;
    }
    __syncthreads();
    const int P = L * n * n;
    const int nq4 = n / 4;
    const int ntiles = W16 * (W16 + 1) / 2;
    const int i16 = lane & 15, kq = 4 * (lane >> 4);
    for (int l = 0; l < L; l++) {
        const int kcount = 2 * l + 1;
        const float* Cl = Cm + (size_t)l * l * Wp;
        for (int tile = wave; tile < ntiles; tile += PSV_THREADS / WAVE) {
            int bi = 0, rem = tile;
            while (rem >= W16 - bi) {
                rem -= W16 - bi;
                bi++;
            }
            const int bj = bi + rem;
            float4_t acc = {0.f, 0.f, 0.f, 0.f};
            // v_mfma_f32_16x16x32_bf16, per 16 values of m: lane i + 16 g holds K values
            // 8 g .. 8 g + 7; K < 16 carries the hi parts of A, K >= 16 its lo parts, both
            // against the hi parts of B (hi hi + lo hi), then the hi parts of A against the
            // lo parts of B on K < 16 (hi lo)
            const bool upper = lane >= 32;
            for (int kb = 0; kb < kcount; kb += 16) {
                short8_t a1, b1, a2, b2;
#pragma unroll
                for (int e = 0; e < 8; e++) {
                    const int mi = kb + 8 * ((lane >> 4) & 1) + e;
                    const float av = mi < kcount ? Cl[mi * Wp + 16 * bi + i16] : 0.f;
                    const float bv = mi < kcount ? Cl[mi * Wp + 16 * bj + i16] : 0.f;
                    short ah, al, bh, bl;
                    bf16_split(av, ah, al);
                    bf16_split(bv, bh, bl);
                    a1[e] = upper ? al : ah;
                    b1[e] = bh;
                    a2[e] = upper ? (short)0 : ah;
                    b2[e] = upper ? (short)0 : bl;
                }
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8_t, a1),
                                                              __builtin_bit_cast(bf16x8_t, b1), acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8_t, a2),
                                                              __builtin_bit_cast(bf16x8_t, b2), acc, 0, 0, 0);
            }
#pragma unroll
            for (int v = 0; v < 4; v++) G[(size_t)(16 * bi + kq + v) * Wp + 16 * bj + i16] = acc[v];
        }
        __syncthreads();
        const float scale = (float)((l % 2 == 0 ? 1.0 : -1.0) / sqrt(2.0 * l + 1.0));
        const int per_pair = n * nq4;
        for (int k = tid; k < npairs * per_pair; k += PSV_THREADS) {
            const int p = k / per_pair, rem = k - p * per_pair;
            const long long s = pair_row[p];
            if (s < 0) continue;
            const int n1 = rem / nq4, h = rem - n1 * nq4;
            const int t1 = pair_t1[p], t2 = pair_t2[p];
            const int rho = t1 * n + n1, kap = t2 * n + 4 * h;
            const float sc = t1 != t2 ? scale * 1.4142135623730951f : scale;
            float v[4];
            if ((rho >> 4) > (kap >> 4)) {
                // lower tile of a diagonal (t1 == t2) block: G is symmetric there
#pragma unroll
                for (int c = 0; c < 4; c++) v[c] = G[(size_t)(kap + c) * Wp + rho];
            } else {
#pragma unroll
                for (int c = 0; c < 4; c++) v[c] = G[(size_t)rho * Wp + kap + c];
            }
            store4<float>(out + s * P + l * n * n + n1 * n + 4 * h, v[0] * sc, v[1] * sc, v[2] * sc, v[3] * sc);
        }
        __syncthreads();
    }
}

// Angular factors of a chunk of neighbour entries (per-entry data in LDS), one
// item per (entry, l, m >= 0): ang[e][0][lm] = Y_lm and, with `grads`,
// ang[e][1 + x][lm] = r dY_lm/dx (spherical_harmonics.rs:256-298, :304-424).
template <typename T>
__device__ void chunk_angular_factors(const SoapArgs<T>& a, const TileEntryLayout& lay, int cnt, const int* leg_l,
                                      const int* leg_m, const T* ent, T* ang, bool grads) {
    const int leg = lay.leg, L2 = lay.L2;
    const ShTable tb(lay.L);
    const T SQRT_2 = (T)1.4142135623730951;
    const int tid = threadIdx.x;
    for (int item = tid; item < cnt * leg; item += TILE_THREADS) {
        const int e = item / leg, k = item - e * leg;
        const int l = leg_l[k], m = leg_m[k];
        const int lm0 = l * l + l;
        const T* base = ent + e * lay.per_entry;
        const T* P = base + lay.p;
        const T cos_theta = base[lay.trig + 0], sin_theta = base[lay.trig + 1];
        const T cos_phi = base[lay.trig + 2], sin_phi = base[lay.trig + 3];
        T* A = ang + (size_t)e * 4 * L2;
        if (m == 0) {
            A[lm0] = P[k] / SQRT_2;
            if (grads) {
                T g[3] = {0, 0, 0};
                if (l > 0) {
                    const T lf = a.sh_c[tb.gl + l] * P[leg_index(l, 1)];
                    g[0] = cos_phi * cos_theta * lf;
                    g[1] = sin_phi * cos_theta * lf;
                    g[2] = -sin_theta * lf;
                }
                for (int x = 0; x < 3; x++) A[(1 + x) * L2 + lm0] = g[x];
            }
        } else {
            T cos_1 = 1, sin_1 = 0, cos_2 = -cos_phi, sin_2 = sin_phi;
            const T minus_two_cos = -2 * cos_phi;
            T sin_m = 0, cos_m = 1;
            for (int mm = 1; mm <= m; mm++) {
                sin_m = minus_two_cos * sin_1 - sin_2;
                cos_m = minus_two_cos * cos_1 - cos_2;
                sin_2 = sin_1;
                sin_1 = sin_m;
                cos_2 = cos_1;
                cos_1 = cos_m;
            }
            const T p = P[k];
            A[lm0 + m] = p * cos_m;
            A[lm0 - m] = p * sin_m;
            if (grads) {
                const T c_dpa = a.sh_c[tb.dpa + k], c_dpb = a.sh_c[tb.dpb + k];
                const T pm1 = P[leg_index(l, m - 1)];
                const T pp1 = m < l ? P[leg_index(l, m + 1)] : (T)0;
                const T dp = c_dpa * pm1 - c_dpb * pp1;
                const T pt = (sin_theta > (T)0.1) ? (T)m / sin_theta * p : -(T)0.5 / cos_theta * dp;
                const T s_dp = sin_m * dp, c_dp = cos_m * dp;
                A[L2 + lm0 + m] = sin_phi * pt * sin_m - (T)0.5 * cos_theta * cos_phi * c_dp;
                A[L2 + lm0 - m] = -sin_phi * pt * cos_m - (T)0.5 * cos_theta * cos_phi * s_dp;
                A[2 * L2 + lm0 + m] = -cos_phi * pt * sin_m - (T)0.5 * cos_theta * sin_phi * c_dp;
                A[2 * L2 + lm0 - m] = cos_phi * pt * cos_m - (T)0.5 * cos_theta * sin_phi * s_dp;
                A[3 * L2 + lm0 + m] = (T)0.5 * sin_theta * c_dp;
                A[3 * L2 + lm0 - m] = (T)0.5 * sin_theta * s_dp;
            }
        }
    }
}

// Spherical expansion c_i and self gradients d c_i / d r_i of one center per
// 256-thread workgroup, from the per-entry data of entry_prep_kernel
// (soap/spherical_expansion.rs:101-358, spherical_expansion_pair.rs:323-422).
// The center's entries are read in contiguous chunks. Per chunk:
//   A. one item per (entry, l, m >= 0) forms Y_l(±m) (spherical_harmonics.rs:
//      304-424) and the three r dY_l(±m)/dx once, into LDS;
//   B. one item per (component, l m, NC / NH radial values) accumulates
//      f R Y (values) or -((df R + f dR) Y d_x + (f / r) R r dY/dx) (self
//      gradients; not for periodic images of the center itself) for the
//      neighbour type of each entry, keeping its sums in registers while the type
//      stays the same and adding them to the per-type sums in LDS when it changes.
template <typename T, int NC, int NH>
__global__ void __launch_bounds__(TILE_THREADS) expansion_tiled_kernel(SoapArgs<T> a, int max_entries,
                                                                       const T* __restrict__ entry_data,
                                                                       T* __restrict__ out_coef, T* __restrict__ out_grad) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int chunk_t[WAVE];
    __shared__ int chunk_self[WAVE];
    __shared__ int leg_l[TILE_MAX_L * (TILE_MAX_L + 1) / 2];
    __shared__ int leg_m[TILE_MAX_L * (TILE_MAX_L + 1) / 2];
    constexpr int NW = NC / NH;  // radial values per item
    const int i = blockIdx.x;
    const int tid = threadIdx.x;
    const int S = a.n_species, L = a.L, n = a.n;
    const int L2 = L * L;
    const int CS = L2 * NC;
    const TileEntryLayout lay(L, n, NC);
    const int leg = lay.leg;
    const bool grads = out_grad != nullptr;
    const int ncomp = grads ? 4 : 1;
    T* acc = reinterpret_cast<T*>(smem);                   // [S][4][L2][NC]
    T* ent = acc + (size_t)S * 4 * CS;                      // [max_entries][per_entry]
    T* ang = ent + (size_t)max_entries * lay.per_entry;    // [max_entries][4][L2]: Y, r dY/dx, r dY/dy, r dY/dz
    const long long e0 = a.offsets[i];
    const int len = (int)(a.offsets[i + 1] - e0);
    const ShTable tb(L);
    const T SQRT_2 = (T)1.4142135623730951;

    for (int k = tid; k < S * 4 * CS; k += TILE_THREADS) acc[k] = 0;
    for (int k = tid; k < leg; k += TILE_THREADS) {
        int l = 0;
        while ((l + 1) * (l + 2) / 2 <= k) l++;
        leg_l[k] = l;
        leg_m[k] = k - l * (l + 1) / 2;
    }
    for (int cb = 0; cb < len; cb += max_entries) {
        const int cnt = min(max_entries, len - cb);
        lds_barrier();
        {
            typedef typename Vec4Of<T>::type V;
            constexpr int VEC = 16 / sizeof(T);
            const V* src = reinterpret_cast<const V*>(entry_data + (e0 + cb) * lay.per_entry);
            V* dst = reinterpret_cast<V*>(ent);
            const int count = cnt * lay.per_entry / VEC;
            for (int k = tid; k < count; k += TILE_THREADS) dst[k] = src[k];
        }
        for (int c = tid; c < cnt; c += TILE_THREADS) {
            const int j = a.entry_j[e0 + cb + c];
            chunk_t[c] = a.species[j];
            chunk_self[c] = j == i ? 1 : 0;
        }
        lds_barrier();
        // A. angular factors of every (entry, l, m >= 0)
        chunk_angular_factors<T>(a, lay, cnt, leg_l, leg_m, ent, ang, grads);
        lds_barrier();
        // B. accumulation per (component, l m, part of n)
        for (int item = tid; item < ncomp * L2 * NH; item += TILE_THREADS) {
            const int comp = item / (L2 * NH);
            const int rest = item - comp * L2 * NH;
            const int lm = rest / NH, h = rest - lm * NH;
            const int l = l_of_lm(lm);
            const int x = comp - 1;
            T sum[NW];
            int cur_t = -1;
            auto flush = [&]() {
                T* d = acc + ((size_t)cur_t * 4 + comp) * CS + lm * NC + h * NW;
#pragma unroll
                for (int c = 0; c < NW; c += 4) {
                    T o[4];
                    load4<T>(d + c, o);
                    store4<T>(d + c, o[0] + sum[c], o[1] + sum[c + 1], o[2] + sum[c + 2], o[3] + sum[c + 3]);
                }
            };
            for (int e = 0; e < cnt; e++) {
                const int t = chunk_t[e];
                if (comp > 0 && chunk_self[e]) continue;
                if (t != cur_t) {
                    if (cur_t >= 0) flush();
                    cur_t = t;
#pragma unroll
                    for (int c = 0; c < NW; c++) sum[c] = 0;
                }
                const T* base = ent + e * lay.per_entry;
                const T* A = ang + (size_t)e * 4 * L2;
                const T f = base[lay.scal + 0];
                const T y = A[lm];
                T Rv[NW];
#pragma unroll
                for (int c = 0; c < NW; c += 4) load4<T>(base + lay.rv + l * NC + h * NW + c, Rv + c);
                if (comp == 0) {
#pragma unroll
                    for (int c = 0; c < NW; c++) {
                        const T v = f * Rv[c];
                        sum[c] += v * y;
                    }
                } else {
                    const T df = base[lay.scal + 1], f_r = base[lay.scal + 2];
                    const T yd = y * base[lay.dir + x];
                    const T g = A[comp * L2 + lm];
                    T dR[NW];
#pragma unroll
                    for (int c = 0; c < NW; c += 4) load4<T>(base + lay.rd + l * NC + h * NW + c, dR + c);
#pragma unroll
                    for (int c = 0; c < NW; c++) {
                        const T ra = df * Rv[c] + f * dR[c];
                        const T rb = f_r * Rv[c];
                        sum[c] -= ra * yd + rb * g;
                    }
                }
            }
            if (cur_t >= 0) flush();
        }
    }
    lds_barrier();
    // write out, dropping the padding; the self contribution goes to the l = 0
    // values of the center's own type (spherical_expansion_pair.rs:222-252)
    const int ct = a.species[i];
    const int slots = L2 * n;
    for (int k = tid; k < S * slots; k += TILE_THREADS) {
        const int t = k / slots;
        const int s = k - t * slots;
        const int lm = s / n, c = s - lm * n;
        T v = acc[((size_t)t * 4) * CS + lm * NC + c];
        if (t == ct && s < n) v += a.self_contribution[s];
        out_coef[(long long)i * S * slots + k] = v;
    }
    if (grads) {
        for (int k = tid; k < S * 3 * slots; k += TILE_THREADS) {
            const int t = k / (3 * slots);
            const int r = k - t * 3 * slots;
            const int x = r / slots;
            const int s = r - x * slots;
            const int lm = s / n, c = s - lm * n;
            out_grad[(long long)i * S * 3 * slots + k] = acc[((size_t)t * 4 + 1 + x) * CS + lm * NC + c];
        }
    }
}

// ---------------------------------------------------------------------------
// Spherical expansion c_i and self gradients d c_i / d r_i on the matrix cores (FP32,
// n = 8, l < 9), from the per-entry data of entry_prep_kernel (soap/
// spherical_expansion.rs:101-358, spherical_expansion_pair.rs:323-422). One center
// per 256-thread workgroup, its entries in chunks of E. Per chunk, every thread
// stages the records (loaded into registers during the previous chunk) and builds
// the chunk's tables in LDS:
//   (Y, T1, T2) per (entry, l m), with r dY_lm/dx = ga_x T1 + gb_x T2
//   (ga = (sin phi, -cos phi, 0), gb = (-cos theta cos phi, -cos theta sin phi,
//   sin theta) / 2, the m = 0 derivative as T2 = -2 c_l P_l1; spherical_harmonics.rs:
//   256-424); f R, and (ra, rb) = (df R + f dR, (f / r) R) per (entry, l, c); d, ga, gb
//   per entry.
// Wave 0 then sums the values c[t][lm][c] = sum_e f R_l[c] Y_lm and wave 1 + x the
// self gradients -sum_e ra_l[c] Y_lm d_x + rb_l[c] r dY_lm/dx (not for periodic
// images of the center itself), as outer products on v_mfma_f32_4x4x1f32: block
// b = (l, quad mq of m, quad cq of c) takes A = four radial values and B = four
// angular values, and lane 4 b + j holds out[l^2 + 4 mq + j][4 cq .. 4 cq + 3], one
// 16-byte piece of the output. The chunk's entries are visited grouped by type, one
// set of accumulators per type of a group of four; more types take one pass over the
// entries per group. For the power spectrum (at most 16 types), the coefficients also
// stay in LDS and the center's power-spectrum values follow (expansion_ps_values).
constexpr int EXM_THREADS = 256;
constexpr int EXM_MAX_L = 9;
constexpr int EXM_MAX_VALUE_TYPES = 16;

struct ExpansionMfmaLayout {
    // float offsets in the dynamic LDS
    int L, L2, per_entry, raw, ytab, fr, rab, geo, csp, cval, total;
    // SC: types whose coefficients stay in LDS for the power-spectrum values (0: none)
    __host__ __device__ ExpansionMfmaLayout(int L_, int E, int SC) : L(L_), L2(L_ * L_) {
        per_entry = TileEntryLayout(L, 8, 8).per_entry;
        raw = 0;                     // the chunk's records
        ytab = raw + E * per_entry;  // (Y, T1, T2, -) per (entry, lm)
        fr = ytab + E * L2 * 4;      // f R per (entry, l, c)
        rab = fr + E * L * 8;        // (ra, rb) per (entry, l, c)
        geo = rab + E * L * 16;      // d, ga, gb per entry
        csp = L2 * 8 + 8;            // per type, padded: two types' column reads meet no bank conflict
        cval = (geo + E * 12 + 3) / 4 * 4;  // the center's coefficients [type][lm][c]
        total = cval + SC * csp;
    }
};

// Power-spectrum values of one center from its coefficients C in LDS (all types,
// padded rows of `csp`), written into the samples of its present type pairs
// (soap/power_spectrum.rs:340-348): one v_mfma_f32_16x16x4f32 product per (l, 2 x 2
// block of type pairs, 4 m), rows (t2, n2) and columns (t1, n1), so that lane j holds
// PS[l][n1][4 h .. 4 h + 3] of the pair (t1, t2), a 16-byte piece of its sample.
// With `bf16`, the products over m are split into bfloat16 hi + lo parts
// (hi hi + lo hi + hi lo, FP32 accumulation) on gfx950's v_mfma_f32_16x16x32_bf16: two
// products of K = 32 per 16 values of m, the hi and lo parts of A side by side in K.
__device__ __forceinline__ void expansion_ps_values(int wave, int n_waves, int lane, int L, int S, const float* C,
                                                    int csp, const long long* pair_row, bool bf16,
                                                    float* __restrict__ out) {
    const int nb = (S + 1) / 2;
    const int nbp = nb * (nb + 1) / 2;
    const int P = L * 64;
    const int j16 = lane & 15, jk = lane >> 4;
    for (int u = wave; u < L * nbp; u += n_waves) {
        const int bp = u % nbp, l = u / nbp;
        int b1 = 0, rem = bp;
        while (rem >= nb - b1) {
            rem -= nb - b1;
            b1++;
        }
        const int b2 = b1 + rem;
        const int k2 = 2 * b2 + (j16 >> 3), k1 = 2 * b1 + (j16 >> 3);
        const bool kv2 = k2 < S, kv1 = k1 < S;
        const float* C2 = C + (kv2 ? k2 : 0) * csp + l * l * 8 + (lane & 7);
        const float* C1 = C + (kv1 ? k1 : 0) * csp + l * l * 8 + (lane & 7);
        float4_t acc = {0.f, 0.f, 0.f, 0.f};
        if (bf16) {
            // v_mfma_f32_16x16x32_bf16, per 16 values of m (two passes at l = 8): lane i + 16 g
            // holds K values 8 g .. 8 g + 7 of row / column i. K < 16 carries the hi parts of
            // m = 16 kb + K, K >= 16 the lo parts (A) or again the hi parts (B) of m = 16 kb + K
            // - 16, so that one product gives hi hi + lo hi; a second, with the lo parts of B
            // against the hi parts of A on K < 16, gives hi lo.
            const bool upper = jk >= 2;
            for (int kb = 0; kb < (2 * l + 1 + 15) / 16; kb++) {
                short8_t a1, b1, a2, b2;
#pragma unroll
                for (int q = 0; q < 8; q++) {
                    const int mp = 16 * kb + 8 * (jk & 1) + q;
                    const bool mv = mp <= 2 * l;
                    const int off = (mv ? mp : 0) * 8;
                    short ah, al, bh, bl;
                    bf16_split((mv && kv2) ? C2[off] : 0.f, ah, al);
                    bf16_split((mv && kv1) ? C1[off] : 0.f, bh, bl);
                    a1[q] = upper ? al : ah;
                    b1[q] = bh;
                    a2[q] = upper ? (short)0 : ah;
                    b2[q] = upper ? (short)0 : bl;
                }
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8_t, a1),
                                                              __builtin_bit_cast(bf16x8_t, b1), acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8_t, a2),
                                                              __builtin_bit_cast(bf16x8_t, b2), acc, 0, 0, 0);
            }
        } else {
            for (int ks = 0; ks < (2 * l + 1 + 3) / 4; ks++) {
                const int mp = 4 * ks + jk;
                const bool mv = mp <= 2 * l;
                const int off = (mv ? mp : 0) * 8;
                acc = mfma16x16((mv && kv2) ? C2[off] : 0.f, (mv && kv1) ? C1[off] : 0.f, acc);
            }
        }
        // lane j: rows 4 (j / 16) + v = (t2 = 2 b2 + j / 32, n2 = 4 ((j / 16) % 2) + v),
        // column j % 16 = (t1 = 2 b1 + (j % 16) / 8, n1 = j % 8)
        const int t2 = 2 * b2 + (lane >> 5), t1 = 2 * b1 + (j16 >> 3);
        const long long row = (t2 < S && t1 <= t2) ? pair_row[pair_index(t1, t2, S)] : -1;
        if (row >= 0) {
            const float kappa = t1 != t2 ? 1.4142135623730951f : 1.f;
            const float norm = (l % 2 == 0 ? 1.f : -1.f) * (float)sqrt((double)(2 * l + 1));
            float* dst = out + row * P + l * 64 + (lane & 7) * 8 + 4 * ((lane >> 4) & 1);
            store4<float>(dst, (acc[0] * kappa) / norm, (acc[1] * kappa) / norm, (acc[2] * kappa) / norm,
                          (acc[3] * kappa) / norm);
        }
    }
}

// 16-block halves of v_mfma_f32_4x4x1f32 needed for the (l, m quad, c quad) blocks
__host__ __device__ inline int expansion_mfma_halves(int L) {
    int q = 0;
    for (int l = 0; l < L; l++) q += (2 * l + 1 + 3) / 4;
    return (2 * q + 15) / 16;
}

template <bool B>
struct BoolTag {
    static constexpr bool value = B;
};

template <int E, int NH>
__global__ void __launch_bounds__(EXM_THREADS) __attribute__((amdgpu_waves_per_eu(E == 8 && NH <= 2 ? 5 : 1, 8)))
expansion_mfma_kernel(
    SoapArgs<float> a, const float* __restrict__ entry_data, float* __restrict__ out_coef, float* __restrict__ out_grad,
    const unsigned* __restrict__ type_mask, const long long* __restrict__ sample_flag,
    const long long* __restrict__ sample_scan, float* __restrict__ ps_values, int bf16_values) {
    extern __shared__ __align__(16) unsigned char smem[];
    // the chunk's entries grouped by type within the group of four types (values), and
    // the same without the center's own periodic images (gradients)
    __shared__ int vorder[E];
    __shared__ int gorder[E];
    __shared__ int vstart[5];
    __shared__ int gstart[5];
    float* lds = reinterpret_cast<float*>(smem);
    // records of up to E entries at l <= 8 (204 floats each), in float4 pieces per thread
    constexpr int NV = (E * 204 / 4 + EXM_THREADS - 1) / EXM_THREADS;
    const int i = blockIdx.x;
    const int tid = threadIdx.x;
    const int lane = tid & (WAVE - 1);
    const int comp = uniform(tid / WAVE);  // 0: values, 1 + x: self gradient along x
    const int S = a.n_species, L = a.L;
    const int L2 = L * L;
    const int slots = L2 * 8;
    // with `ps_values`, the power-spectrum values follow from the coefficients in LDS
    const bool fused = ps_values != nullptr;
    const ExpansionMfmaLayout ly(L, E, fused ? S : 0);
    const TileEntryLayout lay(L, 8, 8);
    const ShTable tb(L);
    const bool grads = out_grad != nullptr;
    const long long e0 = a.offsets[i];
    const int len = (int)(a.offsets[i + 1] - e0);
    const int ct = a.species[i];
    float* raw = lds + ly.raw;
    float* ytab = lds + ly.ytab;
    float* fr = lds + ly.fr;
    float* rab = lds + ly.rab;
    float* geo = lds + ly.geo;
    // the spherical-harmonics coefficients, and the sample row of each type pair's
    // power-spectrum values (-1: no sample), in LDS before the first chunk's barrier
    __shared__ float shc[EXM_MAX_L * (EXM_MAX_L + 1) + 4 * EXM_MAX_L];
    __shared__ long long pair_row[EXM_MAX_VALUE_TYPES * (EXM_MAX_VALUE_TYPES + 1) / 2];
    for (int k = tid; k < tb.size; k += EXM_THREADS) shc[k] = a.sh_c[k];
    if (fused && tid < S * (S + 1) / 2) {
        int t1 = 0, rem = tid;
        while (rem >= S - t1) {
            rem -= S - t1;
            t1++;
        }
        const int t2 = t1 + rem;
        const unsigned mask = type_mask[i];
        const long long idx = (long long)(ct * (S * (S + 1) / 2) + tid) * a.n_atoms + i;
        const long long flag = sample_flag[idx], row = sample_scan[idx];
        pair_row[tid] = (((mask >> t1) & 1u) && ((mask >> t2) & 1u) && flag) ? row : -1;
    }

    // this lane's blocks: block b = 16 h + lane / 4 enumerates (l, m quad, c quad) with
    // the c quad fastest; the A operand is the radial value c = 4 cq + lane % 4, the B
    // operand the angular value of m index 4 mq + lane % 4, which is also the row of
    // the lane's accumulators
    int roff[NH], lmh[NH], cqh[NH];
    bool valid[NH];
#pragma unroll
    for (int h = 0; h < NH; h++) {
        const int b = 16 * h + (lane >> 2);
        roff[h] = 0;
        lmh[h] = 0;
        cqh[h] = 0;
        valid[h] = false;
        int first = 0;
        for (int l = 0; l < L; l++) {
            const int nb = 2 * ((2 * l + 1 + 3) / 4);
            if (b >= first && b < first + nb) {
                const int mq = (b - first) >> 1, cq = (b - first) & 1;
                const int mi = 4 * mq + (lane & 3);
                roff[h] = l * 8 + 4 * cq + (lane & 3);
                valid[h] = mi <= 2 * l;
                lmh[h] = valid[h] ? l * l + mi : 0;
                cqh[h] = cq;
            }
            first += nb;
        }
    }

    float4_t pre[NV];
    auto fetch = [&](int cb) {
        const int c = min(E, len - cb);
        const float4_t* src = reinterpret_cast<const float4_t*>(entry_data + (e0 + cb) * ly.per_entry);
        const int count = c * ly.per_entry / 4;
#pragma unroll
        for (int k = 0; k < NV; k++) {
            const int idx = k * EXM_THREADS + tid;
            if (idx < count) pre[k] = src[idx];
        }
    };

    // one instantiation per role, so that the accumulators of each stay in place
    auto centre = [&](auto role) {
        constexpr bool VAL = decltype(role)::value;  // values (wave 0) or self gradients
        const int n_groups = (S + 3) / 4;
        for (int g = 0; g < n_groups; g++) {
            float4_t acc[4][NH];
#pragma unroll
            for (int k = 0; k < 4; k++) {
#pragma unroll
                for (int h = 0; h < NH; h++) acc[k][h] = float4_t{0.f, 0.f, 0.f, 0.f};
            }
            if (len > 0) fetch(0);
            for (int cb = 0; cb < len; cb += E) {
                const int cnt = min(E, len - cb);
                lds_barrier();  // the previous chunk's tables are read no more
                {
                    float4_t* dst = reinterpret_cast<float4_t*>(raw);
                    const int count = cnt * ly.per_entry / 4;
#pragma unroll
                    for (int k = 0; k < NV; k++) {
                        const int idx = k * EXM_THREADS + tid;
                        if (idx < count) dst[idx] = pre[k];
                    }
                }
                if (cb + E < len) fetch(cb + E);
                lds_barrier();
                if (tid < WAVE) {
                    // wave 0: positions from ballots, entry `tid` of type 4 g + k after the
                    // chunk's entries of types below it and its own type's earlier entries
                    const bool have = tid < cnt;
                    const float* rec = raw + (have ? tid : 0) * ly.per_entry;
                    const int t = have ? (int)rec[lay.dir + 3] - 4 * g : -1;
                    const bool own = have && __float_as_int(rec[lay.scal + 3]) == i;
                    const unsigned long long below = (1ull << tid) - 1ull;
                    int vbase = 0, gbase = 0;
#pragma unroll
                    for (int k = 0; k < 4; k++) {
                        const unsigned long long vm = __ballot(t == k);
                        const unsigned long long gm = __ballot(t == k && !own);
                        if (t == k) {
                            vorder[vbase + __popcll(vm & below)] = tid;
                            if (!own) gorder[gbase + __popcll(gm & below)] = tid;
                        }
                        if (tid == 0) {
                            vstart[k] = vbase;
                            gstart[k] = gbase;
                        }
                        vbase += __popcll(vm);
                        gbase += __popcll(gm);
                    }
                    if (tid == 0) {
                        vstart[4] = vbase;
                        gstart[4] = gbase;
                    }
                }
                // (Y, T1, T2) per (entry, l m)
#pragma unroll 1
                for (int it = tid; it < cnt * L2; it += EXM_THREADS) {
                    const int e = it / L2, lmp = it - e * L2;
                    const int l = l_of_lm(lmp);
                    const int m = lmp - l * l - l;
                    const int am = m < 0 ? -m : m;
                    const float* base = raw + e * ly.per_entry;
                    const float* Pl = base + lay.p;
                    const float cos_theta = base[lay.trig + 0], sin_theta = base[lay.trig + 1];
                    const float cos_phi = base[lay.trig + 2], sin_phi = base[lay.trig + 3];
                    const int k = leg_index(l, am);
                    const float p = Pl[k];
                    float Y, T1, T2;
                    if (m == 0) {
                        Y = p / 1.4142135623730951f;
                        T1 = 0.f;
                        T2 = l == 0 ? 0.f : -2.f * (shc[tb.gl + l] * Pl[leg_index(l, 1)]);
                    } else {
                        // cos (m phi), sin (m phi) by the recurrence of spherical_harmonics.rs
                        float cos_1 = 1, sin_1 = 0, cos_2 = -cos_phi, sin_2 = sin_phi;
                        const float minus_two_cos = -2 * cos_phi;
                        float sin_m = 0, cos_m = 1;
                        for (int mm = 1; mm <= am; mm++) {
                            sin_m = minus_two_cos * sin_1 - sin_2;
                            cos_m = minus_two_cos * cos_1 - cos_2;
                            sin_2 = sin_1;
                            sin_1 = sin_m;
                            cos_2 = cos_1;
                            cos_1 = cos_m;
                        }
                        const float pm1 = Pl[leg_index(l, am - 1)];
                        const float pp1 = am < l ? Pl[leg_index(l, am + 1)] : 0.f;
                        const float dp = shc[tb.dpa + k] * pm1 - shc[tb.dpb + k] * pp1;
                        const float pt = (sin_theta > 0.1f) ? (float)am / sin_theta * p : -0.5f / cos_theta * dp;
                        const float c1 = m > 0 ? cos_m : sin_m;
                        const float s1 = m > 0 ? sin_m : -cos_m;
                        Y = p * c1;
                        T1 = pt * s1;
                        T2 = c1 * dp;
                    }
                    store4<float>(ytab + (size_t)it * 4, Y, T1, T2, 0.f);
                }
                // f R and (ra, rb) per (entry, l, c)
#pragma unroll 1
                for (int it = tid; it < cnt * L * 8; it += EXM_THREADS) {
                    const int e = it / (L * 8), lc = it - e * L * 8;
                    const float* base = raw + e * ly.per_entry;
                    const float f = base[lay.scal + 0], df = base[lay.scal + 1], f_r = base[lay.scal + 2];
                    const float R = base[lay.rv + lc], dR = base[lay.rd + lc];
                    fr[it] = f * R;
                    rab[2 * it + 0] = df * R + f * dR;
                    rab[2 * it + 1] = f_r * R;
                }
                // d, ga, gb per entry
                for (int e = tid; e < cnt; e += EXM_THREADS) {
                    const float* base = raw + e * ly.per_entry;
                    const float cos_theta = base[lay.trig + 0], sin_theta = base[lay.trig + 1];
                    const float cos_phi = base[lay.trig + 2], sin_phi = base[lay.trig + 3];
                    float* gm = geo + e * 12;
                    gm[0] = base[lay.dir + 0];
                    gm[1] = base[lay.dir + 1];
                    gm[2] = base[lay.dir + 2];
                    gm[3] = sin_phi;
                    gm[4] = -cos_phi;
                    gm[5] = 0.f;
                    gm[6] = -0.5f * cos_theta * cos_phi;
                    gm[7] = -0.5f * cos_theta * sin_phi;
                    gm[8] = 0.5f * sin_theta;
                }
                lds_barrier();
                // the type is a constant of each unrolled k, so each set of accumulators
                // stays in its registers; the chunk's order and type starts are read once
                // (lane q holds entry q of the order)
                const int ord = lane < E ? (VAL ? vorder[lane] : gorder[lane]) : 0;
                const int starts = lane < 5 ? (VAL ? vstart[lane] : gstart[lane]) : 0;
                if constexpr (VAL) {
#pragma unroll
                    for (int k = 0; k < 4; k++) {
                        const int lo = __builtin_amdgcn_readlane(starts, k);
                        const int hi = __builtin_amdgcn_readlane(starts, k + 1);
                        // (two entries per pass: the second's operand reads issue before the
                        // first's products)
#pragma unroll 2
                        for (int q = lo; q < hi; q++) {
                            const int e = __builtin_amdgcn_readlane(ord, q);
#pragma unroll
                            for (int h = 0; h < NH; h++) {
                                const float y = ytab[(size_t)(e * L2 + lmh[h]) * 4];
                                const float av = fr[e * L * 8 + roff[h]];
                                acc[k][h] = mfma4x4(av, valid[h] ? y : 0.f, acc[k][h]);
                            }
                        }
                    }
                } else if (grads) {
                    const int x = comp - 1;  // waves 1 to 3
#pragma unroll
                    for (int k = 0; k < 4; k++) {
                        const int lo = __builtin_amdgcn_readlane(starts, k);
                        const int hi = __builtin_amdgcn_readlane(starts, k + 1);
#pragma unroll 2
                        for (int q = lo; q < hi; q++) {
                            const int e = __builtin_amdgcn_readlane(ord, q);
                            const float dx = geo[e * 12 + x], gax = geo[e * 12 + 3 + x], gbx = geo[e * 12 + 6 + x];
#pragma unroll
                            for (int h = 0; h < NH; h++) {
                                const float* y4 = ytab + (size_t)(e * L2 + lmh[h]) * 4;
                                const float Y = y4[0], T1 = y4[1], T2 = y4[2];
                                const float* r2 = rab + 2 * (e * L * 8 + roff[h]);
                                acc[k][h] = mfma4x4(r2[0], valid[h] ? Y * dx : 0.f, acc[k][h]);
                                acc[k][h] = mfma4x4(r2[1], valid[h] ? gax * T1 + gbx * T2 : 0.f, acc[k][h]);
                            }
                        }
                    }
                }
            }
            // the group's sums; the self contribution goes to the l = 0 values of the
            // center's own type (spherical_expansion_pair.rs:222-252)
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int t = 4 * g + k;
                if (t >= S) break;
#pragma unroll
                for (int h = 0; h < NH; h++) {
                    if (!valid[h]) continue;
                    float4_t v = acc[k][h];
                    if constexpr (VAL) {
                        if (t == ct && lmh[h] == 0) {
#pragma unroll
                            for (int q = 0; q < 4; q++) v[q] += a.self_contribution[4 * cqh[h] + q];
                        }
                        store4<float>(out_coef + ((long long)i * S + t) * slots + lmh[h] * 8 + 4 * cqh[h], v[0], v[1],
                                      v[2], v[3]);
                        if (fused) store4<float>(lds + ly.cval + t * ly.csp + lmh[h] * 8 + 4 * cqh[h], v[0], v[1], v[2], v[3]);
                    } else if (grads) {
                        store4<float>(out_grad + (((long long)i * S + t) * 3 + comp - 1) * slots + lmh[h] * 8 + 4 * cqh[h],
                                      -v[0], -v[1], -v[2], -v[3]);
                    }
                }
            }
        }
    };
    if (comp == 0) {
        centre(BoolTag<true>{});
    } else {
        centre(BoolTag<false>{});
    }
    if (fused) {
        lds_barrier();  // the coefficients of every type are in LDS
        expansion_ps_values(comp, EXM_THREADS / WAVE, lane, L, S, lds + ly.cval, ly.csp, pair_row, bf16_values != 0,
                            ps_values);
    }
}

// ---------------------------------------------------------------------------
// Vector-Jacobian product of the position gradients without the gradient rows:
// dA/dr_atom = Σ_blocks Σ_rows (dX/dr)[row] · (dA/dX)[sample of row], the
// contraction of featomic-torch's PositionsGrad (featomic-torch/src/
// autograd.cpp:237-300). With H_i^t = dA/dc_i^t (the loss gradient with respect to
// the expansion coefficients), each neighbour entry (i, j) contributes
// g_ij = Σ_{lm, n} (d c_i^{t_j} / d r_j)[x][lm][n] H_i^{t_j}[lm][n] to dA/dr_j and
// -g_ij to dA/dr_i (translation invariance of c_i; soap/spherical_expansion.rs:
// 276-354 for the per-pair derivative and its self counterpart).

// H from dA/dPS: for each block (t1 <= t2) with a sample at this center,
// dPS[l n1 n2] = κ / ((-1)^l √(2l+1)) Σ_m c^{t1}[lm n1] c^{t2}[lm n2] gives
// H^{t1}[lm n1] += Σ_{n2} s W[l n1 n2] c^{t2}[lm n2] and
// H^{t2}[lm n2] += Σ_{n1} s W[l n1 n2] c^{t1}[lm n1] (soap/power_spectrum.rs:529-618).
template <typename T>
__global__ void __launch_bounds__(TILE_THREADS) ps_vjp_coef_kernel(
    int n_atoms, int S, int L, int n, const T* __restrict__ coef, const int* __restrict__ species,
    const unsigned* __restrict__ type_mask, const long long* __restrict__ sample_flag,
    const long long* __restrict__ sample_scan, const T* __restrict__ grad_values, T* __restrict__ H) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int i = blockIdx.x, tid = threadIdx.x;
    const int L2 = L * L, slots = L2 * n, P = L * n * n;
    T* C = reinterpret_cast<T*>(smem);  // [S][L2][n]
    T* Hs = C + (size_t)S * slots;      // [S][L2][n]
    T* W = Hs + (size_t)S * slots;      // [P]
    for (int k = tid; k < S * slots; k += TILE_THREADS) {
        C[k] = coef[(long long)i * S * slots + k];
        Hs[k] = 0;
    }
    const int ct = species[i];
    const unsigned mask = type_mask[i];
    const int npairs = S * (S + 1) / 2;
    for (int t1 = 0; t1 < S; t1++) {
        if (!((mask >> t1) & 1u)) continue;
        for (int t2 = t1; t2 < S; t2++) {
            if (!((mask >> t2) & 1u)) continue;
            const long long idx = (long long)(ct * npairs + pair_index(t1, t2, S)) * n_atoms + i;
            if (!sample_flag[idx]) continue;
            const long long smp = sample_scan[idx];
            __syncthreads();
            for (int k = tid; k < P; k += TILE_THREADS) W[k] = grad_values[smp * P + k];
            __syncthreads();
            const T kappa = t1 != t2 ? (T)1.4142135623730951 : (T)1;
            for (int item = tid; item < slots; item += TILE_THREADS) {
                const int lm = item / n, c = item - lm * n;
                const int l = l_of_lm(lm);
                const T sc = kappa * (T)((l % 2 == 0 ? 1.0 : -1.0) / sqrt(2.0 * l + 1.0));
                const T* Wl = W + l * n * n;
                T h1 = 0, h2 = 0;
                for (int o = 0; o < n; o++) {
                    h1 += Wl[c * n + o] * C[t2 * slots + lm * n + o];
                    h2 += Wl[o * n + c] * C[t1 * slots + lm * n + o];
                }
                // the same thread owns (lm, c) of both types: no race when t1 == t2
                Hs[t1 * slots + item] += sc * h1;
                Hs[t2 * slots + item] += sc * h2;
            }
        }
    }
    __syncthreads();
    for (int k = tid; k < S * slots; k += TILE_THREADS) H[(long long)i * S * slots + k] = Hs[k];
}

// H from dA/dc given per spherical-expansion block (l, center type, neighbour type):
// a gather from the blocks' [samples][2l+1][n] layout into [atom][type][L2][n].
template <typename T>
__global__ void se_vjp_coef_kernel(int n_atoms, int S, int L, int n, const int* __restrict__ species,
                                   const long long* __restrict__ sample_flag, const long long* __restrict__ sample_scan,
                                   const long long* __restrict__ block_sample_base,
                                   const long long* __restrict__ value_base, const T* __restrict__ grad_values,
                                   T* __restrict__ H) {
    const int L2 = L * L, slots = L2 * n;
    const long long total = (long long)n_atoms * S * slots;
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < total; k += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(k / ((long long)S * slots));
        const int r = (int)(k - (long long)i * S * slots);
        const int nt = r / slots, s = r - nt * slots;
        const int lm = s / n, c = s - lm * n;
        const int l = l_of_lm(lm);
        const long long tb = (long long)species[i] * S + nt;
        const long long idx = tb * n_atoms + i;
        T v = 0;
        if (sample_flag[idx]) {
            const long long local = sample_scan[idx] - block_sample_base[tb];
            v = grad_values[value_base[tb * L + l] + local * (2 * l + 1) * n + (lm - l * l) * n + c];
        }
        H[k] = v;
    }
}

// g_e[x] for every neighbour entry e = (i, j): one workgroup per center, the
// entries in chunks; angular factors as in the expansion, then per (entry, l m)
// A1 = Σ_n (df R + f dR)[l n] H[lm n] and A2 = Σ_n (f / r) R[l n] H[lm n], and per
// (entry, x) g = d_x Σ_lm Y A1 + Σ_lm (r dY/dx) A2 in a fixed order. Periodic
// images of the center itself move with it: g = 0.
template <typename T, int NC>
__global__ void __launch_bounds__(TILE_THREADS) vjp_pair_kernel(SoapArgs<T> a, int max_entries,
                                                                const T* __restrict__ entry_data, const T* __restrict__ H,
                                                                T* __restrict__ g_out) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int chunk_t[WAVE];
    __shared__ int chunk_self[WAVE];
    __shared__ int leg_l[TILE_MAX_L * (TILE_MAX_L + 1) / 2];
    __shared__ int leg_m[TILE_MAX_L * (TILE_MAX_L + 1) / 2];
    const int i = blockIdx.x, tid = threadIdx.x;
    const int S = a.n_species, L = a.L, n = a.n;
    const int L2 = L * L;
    const TileEntryLayout lay(L, n, NC);
    T* Hc = reinterpret_cast<T*>(smem);                 // [S][L2][NC]
    T* ent = Hc + (size_t)S * L2 * NC;                   // [max_entries][per_entry]
    T* ang = ent + (size_t)max_entries * lay.per_entry;  // [max_entries][4][L2]
    T* part = ang + (size_t)max_entries * 4 * L2;        // [max_entries][L2][4]
    const long long e0 = a.offsets[i];
    const int len = (int)(a.offsets[i + 1] - e0);
    for (int k = tid; k < S * L2 * NC; k += TILE_THREADS) {
        const int t = k / (L2 * NC), r = k - t * L2 * NC;
        const int lm = r / NC, c = r - lm * NC;
        Hc[k] = c < n ? H[((long long)i * S + t) * L2 * n + lm * n + c] : (T)0;
    }
    for (int k = tid; k < lay.leg; k += TILE_THREADS) {
        int l = 0;
        while ((l + 1) * (l + 2) / 2 <= k) l++;
        leg_l[k] = l;
        leg_m[k] = k - l * (l + 1) / 2;
    }
    for (int cb = 0; cb < len; cb += max_entries) {
        const int cnt = min(max_entries, len - cb);
        lds_barrier();
        {
            typedef typename Vec4Of<T>::type V;
            constexpr int VEC = 16 / sizeof(T);
            const V* src = reinterpret_cast<const V*>(entry_data + (e0 + cb) * lay.per_entry);
            V* dst = reinterpret_cast<V*>(ent);
            const int count = cnt * lay.per_entry / VEC;
            for (int k = tid; k < count; k += TILE_THREADS) dst[k] = src[k];
        }
        for (int c = tid; c < cnt; c += TILE_THREADS) {
            const int j = a.entry_j[e0 + cb + c];
            chunk_t[c] = a.species[j];
            chunk_self[c] = j == i ? 1 : 0;
        }
        lds_barrier();
        chunk_angular_factors<T>(a, lay, cnt, leg_l, leg_m, ent, ang, true);
        lds_barrier();
        for (int item = tid; item < cnt * L2; item += TILE_THREADS) {
            const int e = item / L2, lm = item - e * L2;
            const int l = l_of_lm(lm);
            const T* base = ent + e * lay.per_entry;
            const T* Ht = Hc + ((size_t)chunk_t[e] * L2 + lm) * NC;
            const T f = base[lay.scal + 0], df = base[lay.scal + 1], f_r = base[lay.scal + 2];
            const T* Rv = base + lay.rv + l * NC;
            const T* dR = base + lay.rd + l * NC;
            T a1 = 0, a2 = 0;
#pragma unroll
            for (int c = 0; c < NC; c++) {
                a1 += (df * Rv[c] + f * dR[c]) * Ht[c];
                a2 += (f_r * Rv[c]) * Ht[c];
            }
            const T* A = ang + (size_t)e * 4 * L2;
            T* pp = part + ((size_t)e * L2 + lm) * 4;
            pp[0] = A[lm] * a1;
            pp[1] = A[L2 + lm] * a2;
            pp[2] = A[2 * L2 + lm] * a2;
            pp[3] = A[3 * L2 + lm] * a2;
        }
        lds_barrier();
        for (int item = tid; item < cnt * 3; item += TILE_THREADS) {
            const int e = item / 3, x = item - e * 3;
            T g = 0;
            if (!chunk_self[e]) {
                const T* base = ent + e * lay.per_entry;
                const T* pe = part + (size_t)e * L2 * 4;
                T sy = 0, sg = 0;
                for (int lm = 0; lm < L2; lm++) {
                    sy += pe[lm * 4];
                    sg += pe[lm * 4 + 1 + x];
                }
                g = base[lay.dir + x] * sy + sg;
            }
            g_out[(e0 + cb + e) * 3 + x] = g;
        }
    }
}

// dA/dr_j = Σ over j's entries e = (j, k, s) of g[mirror of e] - g[e], where the
// mirror (k, j, -s) is found in k's entries (sorted by neighbour); fixed order,
// FP64 sums. An entry without a mirror (a neighbour list that is not symmetric)
// makes the atom's result NaN.
template <typename T>
__global__ void vjp_gather_kernel(int n_atoms, const long long* __restrict__ offsets, const int* __restrict__ entry_j,
                                  const int* __restrict__ entry_shift, const T* __restrict__ g, T* __restrict__ out) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n_atoms) return;
    double acc[3] = {0.0, 0.0, 0.0};
    for (long long e = offsets[j]; e < offsets[j + 1]; e++) {
        const int k = entry_j[e];
        if (k == j) continue;
        const int s0 = entry_shift[3 * e], s1 = entry_shift[3 * e + 1], s2 = entry_shift[3 * e + 2];
        // first entry of k's list with neighbour j
        long long lo = offsets[k], hi = offsets[k + 1];
        while (lo < hi) {
            const long long mid = (lo + hi) / 2;
            if (entry_j[mid] < j) lo = mid + 1; else hi = mid;
        }
        long long m = -1;
        for (long long q = lo; q < offsets[k + 1] && entry_j[q] == j; q++) {
            if (entry_shift[3 * q] == -s0 && entry_shift[3 * q + 1] == -s1 && entry_shift[3 * q + 2] == -s2) {
                m = q;
                break;
            }
        }
        if (m < 0) {
            acc[0] = acc[1] = acc[2] = __builtin_nan("");
            break;
        }
        for (int x = 0; x < 3; x++) acc[x] += (double)g[3 * m + x] - (double)g[3 * e + x];
    }
    for (int x = 0; x < 3; x++) out[3 * (long long)j + x] = (T)acc[x];
}

// Power spectrum strain / cell gradients: one row of [3][3][P] per sample.
template <typename T>
__global__ void __launch_bounds__(WAVE) ps_cell_strain_kernel(int n_atoms, int S, int L, int n, const T* __restrict__ coef,
                                                               const T* __restrict__ dcs, const int* __restrict__ species,
                                                               const unsigned* __restrict__ type_mask,
                                                               const long long* __restrict__ sample_flag,
                                                               const long long* __restrict__ sample_scan, T* __restrict__ out) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int i = blockIdx.x;
    const int slots = L * L * n;
    T* C = reinterpret_cast<T*>(smem);
    T* D1 = C + S * slots;
    T* D2 = D1 + 9 * slots;
    for (int k = threadIdx.x; k < S * slots; k += WAVE) C[k] = coef[(long long)i * S * slots + k];
    __syncthreads();
    const int ct = species[i];
    const unsigned mask = type_mask[i];
    const int P = L * n * n;
    const int npairs = S * (S + 1) / 2;
    for (int t1 = 0; t1 < S; t1++) {
        if (!((mask >> t1) & 1u)) continue;
        for (int t2 = t1; t2 < S; t2++) {
            if (!((mask >> t2) & 1u)) continue;
            const long long idx = (long long)(ct * npairs + pair_index(t1, t2, S)) * n_atoms + i;
            if (!sample_flag[idx]) continue;
            const long long s = sample_scan[idx];
            for (int k = threadIdx.x; k < 9 * slots; k += WAVE) {
                D1[k] = dcs[((long long)i * S + t1) * 9 * slots + k];
                D2[k] = dcs[((long long)i * S + t2) * 9 * slots + k];
            }
            __syncthreads();
            const T kappa = t1 != t2 ? (T)1.4142135623730951 : (T)1;
            ps_contract<T>(9, L, n, C + t1 * slots, C + t2 * slots, D1, D2, true, true, kappa, slots, out + s * 9 * P);
            __syncthreads();
        }
    }
}

// Spherical expansion values / cell / strain: gather from the dense layouts into
// per-(type block, l) regions.
template <typename T>
__global__ void se_gather_kernel(int n_atoms, int S, int L, int n, int ncomp, const T* __restrict__ dense,
                                 const int* __restrict__ species, const long long* __restrict__ sample_flag,
                                 const long long* __restrict__ sample_scan, const long long* __restrict__ block_sample_base,
                                 const long long* __restrict__ region_base, T* __restrict__ out) {
    // one workgroup per (type block, atom)
    const long long idx = blockIdx.x;
    if (!sample_flag[idx]) return;
    const int tb = (int)(idx / n_atoms), i = (int)(idx % n_atoms);
    const int nt = tb % S;
    const long long s = sample_scan[idx] - block_sample_base[tb];
    const int slots = L * L * n;
    for (int l = 0; l < L; l++) {
        const int w = (2 * l + 1) * n;
        T* o = out + region_base[(long long)tb * L + l] + s * ncomp * w;
        for (int k = threadIdx.x; k < ncomp * w; k += blockDim.x) {
            const int x = k / w, r = k - x * w;
            o[k] = dense[(((long long)i * S + nt) * ncomp + x) * slots + l * l * n + r];
        }
    }
}

template <typename T>
SoapArgs<T> make_args(const SoapLaunch<T>& p) {
    SoapArgs<T> a;
    a.n_atoms = p.n_atoms;
    a.n_species = p.n_species;
    a.L = p.L;
    a.n = p.n;
    a.offsets = p.offsets;
    a.entry_j = p.entry_j;
    a.entry_shift = p.entry_shift;
    a.entry_vec = p.entry_vec;
    a.entry_dist = p.entry_dist;
    a.species = p.species;
    a.radial = p.radial;
    a.scalars = p.scalars;
    a.sh_a = p.sh_a;
    a.sh_b = p.sh_b;
    a.sh_c = p.sh_c;
    a.self_contribution = p.self_contribution;
    a.chunk = p.chunk;
    a.error_flag = p.error_flag;
    return a;
}

template <typename T, int MS, int MODE>
void launch_center_ms(const SoapLaunch<T>& p, T* coef, T* grad) {
    SoapArgs<T> a = make_args(p);
    const size_t shmem = (size_t)ChunkLayout(p.L, p.n, p.chunk).per_entry * p.chunk * sizeof(T);
    soap_center_kernel<T, MS, MODE><<<p.n_atoms, WAVE, shmem, p.stream>>>(a, coef, grad);
}

// Exact GTO radial integrals (spline_accuracy = null): one thread per (entry, l),
// FP64 hypergeometric evaluation, stored per entry for the SOAP kernels.
template <typename T>
__global__ void radial_exact_kernel(long long n_entries, int L, int n, const double* __restrict__ entry_dist,
                                    double sigma2_atomic, const double* __restrict__ gto, const double* __restrict__ ortho,
                                    T* __restrict__ out_val, T* __restrict__ out_der) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n_entries * L) return;
    const long long e = idx / L;
    const int l = (int)(idx - e * L);
    const double r = entry_dist[e];
    double raw_v[64], raw_g[64];
    for (int k = 0; k < n; k++) soap_gto_raw(l, k, r, sigma2_atomic, gto[k], &raw_v[k], &raw_g[k]);
    for (int j = 0; j < n; j++) {
        double v = 0.0, g = 0.0;
        for (int k = 0; k < n; k++) {
            v += raw_v[k] * ortho[j * n + k];
            g += raw_g[k] * ortho[j * n + k];
        }
        out_val[idx * n + j] = (T)v;
        out_der[idx * n + j] = (T)g;
    }
}

// Radial integral (spline or precomputed values) at arbitrary distances: one
// thread per (distance, l), output [i][L][n].
template <typename T>
__global__ void radial_eval_kernel(long long count, DeviceRadial<T> rad, const double* __restrict__ r,
                                   T* __restrict__ out_val, T* __restrict__ out_der) {
    const int L = rad.max_angular + 1;
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= count * L) return;
    const long long i = idx / L;
    const int l = (int)(idx - i * L);
    radial_integral<T>(rad, l, i, r[i], out_val + idx * rad.n, out_der + idx * rad.n);
}

}  // namespace

template <typename T>
void launch_radial_eval(hipStream_t stream, long long count, const DeviceRadial<T>& radial, const double* r,
                        T* out_val, T* out_der) {
    const long long total = count * (radial.max_angular + 1);
    if (total == 0) return;
    radial_eval_kernel<T><<<(unsigned)((total + 127) / 128), 128, 0, stream>>>(count, radial, r, out_val, out_der);
}
template void launch_radial_eval<float>(hipStream_t, long long, const DeviceRadial<float>&, const double*, float*,
                                        float*);
template void launch_radial_eval<double>(hipStream_t, long long, const DeviceRadial<double>&, const double*, double*,
                                         double*);

template <typename T>
void launch_radial_exact(hipStream_t stream, long long n_entries, int L, int n, const double* entry_dist,
                         double sigma2_atomic, const double* gto, const double* ortho, T* out_val, T* out_der) {
    const long long total = n_entries * L;
    if (total == 0) return;
    radial_exact_kernel<T><<<(unsigned)((total + 127) / 128), 128, 0, stream>>>(n_entries, L, n, entry_dist, sigma2_atomic,
                                                                               gto, ortho, out_val, out_der);
}
template void launch_radial_exact<float>(hipStream_t, long long, int, int, const double*, double, const double*,
                                         const double*, float*, float*);
template void launch_radial_exact<double>(hipStream_t, long long, int, int, const double*, double, const double*,
                                          const double*, double*, double*);

namespace {
}  // namespace

template <typename T>
size_t soap_chunk_bytes(int L, int n, int chunk) {
    return (size_t)ChunkLayout(L, n, chunk).per_entry * chunk * sizeof(T);
}
template size_t soap_chunk_bytes<float>(int, int, int);
template size_t soap_chunk_bytes<double>(int, int, int);

template <typename T>
void launch_soap_center(const SoapLaunch<T>& p, int mode, T* coef, T* grad) {
    if (p.n_atoms == 0) return;
    if (mode == 0) {
        launch_center_ms<T, 16, 0>(p, coef, grad);
    } else if (mode == 1) {
        launch_center_ms<T, 8, 1>(p, coef, grad);
    } else {
        launch_center_ms<T, 8, 2>(p, coef, grad);
    }
}

template <typename T>
void launch_ps_values(const SoapLaunch<T>& p, const T* coef, const unsigned* type_mask, const long long* sample_flag,
                      const long long* sample_scan, T* out, bool fast, int lds_limit) {
    if (p.n_atoms == 0) return;
    if constexpr (sizeof(T) == 4) {
        // mixed precision on MFMA: n a multiple of 4, at most 16 types, Gram matrix in LDS
        const int W16 = (p.n_species * p.n + 15) / 16;
        const size_t mshm = ((size_t)p.L * p.L * 16 * W16 + (size_t)256 * W16 * W16) * sizeof(float);
        if (fast && p.n % 4 == 0 && p.n_species <= PSV_MAX_TYPES && mshm + 8192 <= (size_t)lds_limit) {
            ps_values_mfma_kernel<<<p.n_atoms, PSV_THREADS, mshm, p.stream>>>(
                p.n_atoms, p.n_species, p.L, p.n, W16, coef, p.species, type_mask, sample_flag, sample_scan, out);
            return;
        }
    }
    const size_t shmem = (size_t)p.n_species * p.L * p.L * p.n * sizeof(T);
    ps_values_kernel<T><<<p.n_atoms, WAVE, shmem, p.stream>>>(p.n_atoms, p.n_species, p.L, p.n, coef, p.species,
                                                              type_mask, sample_flag, sample_scan, out);
}

template <typename T>
size_t gradient_rows_shmem(const SoapLaunch<T>& p, bool ps) {
    const size_t slots = (size_t)p.L * p.L * p.n;
    return ((ps ? p.n_species * slots : 0) + 6 * slots) * sizeof(T) + soap_chunk_bytes<T>(p.L, p.n, p.chunk);
}

template <typename T>
void launch_gradient_rows(const SoapLaunch<T>& p, bool ps, const T* coef, const T* dself, const unsigned* type_mask,
                          const long long* sample_flag, const long long* sample_scan, const long long* row_scan,
                          const long long* block_sample_base, const long long* se_grad_base, const int* atom_system,
                          const int* atom_local, T* out, int* labels) {
    if (p.n_atoms == 0) return;
    SoapArgs<T> a = make_args(p);
    const size_t shmem = gradient_rows_shmem<T>(p, ps);
    if (ps) {
        gradient_rows_kernel<T, true><<<p.n_atoms, WAVE, shmem, p.stream>>>(a, coef, dself, type_mask, sample_flag,
                                                                            sample_scan, row_scan, block_sample_base,
                                                                            se_grad_base, atom_system, atom_local, out, labels);
    } else {
        gradient_rows_kernel<T, false><<<p.n_atoms, WAVE, shmem, p.stream>>>(a, coef, dself, type_mask, sample_flag,
                                                                             sample_scan, row_scan, block_sample_base,
                                                                             se_grad_base, atom_system, atom_local, out, labels);
    }
}

template <typename T>
size_t ps_gradient_shmem(const SoapLaunch<T>& p, int chunk) {
    const size_t slots = (size_t)p.L * p.L * p.n;
    const size_t npairs = (size_t)p.n_species * (p.n_species + 1) / 2;
    const size_t meta = (npairs * (8 + 12) + 15) / 16 * 16;
    return meta + ((size_t)p.n_species * slots + 3 * slots + (size_t)GradLayout(p.L, p.n).per_entry * chunk) * sizeof(T);
}

template <typename T>
bool launch_ps_gradient_fast(const SoapLaunch<T>& p, int chunk, const T* coef, const T* dself, const unsigned* type_mask,
                             const long long* sample_scan, const long long* row_scan, const long long* block_sample_base,
                             const int* atom_system, const int* atom_local, T* out, int* labels) {
    if (p.n_atoms == 0) return true;
    if (p.n > 16 || p.n_species > MAX_ROW_BLOCKS || chunk > 64 || chunk * p.L > GRAD_THREADS - WAVE) return false;
    SoapArgs<T> a = make_args(p);
    const size_t shmem = ps_gradient_shmem<T>(p, chunk);
#define RA_LAUNCH_PSG(NM)                                                                                           \
    ps_gradient_kernel<T, NM, 2><<<p.n_atoms, GRAD_THREADS, shmem, p.stream>>>(a, chunk, coef, dself, type_mask,    \
                                                                               sample_scan, row_scan, block_sample_base, \
                                                                               atom_system, atom_local, out, labels)
    if (p.n <= 4) {
        RA_LAUNCH_PSG(4);
    } else if (p.n <= 8) {
        RA_LAUNCH_PSG(8);
    } else if (p.n <= 12) {
        RA_LAUNCH_PSG(12);
    } else {
        RA_LAUNCH_PSG(16);
    }
#undef RA_LAUNCH_PSG
    return true;
}

// `rows_per_pass`: rows of each 4 x NC tile staged at once (1, 2 or 4)
template <typename T>
size_t ps_gradient_tiled_shmem(const SoapLaunch<T>& p, int max_runs, int max_entries, int rows_per_pass) {
    const int NC = (p.n + 3) / 4 * 4;
    const size_t cs = (size_t)p.L * p.L * NC;
    const size_t dc = (size_t)(max_runs > p.n_species ? max_runs : p.n_species) * 3 * cs;
    const int per_entry = TileEntryLayout(p.L, p.n, NC).per_entry;
    const size_t entries = (size_t)per_entry * max_entries * sizeof(T);
    const int spw = (sizeof(T) == 4 && NC == 8 && rows_per_pass == 1) ? 8 * ORIENTED_SLOT : WAVE * rows_per_pass * NC;
    const size_t staging = (size_t)(TILE_THREADS / WAVE) * spw * sizeof(T) + (size_t)TILE_THREADS * sizeof(void*);
    return ((size_t)p.n_species * cs + dc) * sizeof(T) + (entries > staging ? entries : staging);
}

// dynamic plus static LDS of the gradient-row kernel (the static tables, with a margin
// for the padding the compiler may add)
template <typename T>
size_t ps_gradient_tiled_lds(const SoapLaunch<T>& p, int max_runs, int max_entries, int rows_per_pass) {
    const size_t fixed = TILE_MAX_RUNS * (4 + TILE_MAX_TYPES) * sizeof(int) + (TILE_MAX_TYPES + 1) * sizeof(int) +
                         TILE_MAX_L * sizeof(T) + TILE_MAX_L * (TILE_MAX_L + 1) * sizeof(int);
    return ps_gradient_tiled_shmem<T>(p, max_runs, max_entries, rows_per_pass) + fixed + 256;
}

bool ps_gradient_tiled_supported(int L, int n, int n_species, int max_runs, int max_entries, int rows_per_pass) {
    return n <= 16 && L <= TILE_MAX_L && n_species <= TILE_MAX_TYPES && max_runs >= 1 && max_runs <= TILE_MAX_RUNS &&
           max_entries >= 1 && (rows_per_pass == 1 || rows_per_pass == 2 || rows_per_pass == 4);
}

void launch_ps_grad_plan(hipStream_t stream, int n_atoms, int n_species, const long long* offsets, const int* entry_j,
                         const int* species, const unsigned* type_mask, const long long* sample_scan,
                         const long long* row_scan, const long long* block_sample_base, const int* atom_system,
                         const int* atom_local, int max_runs, int max_entries, int* runs, int* run_rows, int* windows,
                         int* n_windows, int* self_rows, int* labels, bool power_spectrum, int* split) {
    if (n_atoms == 0) return;
    const size_t npairs = (size_t)n_species * (n_species + 1) / 2;
    const size_t shmem = npairs * (8 + 4 * 3);
    if (power_spectrum) {
        ps_grad_plan_kernel<true><<<n_atoms, WAVE, shmem, stream>>>(
            n_atoms, n_species, offsets, entry_j, species, type_mask, sample_scan, row_scan, block_sample_base,
            atom_system, atom_local, max_runs, max_entries, reinterpret_cast<int4*>(runs), run_rows,
            reinterpret_cast<int4*>(windows), n_windows, self_rows, labels, split);
    } else {
        ps_grad_plan_kernel<false><<<n_atoms, WAVE, shmem, stream>>>(
            n_atoms, n_species, offsets, entry_j, species, type_mask, sample_scan, row_scan, block_sample_base,
            atom_system, atom_local, max_runs, max_entries, reinterpret_cast<int4*>(runs), run_rows,
            reinterpret_cast<int4*>(windows), n_windows, self_rows, labels, split);
    }
}

template <typename T>
size_t entry_prep_size(const SoapLaunch<T>& p, long long n_entries) {
    const int NC = (p.n + 3) / 4 * 4;
    return (size_t)TileEntryLayout(p.L, p.n, NC).per_entry * (size_t)n_entries;
}

template <typename T>
void launch_entry_prep(const SoapLaunch<T>& p, long long n_entries, T* data) {
    if (n_entries == 0) return;
    SoapArgs<T> a = make_args(p);
    const int NC = (p.n + 3) / 4 * 4;
    const int per_block = 128 / p.L;
    const size_t shmem = (size_t)per_block * TileEntryLayout(p.L, p.n, NC).per_entry * sizeof(T);
    const long long blocks = (n_entries + per_block - 1) / per_block;
    entry_prep_kernel<T><<<(unsigned)blocks, 256, shmem, p.stream>>>(a, n_entries, NC, per_block, data);
}

template <typename T>
void launch_ps_gradient_tiled(const SoapLaunch<T>& p, int max_runs, int max_entries, int rows_per_pass, const T* coef,
                              const T* dself, const unsigned* type_mask, const int* runs, const int* run_rows,
                              const int* windows, const int* n_windows, const int* self_rows, const T* entry_data,
                              T* out, bool fast, const int* only) {
    if (p.n_atoms == 0) return;
    SoapArgs<T> a = make_args(p);
    const size_t shmem = ps_gradient_tiled_shmem<T>(p, max_runs, max_entries, rows_per_pass);
    const int4* rs = reinterpret_cast<const int4*>(runs);
    const int4* ws = reinterpret_cast<const int4*>(windows);
    // the mixed-precision contraction covers FP32 with n = 8 (windows of up to 4 runs),
    // n = 12 and n = 16
    const bool mixed = fast && sizeof(T) == 4 && ((p.n == 8 && max_runs <= 4) || p.n == 12 || p.n == 16);
#define RA_LAUNCH_TILED(NCV, RPV, FV)                                                                                  \
    ps_gradient_tiled_kernel<T, NCV, RPV, FV><<<p.n_atoms, TILE_THREADS, shmem, p.stream>>>(                          \
        a, max_runs, max_entries, coef, dself, type_mask, rs, run_rows, ws, n_windows, self_rows, entry_data, out, only)
#define RA_LAUNCH_TILED_RP(NCV)       \
    if (rows_per_pass == 4) {         \
        RA_LAUNCH_TILED(NCV, 4, false);      \
    } else if (rows_per_pass == 2) {  \
        RA_LAUNCH_TILED(NCV, 2, false);      \
    } else {                          \
        RA_LAUNCH_TILED(NCV, 1, false);      \
    }
#define RA_LAUNCH_MIXED(NCV)         \
    if (rows_per_pass == 4) {        \
        RA_LAUNCH_TILED(NCV, 4, true);      \
    } else if (rows_per_pass == 2) { \
        RA_LAUNCH_TILED(NCV, 2, true);      \
    } else {                         \
        RA_LAUNCH_TILED(NCV, 1, true);      \
    }
    if (mixed) {
        if (p.n == 8) {
            RA_LAUNCH_MIXED(8)
        } else if (p.n == 12) {
            RA_LAUNCH_MIXED(12)
        } else {
            RA_LAUNCH_MIXED(16)
        }
    } else if (p.n <= 4) {
        RA_LAUNCH_TILED_RP(4)
    } else if (p.n <= 8) {
        RA_LAUNCH_TILED_RP(8)
    } else if (p.n <= 12) {
        RA_LAUNCH_TILED_RP(12)
    } else {
        RA_LAUNCH_TILED_RP(16)
    }
#undef RA_LAUNCH_MIXED
#undef RA_LAUNCH_TILED_RP
#undef RA_LAUNCH_TILED
}

bool ps_gradient_rank2_supported(int L, int n, int n_species, int max_entries) {
    return n == 8 && L >= 1 && L <= R2_MAX_L && n_species <= TILE_MAX_TYPES && (max_entries == 8 || max_entries == 16);
}

// dynamic plus static LDS of the rank-2 kernel
size_t ps_gradient_rank2_lds(const SoapLaunch<float>& p, int max_entries) {
    const Rank2Layout ly(p.L, p.n_species, max_entries);
    const size_t fixed = TILE_MAX_TYPES * sizeof(int) + R2_MAX_L * sizeof(float) +
                         (2 * R2_MAX_L * (R2_MAX_L + 1) / 2 + 4 * R2_MAX_L) * sizeof(float) + R2_MAX_L * R2_MAX_L;
    return (size_t)ly.total * sizeof(float) + fixed + 256;
}

void launch_ps_gradient_rank2(const SoapLaunch<float>& p, int max_entries, const float* coef, const float* dself,
                              const unsigned* type_mask, const int* runs, const int* run_rows, const int* windows,
                              const int* n_windows, const int* self_rows, const int* split, const float* entry_data,
                              float* out) {
    if (p.n_atoms == 0) return;
    if (!ps_gradient_rank2_supported(p.L, p.n, p.n_species, max_entries)) {
        throw std::invalid_argument("rank-2 gradient kernel: unsupported shape");
    }
    SoapArgs<float> a = make_args(p);
    const size_t shmem = (size_t)Rank2Layout(p.L, p.n_species, max_entries).total * sizeof(float);
    const int4* rs = reinterpret_cast<const int4*>(runs);
    const int4* ws = reinterpret_cast<const int4*>(windows);
    if (max_entries == 16) {
        ps_gradient_rank2_kernel<16><<<p.n_atoms, R2_THREADS, shmem, p.stream>>>(
            a, coef, dself, type_mask, rs, run_rows, ws, n_windows, self_rows, split, entry_data, out);
    } else {
        ps_gradient_rank2_kernel<8><<<p.n_atoms, R2_THREADS, shmem, p.stream>>>(
            a, coef, dself, type_mask, rs, run_rows, ws, n_windows, self_rows, split, entry_data, out);
    }
}

template <typename T>
size_t se_gradient_tiled_shmem(const SoapLaunch<T>& p, int max_runs, int max_entries) {
    const int NC = (p.n + 3) / 4 * 4;
    const size_t cs = (size_t)p.L * p.L * NC;
    return ((size_t)max_runs * 3 * cs + (size_t)TileEntryLayout(p.L, p.n, NC).per_entry * max_entries) * sizeof(T);
}

template <typename T>
void launch_se_gradient_tiled(const SoapLaunch<T>& p, int max_runs, int max_entries, const T* dself,
                              const unsigned* type_mask, const int* runs, const int* run_rows, const int* windows,
                              const int* n_windows, const int* self_rows, const long long* se_grad_base,
                              const T* entry_data, T* out) {
    if (p.n_atoms == 0) return;
    SoapArgs<T> a = make_args(p);
    const size_t shmem = se_gradient_tiled_shmem<T>(p, max_runs, max_entries);
    const int4* rs = reinterpret_cast<const int4*>(runs);
    const int4* ws = reinterpret_cast<const int4*>(windows);
#define RA_LAUNCH_SE(NCV)                                                                                             \
    se_gradient_tiled_kernel<T, NCV><<<p.n_atoms, TILE_THREADS, shmem, p.stream>>>(                                  \
        a, max_runs, max_entries, dself, type_mask, rs, run_rows, ws, n_windows, self_rows, se_grad_base, entry_data, out)
    if (p.n <= 4) {
        RA_LAUNCH_SE(4);
    } else if (p.n <= 8) {
        RA_LAUNCH_SE(8);
    } else if (p.n <= 12) {
        RA_LAUNCH_SE(12);
    } else {
        RA_LAUNCH_SE(16);
    }
#undef RA_LAUNCH_SE
}

template <typename T>
size_t expansion_tiled_shmem(const SoapLaunch<T>& p, int max_entries) {
    const int NC = (p.n + 3) / 4 * 4;
    const size_t cs = (size_t)p.L * p.L * NC;
    const size_t per_entry = (size_t)TileEntryLayout(p.L, p.n, NC).per_entry + 4 * (size_t)p.L * p.L;
    return ((size_t)p.n_species * 4 * cs + per_entry * max_entries) * sizeof(T);
}

template <typename T>
void launch_expansion_tiled(const SoapLaunch<T>& p, int max_entries, const T* entry_data, T* coef, T* dself) {
    if (p.n_atoms == 0) return;
    SoapArgs<T> a = make_args(p);
    const size_t shmem = expansion_tiled_shmem<T>(p, max_entries);
#define RA_LAUNCH_EXP(NCV, NHV) \
    expansion_tiled_kernel<T, NCV, NHV><<<p.n_atoms, TILE_THREADS, shmem, p.stream>>>(a, max_entries, entry_data, coef, dself)
    if (p.n <= 4) {
        RA_LAUNCH_EXP(4, 1);
    } else if (p.n <= 8) {
        RA_LAUNCH_EXP(8, 1);
    } else if (p.n <= 12) {
        RA_LAUNCH_EXP(12, 1);
    } else {
        RA_LAUNCH_EXP(16, 2);
    }
#undef RA_LAUNCH_EXP
}

bool expansion_mfma_supported(int L, int n, int S) { return n == 8 && L >= 1 && L <= EXM_MAX_L && S >= 1; }

bool expansion_mfma_values_supported(int n_species) { return n_species <= EXM_MAX_VALUE_TYPES; }

size_t expansion_mfma_lds(const SoapLaunch<float>& p, int E, bool values) {
    return (size_t)ExpansionMfmaLayout(p.L, E, values ? p.n_species : 0).total * sizeof(float);
}

void launch_expansion_mfma(const SoapLaunch<float>& p, int E, const float* entry_data, float* coef, float* dself,
                           const unsigned* type_mask, const long long* sample_flag, const long long* sample_scan,
                           float* ps_values, bool bf16_values) {
    if (p.n_atoms == 0) return;
    if (!expansion_mfma_supported(p.L, p.n, p.n_species) || (E != 8 && E != 16)) {
        throw std::invalid_argument("the matrix-core expansion kernel takes n = 8, l < 9 and chunks of 8 or 16 entries");
    }
    if (ps_values && (!expansion_mfma_values_supported(p.n_species) || !type_mask || !sample_flag || !sample_scan)) {
        throw std::invalid_argument("the expansion kernel forms power-spectrum values for at most 16 types, with the "
                                    "type masks and sample offsets");
    }
    // the record prefetch holds E records of at most 204 floats per thread group
    if (ExpansionMfmaLayout(p.L, E, 0).per_entry > 204) throw std::invalid_argument("expansion record larger than staged");
    SoapArgs<float> a = make_args(p);
    const size_t shmem = expansion_mfma_lds(p, E, ps_values != nullptr);
    const int nh = expansion_mfma_halves(p.L);
#define RA_LAUNCH_EXM(EV, NHV)                                                                        \
    expansion_mfma_kernel<EV, NHV><<<p.n_atoms, EXM_THREADS, shmem, p.stream>>>(a, entry_data, coef, dself, \
                                                                                type_mask, sample_flag,   \
                                                                                sample_scan, ps_values, \
                                                                                bf16_values ? 1 : 0)
#define RA_LAUNCH_EXM_E(NHV)  \
    if (E == 8) {             \
        RA_LAUNCH_EXM(8, NHV); \
    } else {                  \
        RA_LAUNCH_EXM(16, NHV); \
    }
    switch (nh) {
        case 1: RA_LAUNCH_EXM_E(1); break;
        case 2: RA_LAUNCH_EXM_E(2); break;
        case 3: RA_LAUNCH_EXM_E(3); break;
        default: RA_LAUNCH_EXM_E(4); break;
    }
#undef RA_LAUNCH_EXM_E
#undef RA_LAUNCH_EXM
}

template <typename T>
size_t ps_vjp_coef_shmem(const SoapLaunch<T>& p) {
    const size_t slots = (size_t)p.L * p.L * p.n;
    return (2 * (size_t)p.n_species * slots + (size_t)p.L * p.n * p.n) * sizeof(T);
}

template <typename T>
void launch_ps_vjp_coef(const SoapLaunch<T>& p, const T* coef, const unsigned* type_mask, const long long* sample_flag,
                        const long long* sample_scan, const T* grad_values, T* H) {
    if (p.n_atoms == 0) return;
    ps_vjp_coef_kernel<T><<<p.n_atoms, TILE_THREADS, ps_vjp_coef_shmem<T>(p), p.stream>>>(
        p.n_atoms, p.n_species, p.L, p.n, coef, p.species, type_mask, sample_flag, sample_scan, grad_values, H);
}

template <typename T>
void launch_se_vjp_coef(const SoapLaunch<T>& p, const long long* sample_flag, const long long* sample_scan,
                        const long long* block_sample_base, const long long* value_base, const T* grad_values, T* H) {
    const long long total = (long long)p.n_atoms * p.n_species * p.L * p.L * p.n;
    if (total == 0) return;
    const long long blocks = std::min<long long>((total + 255) / 256, 65536);
    se_vjp_coef_kernel<T><<<(unsigned)blocks, 256, 0, p.stream>>>(p.n_atoms, p.n_species, p.L, p.n, p.species,
                                                                  sample_flag, sample_scan, block_sample_base,
                                                                  value_base, grad_values, H);
}

template <typename T>
size_t vjp_pair_shmem(const SoapLaunch<T>& p, int max_entries) {
    const int NC = (p.n + 3) / 4 * 4;
    const size_t L2 = (size_t)p.L * p.L;
    const size_t per_entry = (size_t)TileEntryLayout(p.L, p.n, NC).per_entry + 8 * L2;
    return ((size_t)p.n_species * L2 * NC + per_entry * max_entries) * sizeof(T);
}

template <typename T>
void launch_vjp_pair(const SoapLaunch<T>& p, int max_entries, const T* entry_data, const T* H, T* g) {
    if (p.n_atoms == 0) return;
    SoapArgs<T> a = make_args(p);
    const size_t shmem = vjp_pair_shmem<T>(p, max_entries);
#define RA_LAUNCH_VJP(NCV) \
    vjp_pair_kernel<T, NCV><<<p.n_atoms, TILE_THREADS, shmem, p.stream>>>(a, max_entries, entry_data, H, g)
    if (p.n <= 4) {
        RA_LAUNCH_VJP(4);
    } else if (p.n <= 8) {
        RA_LAUNCH_VJP(8);
    } else if (p.n <= 12) {
        RA_LAUNCH_VJP(12);
    } else {
        RA_LAUNCH_VJP(16);
    }
#undef RA_LAUNCH_VJP
}

template <typename T>
void launch_vjp_gather(hipStream_t stream, int n_atoms, const long long* offsets, const int* entry_j,
                       const int* entry_shift, const T* g, T* out) {
    if (n_atoms == 0) return;
    vjp_gather_kernel<T><<<(n_atoms + 255) / 256, 256, 0, stream>>>(n_atoms, offsets, entry_j, entry_shift, g, out);
}

template <typename T>
void launch_ps_cell_strain(const SoapLaunch<T>& p, const T* coef, const T* dcs, const unsigned* type_mask,
                           const long long* sample_flag, const long long* sample_scan, T* out) {
    if (p.n_atoms == 0) return;
    const size_t slots = (size_t)p.L * p.L * p.n;
    const size_t shmem = (p.n_species + 18) * slots * sizeof(T);
    ps_cell_strain_kernel<T><<<p.n_atoms, WAVE, shmem, p.stream>>>(p.n_atoms, p.n_species, p.L, p.n, coef, dcs, p.species,
                                                                   type_mask, sample_flag, sample_scan, out);
}

template <typename T>
void launch_se_gather(const SoapLaunch<T>& p, int ncomp, const T* dense, const long long* sample_flag,
                      const long long* sample_scan, const long long* block_sample_base, const long long* region_base, T* out) {
    const long long total = (long long)p.n_species * p.n_species * p.n_atoms;
    if (total == 0) return;
    se_gather_kernel<T><<<(unsigned)total, 64, 0, p.stream>>>(p.n_atoms, p.n_species, p.L, p.n, ncomp, dense, p.species,
                                                              sample_flag, sample_scan, block_sample_base, region_base, out);
}


// spherical_expansion_by_pair: one directed pair per sample. A workgroup takes
// `a.chunk` consecutive samples; each of the first lanes fills the per-entry data
// (spherical harmonics, cutoff x radial values) of one sample's neighbour entry into
// LDS, then the wave writes every sample's rows, lanes over the (l, m, n) slots:
//   values   f Y_lm I_nl,             block (type pair, l), row row0 / row1
//   g[x]     d values / d pair vector (soap/spherical_expansion_pair.rs:382-419)
//   positions -g (first atom), +g (second atom); a single zero row for a pair of an
//            atom with its own image
//   strain   v[a] g[x]; cell shift[a] g[x]
// Self pairs (entry < 0) hold the self contribution at l = 0 and zero strain / cell rows.
template <typename T>
__global__ void __launch_bounds__(WAVE) by_pair_kernel(SoapArgs<T> a, long long n_samples,
                                                       const ByPairSample* __restrict__ samples,
                                                       const long long* __restrict__ value_base,
                                                       const long long* __restrict__ pos_base,
                                                       const long long* __restrict__ strain_base,
                                                       const long long* __restrict__ cell_base, T* __restrict__ values,
                                                       T* __restrict__ pos, T* __restrict__ strain, T* __restrict__ cell) {
    extern __shared__ __align__(16) unsigned char smem[];
    const ChunkLayout lay(a.L, a.n, a.chunk);
    T* chunk = reinterpret_cast<T*>(smem);
    const long long first = (long long)blockIdx.x * a.chunk;
    const int cnt = (int)min((long long)a.chunk, n_samples - first);
    const int lane = threadIdx.x;
    if (lane < cnt) {
        const long long e = samples[first + lane].entry;
        if (e >= 0) entry_data<T>(a, lay, e, -1, chunk + lane * lay.per_entry);
    }
    __syncthreads();
    const int n = a.n;
    const int slots = lay.L2 * n;
    for (int c = 0; c < cnt; c++) {
        const ByPairSample sm = samples[first + c];
        const T* en = chunk + c * lay.per_entry;
        for (int s = lane; s < slots; s += WAVE) {
            const int lm = s / n, nn = s - lm * n;
            const int l = l_of_lm(lm);
            const int row = l == 0 ? sm.row0 : sm.row1;
            if (row < 0) continue;
            const long long w = (long long)(2 * l + 1) * n;
            const long long off = (long long)(lm - l * l) * n + nn;
            const long long blk = (long long)sm.type_pair * a.L + l;
            if (sm.entry < 0) {
                values[value_base[blk] + row * w + off] = a.self_contribution[nn];
                for (int q = 0; q < 9; q++) {
                    if (strain) strain[strain_base[blk] + (row * 9LL + q) * w + off] = (T)0;
                    if (cell) cell[cell_base[blk] + (row * 9LL + q) * w + off] = (T)0;
                }
                continue;
            }
            const T y = en[lay.y + lm];
            values[value_base[blk] + row * w + off] = en[lay.rv + l * n + nn] * y;
            if (!pos && !strain && !cell) continue;
            const T ra = en[lay.ra + l * n + nn] * y, rb = en[lay.rb + l * n + nn];
            T g[3];
#pragma unroll
            for (int x = 0; x < 3; x++) g[x] = ra * en[lay.dir + x] + rb * en[lay.dy + x * lay.L2 + lm];
            if (pos && sm.grad_rows > 0) {
                T* out = pos + pos_base[blk] + (long long)sm.grad_row * 3 * w + off;
#pragma unroll
                for (int x = 0; x < 3; x++) {
                    if (sm.grad_rows == 2) {
                        out[x * w] = -g[x];
                        out[(3 + x) * w] = g[x];
                    } else {
                        out[x * w] = (T)0;
                    }
                }
            }
#pragma unroll
            for (int p = 0; p < 3; p++) {
#pragma unroll
                for (int x = 0; x < 3; x++) {
                    if (strain) strain[strain_base[blk] + (row * 9LL + p * 3 + x) * w + off] = en[lay.vec + p] * g[x];
                    if (cell) cell[cell_base[blk] + (row * 9LL + p * 3 + x) * w + off] = en[lay.shift + p] * g[x];
                }
            }
        }
    }
}

template <typename T>
void launch_by_pair(const SoapLaunch<T>& p, long long n_samples, const ByPairSample* samples, const long long* value_base,
                    const long long* pos_base, const long long* strain_base, const long long* cell_base, T* values, T* pos,
                    T* strain, T* cell) {
    if (n_samples == 0) return;
    SoapArgs<T> a = make_args(p);
    const unsigned blocks = (unsigned)((n_samples + p.chunk - 1) / p.chunk);
    by_pair_kernel<T><<<blocks, WAVE, soap_chunk_bytes<T>(p.L, p.n, p.chunk), p.stream>>>(
        a, n_samples, samples, value_base, pos_base, strain_base, cell_base, values, pos, strain, cell);
}

#define RA_INSTANTIATE(T)                                                                                              \
    template void launch_soap_center<T>(const SoapLaunch<T>&, int, T*, T*);                                            \
    template void launch_ps_values<T>(const SoapLaunch<T>&, const T*, const unsigned*, const long long*,              \
                                      const long long*, T*, bool, int);                                                \
    template size_t gradient_rows_shmem<T>(const SoapLaunch<T>&, bool);                                                \
    template size_t ps_gradient_shmem<T>(const SoapLaunch<T>&, int);                                                   \
    template bool launch_ps_gradient_fast<T>(const SoapLaunch<T>&, int, const T*, const T*, const unsigned*,          \
                                             const long long*, const long long*, const long long*, const int*,        \
                                             const int*, T*, int*);                                                    \
    template size_t ps_gradient_tiled_shmem<T>(const SoapLaunch<T>&, int, int, int);                                   \
    template size_t ps_gradient_tiled_lds<T>(const SoapLaunch<T>&, int, int, int);                                     \
    template void launch_ps_gradient_tiled<T>(const SoapLaunch<T>&, int, int, int, const T*, const T*, const unsigned*, \
                                              const int*, const int*, const int*, const int*, const int*, const T*, T*, \
                                              bool, const int*);                                                       \
    template size_t entry_prep_size<T>(const SoapLaunch<T>&, long long);                                               \
    template size_t expansion_tiled_shmem<T>(const SoapLaunch<T>&, int);                                               \
    template size_t se_gradient_tiled_shmem<T>(const SoapLaunch<T>&, int, int);                                        \
    template void launch_se_gradient_tiled<T>(const SoapLaunch<T>&, int, int, const T*, const unsigned*, const int*,    \
                                              const int*, const int*, const int*, const int*, const long long*,        \
                                              const T*, T*);                                                           \
    template void launch_expansion_tiled<T>(const SoapLaunch<T>&, int, const T*, T*, T*);                              \
    template void launch_entry_prep<T>(const SoapLaunch<T>&, long long, T*);                                            \
    template void launch_gradient_rows<T>(const SoapLaunch<T>&, bool, const T*, const T*, const unsigned*,             \
                                          const long long*, const long long*, const long long*, const long long*,      \
                                          const long long*, const int*, const int*, T*, int*);                         \
    template size_t ps_vjp_coef_shmem<T>(const SoapLaunch<T>&);                                                        \
    template void launch_ps_vjp_coef<T>(const SoapLaunch<T>&, const T*, const unsigned*, const long long*,            \
                                        const long long*, const T*, T*);                                               \
    template void launch_se_vjp_coef<T>(const SoapLaunch<T>&, const long long*, const long long*, const long long*,   \
                                        const long long*, const T*, T*);                                               \
    template size_t vjp_pair_shmem<T>(const SoapLaunch<T>&, int);                                                      \
    template void launch_vjp_pair<T>(const SoapLaunch<T>&, int, const T*, const T*, T*);                               \
    template void launch_vjp_gather<T>(hipStream_t, int, const long long*, const int*, const int*, const T*, T*);      \
    template void launch_ps_cell_strain<T>(const SoapLaunch<T>&, const T*, const T*, const unsigned*,                 \
                                           const long long*, const long long*, T*);                                    \
    template void launch_se_gather<T>(const SoapLaunch<T>&, int, const T*, const long long*, const long long*,        \
                                      const long long*, const long long*, T*);                                         \
    template void launch_by_pair<T>(const SoapLaunch<T>&, long long, const ByPairSample*, const long long*,           \
                                    const long long*, const long long*, const long long*, T*, T*, T*, T*);

RA_INSTANTIATE(float)
RA_INSTANTIATE(double)

}  // namespace rascaline_amd
