// LODE spherical expansion kernels (gfx950): structure-factor phases in FP64,
// k-space projection, and the three contractions over k-vectors (values, the
// center's own gradient rows, neighbour gradient rows) as LDS-tiled GEMMs whose
// A operand is generated on the fly from the phases and whose epilogue writes
// straight into the output blocks.
//
// Per system (featomic/src/calculators/lode/spherical_expansion.rs):
//   phases     cos / sin (k . r_i), FP64                           :95-157
//   sums       per neighbour type, Σ_j cos / sin (k . r_j), FP64
//   B matrices Bv[k][(l, m, n)] = (4 pi / V) phase_l rho(k) I_nl(k) Y_lm(k^)
//              Bg[k][(x, l, m, n)] = Bv[k][(l, m, n)] |k| k^_x      :227-342,696-869
//   values     c[i, t] = Σ_k SF_t(i, k) Bv[k]   (real SF for even l, imaginary for odd)
//   self rows  d c[i, t] / d r_i = Σ_k A_self Bg[k]
//   pair rows  d c[i, type j] / d r_j = Σ_k A_pair(i, j, k) Bg[k]
// with, for even / odd l,
//   SF_re = 2 (c_i Σc_t + s_i Σs_t),  SF_im = 2 (s_i Σc_t - c_i Σs_t)
//   A_self  = -SF_im           /  SF_re - 2 [t == type i]
//   A_pair  = 2 (s_i c_j - c_i s_j)  /  -2 (c_i c_j + s_i s_j).
#include <hip/hip_runtime.h>

#include "kernels_lode.hpp"
#include "soap_common.hpp"
#include "special.hpp"

namespace rascaline_amd {

namespace {

constexpr int LT = 256;  // threads per workgroup

// cos / sin (k . r_i) for all (i, k) of one system, FP64
__global__ void lode_phase_kernel(int n_atoms, int n_k, const double* __restrict__ positions, const double* __restrict__ kvec,
                                  double* __restrict__ cosv, double* __restrict__ sinv) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)n_atoms * n_k) return;
    const int i = (int)(idx / n_k), k = (int)(idx - (long long)i * n_k);
    const double* r = positions + 3 * i;
    const double* kv = kvec + 4 * k;
    const double kx = kv[3] * kv[0], ky = kv[3] * kv[1], kz = kv[3] * kv[2];
    const double arg = kx * r[0] + ky * r[1] + kz * r[2];
    double s, c;
    sincos(arg, &s, &c);
    cosv[idx] = c;
    sinv[idx] = s;
}

// per-type sums over atoms, FP64: sumc[t][k], sums[t][k]
__global__ void lode_type_sums_kernel(int n_atoms, int n_k, int S, const int* __restrict__ type_index,
                                      const double* __restrict__ cosv, const double* __restrict__ sinv,
                                      double* __restrict__ sumc, double* __restrict__ sums) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_k) return;
    for (int t = 0; t < S; t++) {
        sumc[(long long)t * n_k + k] = 0.0;
        sums[(long long)t * n_k + k] = 0.0;
    }
    for (int j = 0; j < n_atoms; j++) {
        const int t = type_index[j];
        sumc[(long long)t * n_k + k] += cosv[(long long)j * n_k + k];
        sums[(long long)t * n_k + k] += sinv[(long long)j * n_k + k];
    }
}

// k-space projection: one thread per k-vector writes its rows of the B matrices of
// both parities (columns per parity in (l, m, n) order, l of that parity).
template <typename T>
__global__ void lode_projection_kernel(LodeProjectionArgs<T> p) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= p.n_k) return;
    const int L = p.L, n = p.n;
    const double* kv = p.kvec + 4 * k;
    T d[3] = {(T)kv[0], (T)kv[1], (T)kv[2]};
    const double knorm = kv[3];
    T Y[LODE_MAX_L * LODE_MAX_L];
    T P[LODE_MAX_L * (LODE_MAX_L + 1) / 2];
    spherical_harmonics<T>(L, d, Y, nullptr, P, nullptr, nullptr, p.sh_a, p.sh_b);
    const double scale_k = p.global_factor * p.density_fourier[k];
    int col[2] = {0, 0};
    for (int l = 0; l < L; l++) {
        const int par = l % 2;
        const double phase = par == 0 ? ((l / 2) % 2 == 0 ? 1.0 : -1.0) : (((l + 1) / 2) % 2 == 0 ? 1.0 : -1.0);
        T R[LODE_MAX_N];
        if (p.radial.use_spline) {
            radial_integral<T>(p.radial, l, 0, knorm, R, nullptr);
        } else {
            double raw[LODE_MAX_N];
            for (int q = 0; q < n; q++) {
                double g;
                lode_gto_raw(l, q, knorm, p.radial.gto[q], &raw[q], &g);
            }
            for (int q = 0; q < n; q++) {
                double v = 0.0;
                for (int r = 0; r < n; r++) v += raw[r] * p.radial.ortho[q * n + r];
                R[q] = (T)v;
            }
        }
        T* bv = p.bv[par] + (long long)k * p.ncol_v[par];
        T* bg = p.bg[par] + (long long)k * 3 * p.ncol_v[par];
        const int nv = p.ncol_v[par];
        for (int m = 0; m < 2 * l + 1; m++) {
            const T y = Y[l * l + m];
            for (int q = 0; q < n; q++) {
                const T v = (T)(scale_k * phase) * (R[q] * y);
                const int c = col[par] + m * n + q;
                bv[c] = v;
                const T vk = v * (T)knorm;
                bg[c] = vk * d[0];
                bg[nv + c] = vk * d[1];
                bg[2 * nv + c] = vk * d[2];
            }
        }
        col[par] += (2 * l + 1) * n;
    }
}

// Tiled GEMM out[row][col] = Σ_k A(row, k) B[k][col] for one parity, 64 x 64 tiles,
// 16 k per step, 4 x 4 outputs per thread.
constexpr int GT = 64, GK = 16;

template <typename T, int MODE>
__global__ void __launch_bounds__(LT) lode_gemm_kernel(LodeGemmArgs<T> g) {
    __shared__ T As[GK][GT + 4];
    __shared__ T Bs[GK][GT + 4];
    const int tid = threadIdx.x;
    const int tx = tid % 16, ty = tid / 16;
    const long long row0 = (long long)blockIdx.x * GT;
    const int col0 = blockIdx.y * GT;
    const int N = g.n_atoms, nk = g.n_k;
    const int par = g.parity;
    const long long n_rows = MODE == 2 ? (long long)N * N : (long long)N * g.n_types;
    const int ncol = MODE == 0 ? g.ncol_v : 3 * g.ncol_v;
    const T* B = MODE == 0 ? g.bv : g.bg;
    T acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) acc[a][b] = 0;

    for (int k0 = 0; k0 < nk; k0 += GK) {
        // A tile: 64 rows x 16 k, 4 per thread
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const int idx = tid + e * LT;
            const int r = idx / GK, kk = idx % GK;
            const long long row = row0 + r;
            const int k = k0 + kk;
            double v = 0.0;
            if (row < n_rows && k < nk) {
                if (MODE == 2) {
                    const int i = (int)(row / N), j = (int)(row - (long long)i * N);
                    const double ci = g.cosv[(long long)i * nk + k], si = g.sinv[(long long)i * nk + k];
                    const double cj = g.cosv[(long long)j * nk + k], sj = g.sinv[(long long)j * nk + k];
                    v = par == 0 ? 2.0 * (si * cj - ci * sj) : -2.0 * (ci * cj + si * sj);
                } else {
                    const int i = (int)(row / g.n_types), tl = (int)(row - (long long)i * g.n_types);
                    const int t = g.types_present[tl];
                    const double ci = g.cosv[(long long)i * nk + k], si = g.sinv[(long long)i * nk + k];
                    const double sc = g.sumc[(long long)t * nk + k], ss = g.sums[(long long)t * nk + k];
                    const double re = 2.0 * (ci * sc + si * ss), im = 2.0 * (si * sc - ci * ss);
                    if (MODE == 0) {
                        v = par == 0 ? re : im;
                    } else {
                        v = par == 0 ? -im : re - (t == g.type_index[i] ? 2.0 : 0.0);
                    }
                }
            }
            As[kk][r] = (T)v;
        }
        // B tile: 16 k x 64 cols
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const int idx = tid + e * LT;
            const int kk = idx / GT, c = idx % GT;
            const int k = k0 + kk, col = col0 + c;
            Bs[kk][c] = (k < nk && col < ncol) ? B[(long long)k * ncol + col] : (T)0;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < GK; kk++) {
            T a[4], b[4];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                a[q] = As[kk][ty + 16 * q];
                b[q] = Bs[kk][tx + 16 * q];
            }
#pragma unroll
            for (int x = 0; x < 4; x++)
#pragma unroll
                for (int y = 0; y < 4; y++) acc[x][y] += a[x] * b[y];
        }
        __syncthreads();
    }

    // epilogue: scatter into the blocks
#pragma unroll
    for (int x = 0; x < 4; x++) {
        const long long row = row0 + ty + 16 * x;
        if (row >= n_rows) continue;
        int i, t, j = -1;
        if (MODE == 2) {
            i = (int)(row / N);
            j = (int)(row - (long long)i * N);
            if (j == i) continue;
            t = g.type_index[j];
        } else {
            i = (int)(row / g.n_types);
            t = g.types_present[(int)(row - (long long)i * g.n_types)];
        }
        const int ct = g.type_index[i];
        const int S = g.n_species;
        const long long pair = (long long)ct * S + t;
#pragma unroll
        for (int y = 0; y < 4; y++) {
            const int col = col0 + tx + 16 * y;
            if (col >= ncol) continue;
            const int cv = MODE == 0 ? col : col % g.ncol_v;
            const int xyz = MODE == 0 ? 0 : col / g.ncol_v;
            const int l = g.col_l[cv];
            const int in_l = g.col_in_l[cv];  // m * n + q within the l segment
            const int width = (2 * l + 1) * g.n;
            const long long block = (long long)l * S * S + pair;
            T v = acc[x][y];
            long long dst;
            if (MODE == 0) {
                const long long srow = g.value_row0[pair] + g.type_rank[i];
                if (l == 0) v += g.value_shift[pair * g.n + in_l];
                dst = g.value_base[block] + srow * width + in_l;
                g.values[dst] = v;
            } else {
                // gradient row: the sample's first row plus the position of atom j (or i)
                const long long first = g.grad_row0[pair] + (long long)g.type_rank[i] * g.rows_per_sample[pair];
                long long rank;
                if (MODE == 1) {
                    rank = g.below[(long long)i * S + t];
                } else {
                    rank = g.type_rank[j] + ((ct != t && i < j) ? 1 : 0);
                }
                dst = g.grad_base[block] + (first + rank) * 3 * width + (long long)xyz * width + in_l;
                g.grads[dst] = v;
            }
        }
    }
}

}  // namespace

void launch_lode_phases(hipStream_t stream, int n_atoms, int n_k, const double* positions, const double* kvec, double* cosv,
                        double* sinv) {
    const long long total = (long long)n_atoms * n_k;
    if (total == 0) return;
    lode_phase_kernel<<<(unsigned)((total + LT - 1) / LT), LT, 0, stream>>>(n_atoms, n_k, positions, kvec, cosv, sinv);
}

void launch_lode_type_sums(hipStream_t stream, int n_atoms, int n_k, int S, const int* type_index, const double* cosv,
                           const double* sinv, double* sumc, double* sums) {
    if (n_k == 0) return;
    lode_type_sums_kernel<<<(n_k + LT - 1) / LT, LT, 0, stream>>>(n_atoms, n_k, S, type_index, cosv, sinv, sumc, sums);
}

template <typename T>
void launch_lode_projection(hipStream_t stream, const LodeProjectionArgs<T>& p) {
    if (p.n_k == 0) return;
    lode_projection_kernel<T><<<(p.n_k + 127) / 128, 128, 0, stream>>>(p);
}

template <typename T>
void launch_lode_gemm(hipStream_t stream, int mode, const LodeGemmArgs<T>& g) {
    const long long n_rows = mode == 2 ? (long long)g.n_atoms * g.n_atoms : (long long)g.n_atoms * g.n_types;
    const int ncol = mode == 0 ? g.ncol_v : 3 * g.ncol_v;
    if (n_rows == 0 || ncol == 0) return;
    dim3 grid((unsigned)((n_rows + GT - 1) / GT), (unsigned)((ncol + GT - 1) / GT));
    if (mode == 0) {
        lode_gemm_kernel<T, 0><<<grid, LT, 0, stream>>>(g);
    } else if (mode == 1) {
        lode_gemm_kernel<T, 1><<<grid, LT, 0, stream>>>(g);
    } else {
        lode_gemm_kernel<T, 2><<<grid, LT, 0, stream>>>(g);
    }
}

template void launch_lode_projection<float>(hipStream_t, const LodeProjectionArgs<float>&);
template void launch_lode_projection<double>(hipStream_t, const LodeProjectionArgs<double>&);
template void launch_lode_gemm<float>(hipStream_t, int, const LodeGemmArgs<float>&);
template void launch_lode_gemm<double>(hipStream_t, int, const LodeGemmArgs<double>&);

}  // namespace rascaline_amd
