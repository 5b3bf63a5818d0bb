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

#include <algorithm>

#include "kernels_lode.hpp"
#include "soap_common.hpp"
#include "special.hpp"

namespace rascaline_amd {

namespace {

constexpr int LT = 256;  // threads per workgroup

// cos / sin (k . r_i) for all (i, k) of one system, FP64
__global__ void lode_phase_kernel(int n_atoms, int n_k, const double* __restrict__ positions, const double* __restrict__ kvec,
                                  double* __restrict__ cosv, double* __restrict__ sinv, float* __restrict__ cosf,
                                  float* __restrict__ sinf) {
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
    if (cosf) {
        cosf[idx] = (float)c;
        sinf[idx] = (float)s;
    }
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

// A(row, k) of the three LODE GEMMs, from the FP64 phases (values: the structure
// factor; self rows: -Im / Re minus the centre's own term; pair rows:
// 2 (s_i c_j - c_i s_j) / -2 (c_i c_j + s_i s_j)), rounded once to the GEMM type.
template <typename T, int MODE>
__device__ __forceinline__ double lode_a_value(const LodeGemmArgs<T>& g, long long row, int k) {
    const int N = g.n_atoms, nk = g.n_k;
    const int par = g.parity;
    if (MODE == 2) {
        const int i = (int)(row / N), j = (int)(row - (long long)i * N);
        const double ci = g.cosv[(long long)i * nk + k], si = g.sinv[(long long)i * nk + k];
        const double cj = g.cosv[(long long)j * nk + k], sj = g.sinv[(long long)j * nk + k];
        return par == 0 ? 2.0 * (si * cj - ci * sj) : -2.0 * (ci * cj + si * sj);
    }
    const int i = (int)(row / g.n_types), tl = (int)(row - (long long)i * g.n_types);
    const int t = g.types_present[tl];
    const double ci = g.cosv[(long long)i * nk + k], si = g.sinv[(long long)i * nk + k];
    const double sc = g.sumc[(long long)t * nk + k], ss = g.sums[(long long)t * nk + k];
    const double re = 2.0 * (ci * sc + si * ss), im = 2.0 * (si * sc - ci * ss);
    if (MODE == 0) return par == 0 ? re : im;
    return par == 0 ? -im : re - (t == g.type_index[i] ? 2.0 : 0.0);
}

// Writes GEMM element (row, col) into its output block (values, self rows or pair
// rows; pair rows with j == i are skipped).
template <typename T, int MODE>
__device__ __forceinline__ void lode_store(const LodeGemmArgs<T>& g, long long row, int col, T v) {
    const int N = g.n_atoms;
    int i, t, j = -1;
    if (MODE == 2) {
        i = (int)(row / N);
        j = (int)(row - (long long)i * N);
        if (j == i) return;
        t = g.type_index[j];
    } else {
        i = (int)(row / g.n_types);
        t = g.types_present[(int)(row - (long long)i * g.n_types)];
    }
    const int ct = g.type_index[i];
    const int S = g.n_species;
    const long long pair = (long long)ct * S + t;
    const int cv = MODE == 0 ? col : col % g.ncol_v;
    const int xyz = MODE == 0 ? 0 : col / g.ncol_v;
    const int l = g.col_l[cv];
    const int in_l = g.col_in_l[cv];  // m * n + q within the l segment
    const int width = (2 * l + 1) * g.n;
    const long long block = (long long)l * S * S + pair;
    if (MODE == 0) {
        const long long srow = g.value_row0[pair] + g.type_rank[i];
        if (l == 0) v += g.value_shift[pair * g.n + in_l];
        g.values[g.value_base[block] + srow * width + in_l] = v;
    } else {
        // gradient row: the sample's first row plus the position of atom j (or i)
        const long long first = g.grad_row0[pair] + (long long)g.type_rank[i] * g.rows_per_sample[pair];
        const long long rank = MODE == 1 ? g.below[(long long)i * S + t]
                                         : g.type_rank[j] + ((ct != t && i < j) ? 1 : 0);
        g.grads[g.grad_base[block] + (first + rank) * 3 * width + (long long)xyz * width + in_l] = v;
    }
}

// Tiled GEMM out[row][col] = Σ_k A(row, k) B[k][col] for one parity, 64 x 64 tiles,
// 16 k per step. VALU version (FP64 arithmetic): 4 x 4 outputs per thread.
constexpr int GT = 64, GK = 16;

template <typename T, int MODE>
__device__ __forceinline__ void lode_load_tiles(const LodeGemmArgs<T>& g, long long row0, int col0, int k0,
                                                long long n_rows, int ncol, const T* B, T (*As)[GT + 4],
                                                T (*Bs)[GT + 4]) {
    const int tid = threadIdx.x;
    const int nk = g.n_k;
    // A tile: 64 rows x 16 k, 4 per thread
#pragma unroll
    for (int e = 0; e < 4; e++) {
        const int idx = tid + e * LT;
        const int r = idx / GK, kk = idx % GK;
        const long long row = row0 + r;
        const int k = k0 + kk;
        const double v = (row < n_rows && k < nk) ? lode_a_value<T, MODE>(g, row, k) : 0.0;
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
}

template <typename T, int MODE>
__global__ void __launch_bounds__(LT) lode_gemm_kernel(LodeGemmArgs<T> g) {
    __shared__ T As[GK][GT + 4];
    __shared__ T Bs[GK][GT + 4];
    const int tid = threadIdx.x;
    const int tx = tid % 16, ty = tid / 16;
    const long long row0 = (long long)blockIdx.x * GT;
    const int col0 = blockIdx.y * GT;
    const int N = g.n_atoms, nk = g.n_k;
    const long long n_rows = MODE == 2 ? (long long)N * N : (long long)N * g.n_types;
    const int ncol = MODE == 0 ? g.ncol_v : 3 * g.ncol_v;
    const T* B = MODE == 0 ? g.bv : g.bg;
    T acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) acc[a][b] = 0;

    for (int k0 = 0; k0 < nk; k0 += GK) {
        lode_load_tiles<T, MODE>(g, row0, col0, k0, n_rows, ncol, B, As, Bs);
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
#pragma unroll
    for (int x = 0; x < 4; x++) {
        const long long row = row0 + ty + 16 * x;
        if (row >= n_rows) continue;
#pragma unroll
        for (int y = 0; y < 4; y++) {
            const int col = col0 + tx + 16 * y;
            if (col < ncol) lode_store<T, MODE>(g, row, col, acc[x][y]);
        }
    }
}

// FP32 version on the matrix cores: each wave computes a 32 x 32 quadrant of the
// 64 x 64 tile as 2 x 2 blocks of v_mfma_f32_16x16x4f32 (A[i][k] at lane i + 16 k,
// B[k][j] at lane j + 16 k, D[i][j] at lane j + 16 (i / 4), register i % 4).
typedef float lode_float4 __attribute__((ext_vector_type(4)));

template <int MODE>
__global__ void __launch_bounds__(LT) lode_gemm_mfma_kernel(LodeGemmArgs<float> g) {
    __shared__ float As[GK][GT + 4];
    __shared__ float Bs[GK][GT + 4];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const long long row0 = (long long)blockIdx.x * GT;
    const int col0 = blockIdx.y * GT;
    const int N = g.n_atoms, nk = g.n_k;
    const long long n_rows = MODE == 2 ? (long long)N * N : (long long)N * g.n_types;
    const int ncol = MODE == 0 ? g.ncol_v : 3 * g.ncol_v;
    const float* B = MODE == 0 ? g.bv : g.bg;
    const int rq = (wave >> 1) * 32, cq = (wave & 1) * 32;
    const int i16 = lane & 15, kq = lane >> 4;
    lode_float4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) acc[a][b] = {0.f, 0.f, 0.f, 0.f};

    for (int k0 = 0; k0 < nk; k0 += GK) {
        lode_load_tiles<float, MODE>(g, row0, col0, k0, n_rows, ncol, B, As, Bs);
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < GK; ks += 4) {
            const float a0 = As[ks + kq][rq + i16], a1 = As[ks + kq][rq + 16 + i16];
            const float b0 = Bs[ks + kq][cq + i16], b1 = Bs[ks + kq][cq + 16 + i16];
            acc[0][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b0, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b1, acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b0, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b1, acc[1][1], 0, 0, 0);
        }
        __syncthreads();
    }
#pragma unroll
    for (int rb = 0; rb < 2; rb++)
#pragma unroll
        for (int cb = 0; cb < 2; cb++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const long long row = row0 + rq + 16 * rb + 4 * kq + v;
                const int col = col0 + cq + 16 * cb + i16;
                if (row < n_rows && col < ncol) lode_store<float, MODE>(g, row, col, acc[rb][cb][v]);
            }
}

// Pair rows (i, j != i) in FP32 on the matrix cores, tiled for reuse of the
// neighbours' phases: a workgroup takes 4 centers x 64 neighbours x 64 columns,
// with the k-vectors outermost in chunks of 32. Per chunk the neighbours' phases
// (rounded once from FP64) and B go into LDS and serve all four centers; wave w
// keeps center i0 + w's 64 x 64 block in 16 accumulators and forms its A
// fragments A[j][k] = 2 (s_i c_j - c_i s_j) (even l) or -2 (c_i c_j + s_i s_j)
// (odd l) from them.
constexpr int PJ = 64, PC = 64, PK = 32, PI = 4;

// three waves per SIMD: at most ~168 registers per lane, which the k loop fits when
// unrolled by 4 (fully unrolled, it hoists all LDS reads and needs 212)
__global__ void __launch_bounds__(LT, 3) lode_pair_mfma_kernel(LodeGemmArgs<float> g) {
    __shared__ float cj[PK][PJ + 4], sj[PK][PJ + 4], Bs[PK][PC + 4];
    __shared__ float ci[PI][PK], si[PI][PK];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int N = g.n_atoms, nk = g.n_k;
    const int par = g.parity;
    const int ncol = 3 * g.ncol_v;
    const int i0 = blockIdx.x * PI, j0 = blockIdx.y * PJ, c0 = blockIdx.z * PC;
    const int i16 = lane & 15, kq = lane >> 4;
    lode_float4 acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) acc[a][b] = {0.f, 0.f, 0.f, 0.f};

    // the next chunk's operands in registers: 8 neighbour cosines, sines and B
    // values per thread, and one center phase pair for the first 128 threads
    constexpr int PER = PK * PJ / LT;
    float rc[PER], rs[PER], rb[PER], rci = 0.f, rsi = 0.f;
    auto fetch = [&](int k0) {
#pragma unroll
        for (int e = 0; e < PER; e++) {
            const int idx = tid + e * LT;
            const int kk = idx / PJ, jj = idx - kk * PJ;
            const int k = k0 + kk, j = j0 + jj, col = c0 + jj;
            const bool in = k < nk && j < N;
            rc[e] = in ? g.cosf[(long long)j * nk + k] : 0.f;
            rs[e] = in ? g.sinf[(long long)j * nk + k] : 0.f;
            rb[e] = (k < nk && col < ncol) ? g.bg[(long long)k * ncol + col] : 0.f;
        }
        if (tid < PI * PK) {
            const int ii = tid / PK, kk = tid - ii * PK;
            const int i = i0 + ii, k = k0 + kk;
            const bool in = k < nk && i < N;
            rci = in ? g.cosf[(long long)i * nk + k] : 0.f;
            rsi = in ? g.sinf[(long long)i * nk + k] : 0.f;
        }
    };
    fetch(0);
    for (int k0 = 0; k0 < nk; k0 += PK) {
#pragma unroll
        for (int e = 0; e < PER; e++) {
            const int idx = tid + e * LT;
            const int kk = idx / PJ, jj = idx - kk * PJ;
            cj[kk][jj] = rc[e];
            sj[kk][jj] = rs[e];
            Bs[kk][jj] = rb[e];
        }
        if (tid < PI * PK) {
            const int ii = tid / PK, kk = tid - ii * PK;
            ci[ii][kk] = rci;
            si[ii][kk] = rsi;
        }
        __syncthreads();
        // the following chunk's loads fly while this one is multiplied
        if (k0 + PK < nk) fetch(k0 + PK);
#pragma unroll 4
        for (int ks = 0; ks < PK; ks += 4) {
            const int k = ks + kq;
            const float c_i = ci[wave][k], s_i = si[wave][k];
            float b[4];
#pragma unroll
            for (int cb = 0; cb < 4; cb++) b[cb] = Bs[k][cb * 16 + i16];
#pragma unroll
            for (int jb = 0; jb < 4; jb++) {
                const float c_j = cj[k][jb * 16 + i16], s_j = sj[k][jb * 16 + i16];
                const float a = par == 0 ? 2.0f * (s_i * c_j - c_i * s_j) : -2.0f * (c_i * c_j + s_i * s_j);
#pragma unroll
                for (int cb = 0; cb < 4; cb++)
                    acc[jb][cb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b[cb], acc[jb][cb], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    const int i = i0 + wave;
    if (i >= N) return;
#pragma unroll
    for (int jb = 0; jb < 4; jb++)
#pragma unroll
        for (int cb = 0; cb < 4; cb++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int j = j0 + jb * 16 + 4 * kq + v;
                const int col = c0 + cb * 16 + i16;
                if (j < N && col < ncol) lode_store<float, 2>(g, (long long)i * N + j, col, acc[jb][cb][v]);
            }
}

// ---- vector-Jacobian product of the position gradients (see kernels_lode.hpp)
constexpr int VJP_AT = 8;  // atoms per workgroup of lode_vjp_v_kernel

// V_p[t][x][i][k] for a tile of VJP_AT atoms, one thread per k-vector: each
// thread walks its row of Bg once for all the tile's atoms and directions.
template <typename T>
__global__ void __launch_bounds__(LT) lode_vjp_v_kernel(LodeVjpArgs<T> v, int par, int ncol_max) {
    extern __shared__ __align__(16) unsigned char smem[];
    T* W = reinterpret_cast<T*>(smem);  // [VJP_AT][ncol]
    const int N = v.n_atoms, nk = v.n_k, S = v.n_species;
    const int tl = blockIdx.z, t = v.types_present[tl];
    const int i0 = blockIdx.y * VJP_AT;
    const int ncol = v.ncol_v[par];
    for (int idx = threadIdx.x; idx < VJP_AT * ncol; idx += LT) {
        const int a = idx / ncol, cv = idx - a * ncol;
        const int i = i0 + a;
        T w = 0;
        if (i < N) {
            const int ct = v.type_index[i];
            const long long pair = (long long)ct * S + t;
            const int l = v.col_l[par][cv];
            const int width = (2 * l + 1) * v.n;
            const long long block = (long long)l * S * S + pair;
            w = v.grad_values[v.value_base[block] + (v.value_row0[pair] + v.type_rank[i]) * width + v.col_in_l[par][cv]];
        }
        W[idx] = w;
    }
    __syncthreads();
    const int k = blockIdx.x * LT + threadIdx.x;
    if (k >= nk) return;
    T acc[VJP_AT][3];
#pragma unroll
    for (int a = 0; a < VJP_AT; a++) acc[a][0] = acc[a][1] = acc[a][2] = 0;
    const T* B = v.bg[par] + (long long)k * 3 * ncol;
    for (int cv = 0; cv < ncol; cv++) {
        const T b0 = B[cv], b1 = B[ncol + cv], b2 = B[2 * ncol + cv];
#pragma unroll
        for (int a = 0; a < VJP_AT; a++) {
            const T w = W[a * ncol + cv];
            acc[a][0] += b0 * w;
            acc[a][1] += b1 * w;
            acc[a][2] += b2 * w;
        }
    }
#pragma unroll
    for (int a = 0; a < VJP_AT; a++) {
        const int i = i0 + a;
        if (i >= N) break;
#pragma unroll
        for (int x = 0; x < 3; x++)
            v.V[((((long long)par * v.n_types + tl) * 3 + x) * N + i) * nk + k] = acc[a][x];
    }
}

// U_p[t][x][cos / sin][k] = Σ_i (cos_i, sin_i)(k) V_p[t][x][i][k], in FP64 and atom order
template <typename T>
__global__ void lode_vjp_u_kernel(LodeVjpArgs<T> v) {
    const int N = v.n_atoms, nk = v.n_k;
    const long long total = 2LL * v.n_types * 3 * nk;
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    const int k = (int)(idx % nk);
    const long long ptx = idx / nk;  // (par, tl, x)
    const T* Vr = v.V + ptx * N * nk + k;
    double uc = 0.0, us = 0.0;
    for (int i = 0; i < N; i++) {
        const double w = (double)Vr[(long long)i * nk];
        uc += v.cosv[(long long)i * nk + k] * w;
        us += v.sinv[(long long)i * nk + k] * w;
    }
    v.U[(ptx * 2 + 0) * nk + k] = uc;
    v.U[(ptx * 2 + 1) * nk + k] = us;
}

// dA/dr_j: one wave per atom, lanes over k, FP64 sums with a fixed reduction order.
// Pair rows (i != j), even l: A = 2 (s_i c_j - c_i s_j) -> 2 c_j U_s - 2 s_j U_c; odd l:
// A = -2 (c_i c_j + s_i s_j) -> -2 c_j U_c - 2 s_j U_s, plus 2 V_j for the excluded
// i = j term. Own rows, per type t: -SF_im (even) and SF_re - 2 [t = type j] (odd)
// times V_p[t][x][j].
template <typename T>
__global__ void __launch_bounds__(LT) lode_vjp_atom_kernel(LodeVjpArgs<T> v) {
    const int lane = threadIdx.x & 63;
    const int j = blockIdx.x * (LT / 64) + threadIdx.x / 64;
    if (j >= v.n_atoms) return;
    const int N = v.n_atoms, nk = v.n_k, NT = v.n_types;
    const int tj = v.type_index[j];
    int tlj = 0;
    for (int q = 0; q < NT; q++)
        if (v.types_present[q] == tj) tlj = q;
    double acc[3] = {0.0, 0.0, 0.0};
    auto Vp = [&](int par, int tl, int x, int k) -> double {
        return (double)v.V[((((long long)par * NT + tl) * 3 + x) * N + j) * nk + k];
    };
    auto Up = [&](int par, int tl, int x, int cs, int k) -> double {
        return v.U[((((long long)par * NT + tl) * 3 + x) * 2 + cs) * nk + k];
    };
    for (int k = lane; k < nk; k += 64) {
        const double cj = v.cosv[(long long)j * nk + k], sj = v.sinv[(long long)j * nk + k];
#pragma unroll
        for (int x = 0; x < 3; x++) {
            double g = 2.0 * cj * Up(0, tlj, x, 1, k) - 2.0 * sj * Up(0, tlj, x, 0, k);
            g += -2.0 * cj * Up(1, tlj, x, 0, k) - 2.0 * sj * Up(1, tlj, x, 1, k) + 2.0 * Vp(1, tlj, x, k);
            for (int tl = 0; tl < NT; tl++) {
                const int t = v.types_present[tl];
                const double sc = v.sumc[(long long)t * nk + k], ss = v.sums[(long long)t * nk + k];
                const double im = 2.0 * (sj * sc - cj * ss), re = 2.0 * (cj * sc + sj * ss);
                g += -im * Vp(0, tl, x, k) + (re - (t == tj ? 2.0 : 0.0)) * Vp(1, tl, x, k);
            }
            acc[x] += g;
        }
    }
#pragma unroll
    for (int x = 0; x < 3; x++) {
        double s = acc[x];
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
        acc[x] = s;
    }
    if (lane == 0) {
        for (int x = 0; x < 3; x++) v.out[3LL * j + x] = (T)acc[x];
    }
}

}  // namespace

void launch_lode_phases(hipStream_t stream, int n_atoms, int n_k, const double* positions, const double* kvec, double* cosv,
                        double* sinv, float* cosf, float* sinf) {
    const long long total = (long long)n_atoms * n_k;
    if (total == 0) return;
    lode_phase_kernel<<<(unsigned)((total + LT - 1) / LT), LT, 0, stream>>>(n_atoms, n_k, positions, kvec, cosv, sinv,
                                                                          cosf, sinf);
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
    if constexpr (sizeof(T) == 4) {
        // FP32: products and sums on the matrix cores, FP32 throughout
        if (mode == 0) {
            lode_gemm_mfma_kernel<0><<<grid, LT, 0, stream>>>(g);
        } else if (mode == 1) {
            lode_gemm_mfma_kernel<1><<<grid, LT, 0, stream>>>(g);
        } else {
            const long long n = g.n_atoms;
            dim3 pgrid((unsigned)((n + PI - 1) / PI), (unsigned)((n + PJ - 1) / PJ), (unsigned)((ncol + PC - 1) / PC));
            lode_pair_mfma_kernel<<<pgrid, LT, 0, stream>>>(g);
        }
        return;
    }
    if (mode == 0) {
        lode_gemm_kernel<T, 0><<<grid, LT, 0, stream>>>(g);
    } else if (mode == 1) {
        lode_gemm_kernel<T, 1><<<grid, LT, 0, stream>>>(g);
    } else {
        lode_gemm_kernel<T, 2><<<grid, LT, 0, stream>>>(g);
    }
}

template <typename T>
void launch_lode_vjp(hipStream_t stream, const LodeVjpArgs<T>& v) {
    const int N = v.n_atoms, nk = v.n_k;
    if (N == 0 || nk == 0 || v.n_types == 0) return;
    for (int par = 0; par < 2; par++) {
        const size_t shmem = (size_t)VJP_AT * std::max(v.ncol_v[par], 1) * sizeof(T);
        dim3 grid((unsigned)((nk + LT - 1) / LT), (unsigned)((N + VJP_AT - 1) / VJP_AT), (unsigned)v.n_types);
        lode_vjp_v_kernel<T><<<grid, LT, shmem, stream>>>(v, par, v.ncol_v[par]);
    }
    const long long total = 2LL * v.n_types * 3 * nk;
    lode_vjp_u_kernel<T><<<(unsigned)((total + LT - 1) / LT), LT, 0, stream>>>(v);
    lode_vjp_atom_kernel<T><<<(unsigned)((N + LT / 64 - 1) / (LT / 64)), LT, 0, stream>>>(v);
}

template void launch_lode_projection<float>(hipStream_t, const LodeProjectionArgs<float>&);
template void launch_lode_projection<double>(hipStream_t, const LodeProjectionArgs<double>&);
template void launch_lode_gemm<float>(hipStream_t, int, const LodeGemmArgs<float>&);
template void launch_lode_gemm<double>(hipStream_t, int, const LodeGemmArgs<double>&);
template void launch_lode_vjp<float>(hipStream_t, const LodeVjpArgs<float>&);
template void launch_lode_vjp<double>(hipStream_t, const LodeVjpArgs<double>&);

}  // namespace rascaline_amd
