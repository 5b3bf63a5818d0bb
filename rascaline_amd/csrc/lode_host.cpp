// LODE spherical expansion: host orchestration of the kernels in kernels_lode.hip
// (featomic/src/calculators/lode/spherical_expansion.rs).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "calculator.hpp"
#include "kernels_lode.hpp"
#include "special.hpp"

namespace rascaline_amd {

namespace {

double det3(const double* m) {
    return m[0] * (m[4] * m[8] - m[7] * m[5]) - m[1] * (m[3] * m[8] - m[5] * m[6]) + m[2] * (m[3] * m[7] - m[4] * m[6]);
}

// Half-space k-vectors with |k| < k_cutoff, featomic/src/math/k_vectors.rs:20-83:
// reciprocal vectors b_a = rows of 2 pi H^{-T} (H rows are the cell vectors); the
// enumeration n3 >= 1 (n1 = n2 = 0), then n2 >= 1 with n3 in ±n3max, then n1 >= 1
// with n2, n3 in ±. Output [n_k][4]: unit direction, norm.
std::vector<double> k_vectors(const double* cell, double k_cutoff) {
    const double t[9] = {cell[0], cell[3], cell[6], cell[1], cell[4], cell[7], cell[2], cell[5], cell[8]};
    const double inv_det = 1.0 / det3(t);
    double inv[9];
    inv[0] = (t[4] * t[8] - t[7] * t[5]) * inv_det;
    inv[1] = (t[2] * t[7] - t[1] * t[8]) * inv_det;
    inv[2] = (t[1] * t[5] - t[2] * t[4]) * inv_det;
    inv[3] = (t[5] * t[6] - t[3] * t[8]) * inv_det;
    inv[4] = (t[0] * t[8] - t[2] * t[6]) * inv_det;
    inv[5] = (t[3] * t[2] - t[0] * t[5]) * inv_det;
    inv[6] = (t[3] * t[7] - t[6] * t[4]) * inv_det;
    inv[7] = (t[6] * t[1] - t[0] * t[7]) * inv_det;
    inv[8] = (t[0] * t[4] - t[3] * t[1]) * inv_det;
    double rc[9];
    for (int k = 0; k < 9; k++) rc[k] = 2.0 * RA_PI * inv[k];
    double m[9];
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++)
            m[a * 3 + b] = rc[a * 3 + 0] * rc[b * 3 + 0] + rc[a * 3 + 1] * rc[b * 3 + 1] + rc[a * 3 + 2] * rc[b * 3 + 2];
    const double kv = det3(rc);
    const int n1max = static_cast<int>(std::sqrt(m[4] * m[8] - m[5] * m[5]) / kv * k_cutoff);
    const int n2max = static_cast<int>(std::sqrt(m[8] * m[0] - m[6] * m[6]) / kv * k_cutoff);
    const int n3max = static_cast<int>(std::sqrt(m[0] * m[4] - m[1] * m[1]) / kv * k_cutoff);
    const double* b1 = rc;
    const double* b2 = rc + 3;
    const double* b3 = rc + 6;
    const double cutoff2 = k_cutoff * k_cutoff;
    std::vector<double> out;
    auto push = [&](double kx, double ky, double kz) {
        const double n2 = kx * kx + ky * ky + kz * kz;
        if (n2 < cutoff2) {
            const double norm = std::sqrt(n2);
            out.insert(out.end(), {kx / norm, ky / norm, kz / norm, norm});
        }
    };
    for (int c = 1; c <= n3max; c++) push(c * b3[0], c * b3[1], c * b3[2]);
    for (int b = 1; b <= n2max; b++)
        for (int c = -n3max; c <= n3max; c++)
            push(b * b2[0] + c * b3[0], b * b2[1] + c * b3[1], b * b2[2] + c * b3[2]);
    for (int a = 1; a <= n1max; a++)
        for (int b = -n2max; b <= n2max; b++)
            for (int c = -n3max; c <= n3max; c++)
                push(a * b1[0] + b * b2[0] + c * b3[0], a * b1[1] + b * b2[1] + c * b3[1], a * b1[2] + b * b2[2] + c * b3[2]);
    return out;
}

// Fourier transform of the smeared power-law density,
// lode/spherical_expansion.rs:280-342.
double density_fourier(const LodeParams& p, double k) {
    const double s2 = p.smearing * p.smearing;
    if (p.exponent == 0) return std::pow(4.0 * RA_PI * s2, 0.75) * std::exp(-0.5 * k * k * s2);
    if (p.exponent == 1) {
        const double k2 = k * k;
        return 4.0 * RA_PI * (std::exp(-0.5 * k2 * s2) / k2);
    }
    const double pe = static_cast<double>(p.exponent);
    const double p_eff = 3.0 - pe;
    const double factor = std::pow(RA_PI, 1.5) * std::pow(2.0 * s2, 0.5 * p_eff) / gamma_fn(0.5 * pe);
    const double x = 0.5 * k * k * s2;
    double v = 0.0;
    switch (p.exponent) {
    case 2: v = std::sqrt(RA_PI / x) * std::erfc(std::sqrt(x)); break;
    case 3: v = -expi(-x); break;
    case 4: v = 2.0 * (std::exp(-x) - std::sqrt(RA_PI * x) * std::erfc(std::sqrt(x))); break;
    case 5: v = std::exp(-x) + x * expi(-x); break;
    case 6: v = ((2.0 - 4.0 * x) * std::exp(-x) + 4.0 * std::sqrt(RA_PI) * std::pow(x, 1.5) * std::erfc(std::sqrt(x))) / 3.0; break;
    case 7: v = (1.0 - x) * std::exp(-x) / 2.0 - x * x / 2.0 * expi(-x); break;
    case 8:
        v = -2.0 / 15.0 * ((-3.0 + 2.0 * x - 4.0 * x * x) * std::exp(-x) + 4.0 * std::sqrt(RA_PI) * std::pow(x, 2.5) * std::erfc(std::sqrt(x)));
        break;
    case 9: v = (x * x - x + 2.0) * std::exp(-x) / 6.0 + x * x * x / 6.0 * expi(-x); break;
    default: throw Error::invalid_parameter("LODE is only implemented for potential_exponent < 10");
    }
    return factor * v;
}

// k = 0 term prefactor (lode/spherical_expansion.rs:347-390), non zero for p = 0 and p > 3
double k0_factor(const LodeParams& p) {
    const double s = p.smearing;
    if (p.exponent == 0) {
        const double s2 = s * s;
        return std::pow(2.0 * RA_PI * s2, 1.5) / std::pow(RA_PI * s2, 0.75) / std::sqrt(4.0 * RA_PI);
    }
    const double pe = static_cast<double>(p.exponent);
    const double p_eff = 3.0 - pe;
    return 0.5 * RA_PI * std::pow(2.0, p_eff) / gamma_fn(0.5 * pe) * std::pow(2.0, (pe - 1.0) / 2.0) / -p_eff *
           std::pow(s, -p_eff) / std::pow(s, 2.0 * pe - 6.0);
}

}  // namespace

template <typename T>
mts_tensormap_t* Calculator::run_lode(const HostBatch& batch, const Gradients& grads, hipStream_t stream,
                                      featomic_amd_device_output_t* output) {
    const LodeParams& p = lode_;
    if (!ev_start_) {
        RA_HIP_CHECK(hipEventCreate(&ev_start_));
        RA_HIP_CHECK(hipEventCreate(&ev_stop_));
        RA_HIP_CHECK(hipEventCreate(&ev_main_start_));
        RA_HIP_CHECK(hipEventCreate(&ev_main_stop_));
    }
    RA_HIP_CHECK(hipEventRecord(ev_start_, stream));
    main_ms_sum_ = 0.0;
    main_bytes_ = 0;
    const int L = p.L(), n = p.n_radial;
    if (L > LODE_MAX_L || n > LODE_MAX_N) {
        throw Error::invalid_parameter("max_angular must be < " + std::to_string(LODE_MAX_L) + " and max_radial < " +
                                       std::to_string(LODE_MAX_N) + " for the GPU LODE kernels");
    }
    const Species sp = species_of(batch);
    const int S = static_cast<int>(sp.types.size());
    const size_t n_sys = batch.n_systems();
    if (batch.positions.empty() && !batch.device_positions && batch.n_atoms() > 0) {
        throw Error::internal("LODE needs the positions");
    }

    // types present per system; keys are the type pairs found together in a system
    std::vector<std::vector<int>> count(n_sys, std::vector<int>(S, 0));
    std::vector<char> pair_present(static_cast<size_t>(S) * S, 0);
    for (size_t s = 0; s < n_sys; s++) {
        for (int64_t a = batch.offsets[s]; a < batch.offsets[s + 1]; a++) count[s][sp.index[a]]++;
        for (int a = 0; a < S; a++)
            for (int b = 0; b < S; b++)
                if (count[s][a] > 0 && count[s][b] > 0) pair_present[a * S + b] = 1;
    }
    // rows: per (system, pair) the first sample / gradient row within the block
    std::vector<long long> vrow0(n_sys * S * S, 0), grow0(n_sys * S * S, 0), rps(n_sys * S * S, 0);
    std::vector<long long> n_samples(S * S, 0), n_rows(S * S, 0);
    for (size_t s = 0; s < n_sys; s++) {
        for (int pr = 0; pr < S * S; pr++) {
            const int ct = pr / S, nt = pr % S;
            if (!pair_present[pr] || count[s][nt] == 0) continue;
            vrow0[s * S * S + pr] = n_samples[pr];
            grow0[s * S * S + pr] = n_rows[pr];
            rps[s * S * S + pr] = count[s][nt] + (ct != nt ? 1 : 0);
            n_samples[pr] += count[s][ct];
            n_rows[pr] += static_cast<long long>(count[s][ct]) * rps[s * S * S + pr];
        }
    }
    // block element bases, blocks in (ct, nt)-major, l-minor order
    std::vector<long long> value_base(static_cast<size_t>(L) * S * S, 0), grad_base(static_cast<size_t>(L) * S * S, 0);
    long long values_size = 0, grads_size = 0;
    for (int pr = 0; pr < S * S; pr++) {
        if (!pair_present[pr]) continue;
        for (int l = 0; l < L; l++) {
            value_base[static_cast<size_t>(l) * S * S + pr] = values_size;
            grad_base[static_cast<size_t>(l) * S * S + pr] = grads_size;
            values_size += n_samples[pr] * (2 * l + 1) * n;
            if (grads.positions) grads_size += n_rows[pr] * 3 * (2 * l + 1) * n;
        }
    }

    // columns of each parity: (l, m, n) for the l of that parity
    std::vector<int> col_l[2], col_in_l[2];
    for (int l = 0; l < L; l++)
        for (int m = 0; m < 2 * l + 1; m++)
            for (int q = 0; q < n; q++) {
                col_l[l % 2].push_back(l);
                col_in_l[l % 2].push_back(m * n + q);
            }
    const int ncol_v[2] = {static_cast<int>(col_l[0].size()), static_cast<int>(col_l[1].size())};

    const DeviceTables<T>& tab = device_tables<T>();
    T* d_values = ws_.get<T>("lode_values", std::max<long long>(values_size, 1));
    T* d_grads = grads.positions ? ws_.get<T>("lode_grads", std::max<long long>(grads_size, 1)) : nullptr;
    long long* d_value_base = ws_.get<long long>("lode_value_base", value_base.size());
    long long* d_grad_base = ws_.get<long long>("lode_grad_base", grad_base.size());
    upload(d_value_base, value_base, stream);
    upload(d_grad_base, grad_base, stream);
    int* d_col_l[2];
    int* d_col_in_l[2];
    for (int par = 0; par < 2; par++) {
        d_col_l[par] = ws_.get<int>(par == 0 ? "lode_col_l0" : "lode_col_l1", std::max<size_t>(col_l[par].size(), 1));
        d_col_in_l[par] = ws_.get<int>(par == 0 ? "lode_col_m0" : "lode_col_m1", std::max<size_t>(col_l[par].size(), 1));
        upload(d_col_l[par], col_l[par], stream);
        upload(d_col_in_l[par], col_in_l[par], stream);
    }

    const bool k0 = p.exponent == 0 || p.exponent > 3;
    std::vector<double> ri0(n, 0.0);
    if (k0) tables_.exact(0, 0.0, ri0.data(), nullptr);
    const double kfac = k0 ? k0_factor(p) : 0.0;

    for (size_t s = 0; s < n_sys; s++) {
        const int N = static_cast<int>(batch.offsets[s + 1] - batch.offsets[s]);
        if (N == 0) continue;
        const double* cell = batch.cells.data() + 9 * s;
        bool zero = true;
        for (int k = 0; k < 9; k++) zero = zero && cell[k] == 0.0;
        if (zero) throw Error::invalid_parameter("LODE can only be used with periodic systems");
        const std::vector<double> kv = k_vectors(cell, p.k_cutoff);
        const int nk = static_cast<int>(kv.size() / 4);
        if (nk == 0) throw Error::invalid_parameter("No k-vectors for current combination of hyper parameters.");
        std::vector<double> df(nk);
        for (int k = 0; k < nk; k++) df[k] = density_fourier(p, kv[4 * k + 3]);
        const double volume = std::fabs(det3(cell));
        const double gf = 4.0 * RA_PI / volume;

        // per-atom bookkeeping
        const int64_t a0 = batch.offsets[s];
        std::vector<int> tidx(N), trank(N), below(static_cast<size_t>(N) * S);
        std::vector<int> seen(S, 0);
        for (int a = 0; a < N; a++) {
            const int t = sp.index[a0 + a];
            tidx[a] = t;
            for (int u = 0; u < S; u++) below[static_cast<size_t>(a) * S + u] = seen[u];
            trank[a] = seen[t]++;
        }
        std::vector<int> present;
        for (int t = 0; t < S; t++)
            if (count[s][t] > 0) present.push_back(t);
        std::vector<long long> s_vrow0(vrow0.begin() + s * S * S, vrow0.begin() + (s + 1) * S * S);
        std::vector<long long> s_grow0(grow0.begin() + s * S * S, grow0.begin() + (s + 1) * S * S);
        std::vector<long long> s_rps(rps.begin() + s * S * S, rps.begin() + (s + 1) * S * S);
        // constant l = 0 terms: centre (same type) and k = 0 (p = 0, p > 3)
        std::vector<T> shift(static_cast<size_t>(S) * S * n, static_cast<T>(0));
        for (int ct = 0; ct < S; ct++)
            for (int nt = 0; nt < S; nt++)
                for (int q = 0; q < n; q++) {
                    double v = 0.0;
                    if (ct == nt) v -= (1.0 - p.center_atom_weight) * self_contribution_[q];
                    if (k0) v += count[s][nt] * (gf * (kfac * ri0[q]));
                    shift[(static_cast<size_t>(ct) * S + nt) * n + q] = static_cast<T>(v);
                }

        const double* d_pos = batch.device_positions ? batch.device_positions + 3 * a0 : nullptr;
        if (!d_pos) {
            double* staged = ws_.get<double>("lode_pos", 3 * static_cast<size_t>(N));
            RA_HIP_CHECK(hipMemcpyAsync(staged, batch.positions.data() + 3 * a0, sizeof(double) * 3 * N,
                                        hipMemcpyHostToDevice, stream));
            d_pos = staged;
        }
        double* d_kv = ws_.get<double>("lode_kv", kv.size());
        double* d_df = ws_.get<double>("lode_df", df.size());
        int* d_tidx = ws_.get<int>("lode_tidx", N);
        int* d_trank = ws_.get<int>("lode_trank", N);
        int* d_below = ws_.get<int>("lode_below", below.size());
        int* d_present = ws_.get<int>("lode_present", present.size());
        long long* d_vrow0 = ws_.get<long long>("lode_vrow0", s_vrow0.size());
        long long* d_grow0 = ws_.get<long long>("lode_grow0", s_grow0.size());
        long long* d_rps = ws_.get<long long>("lode_rps", s_rps.size());
        T* d_shift = ws_.get<T>("lode_shift", shift.size());
        upload(d_kv, kv, stream);
        upload(d_df, df, stream);
        upload(d_tidx, tidx, stream);
        upload(d_trank, trank, stream);
        upload(d_below, below, stream);
        upload(d_present, present, stream);
        upload(d_vrow0, s_vrow0, stream);
        upload(d_grow0, s_grow0, stream);
        upload(d_rps, s_rps, stream);
        upload(d_shift, shift, stream);

        double* d_cos = ws_.get<double>("lode_cos", static_cast<size_t>(N) * nk);
        double* d_sin = ws_.get<double>("lode_sin", static_cast<size_t>(N) * nk);
        double* d_sumc = ws_.get<double>("lode_sumc", static_cast<size_t>(S) * nk);
        double* d_sums = ws_.get<double>("lode_sums", static_cast<size_t>(S) * nk);
        float* d_cosf = nullptr;
        float* d_sinf = nullptr;
        if (sizeof(T) == 4) {
            d_cosf = ws_.get<float>("lode_cosf", static_cast<size_t>(N) * nk);
            d_sinf = ws_.get<float>("lode_sinf", static_cast<size_t>(N) * nk);
        }
        launch_lode_phases(stream, N, nk, d_pos, d_kv, d_cos, d_sin, d_cosf, d_sinf);
        launch_lode_type_sums(stream, N, nk, S, d_tidx, d_cos, d_sin, d_sumc, d_sums);

        LodeProjectionArgs<T> pa;
        pa.n_k = nk;
        pa.L = L;
        pa.n = n;
        pa.kvec = d_kv;
        pa.density_fourier = d_df;
        pa.global_factor = gf;
        pa.radial = tab.radial;
        pa.sh_a = tab.sh_a;
        pa.sh_b = tab.sh_b;
        for (int par = 0; par < 2; par++) {
            pa.ncol_v[par] = ncol_v[par];
            pa.bv[par] = ws_.get<T>(par == 0 ? "lode_bv0" : "lode_bv1", std::max<size_t>(static_cast<size_t>(nk) * ncol_v[par], 1));
            pa.bg[par] = ws_.get<T>(par == 0 ? "lode_bg0" : "lode_bg1", std::max<size_t>(static_cast<size_t>(nk) * 3 * ncol_v[par], 1));
        }
        launch_lode_projection<T>(stream, pa);

        if (vjp_grad_) {
            // vector-Jacobian product of the position gradients: O(N n_k) sums instead
            // of the N^2 gradient rows
            const int NT = static_cast<int>(present.size());
            LodeVjpArgs<T> v;
            v.n_atoms = N;
            v.n_k = nk;
            v.n_species = S;
            v.n_types = NT;
            v.n = n;
            v.types_present = d_present;
            v.type_index = d_tidx;
            v.type_rank = d_trank;
            v.cosv = d_cos;
            v.sinv = d_sin;
            v.sumc = d_sumc;
            v.sums = d_sums;
            for (int par = 0; par < 2; par++) {
                v.bg[par] = pa.bg[par];
                v.ncol_v[par] = ncol_v[par];
                v.col_l[par] = d_col_l[par];
                v.col_in_l[par] = d_col_in_l[par];
            }
            v.grad_values = static_cast<const T*>(vjp_grad_);
            v.value_base = d_value_base;
            v.value_row0 = d_vrow0;
            v.V = ws_.get<T>("lode_vjp_v", std::max<size_t>(2ULL * NT * 3 * static_cast<size_t>(N) * nk, 1));
            v.U = ws_.get<double>("lode_vjp_u", std::max<size_t>(2ULL * NT * 3 * 2 * static_cast<size_t>(nk), 1));
            v.out = static_cast<T*>(vjp_out_) + 3 * a0;
            RA_HIP_CHECK(hipEventRecord(ev_main_start_, stream));
            launch_lode_vjp<T>(stream, v);
            RA_HIP_CHECK(hipEventRecord(ev_main_stop_, stream));
            RA_HIP_CHECK(hipGetLastError());
            // the per-system scratch buffers are reused by the next system
            RA_HIP_CHECK(hipStreamSynchronize(stream));
            float ms = 0.0f;
            RA_HIP_CHECK(hipEventElapsedTime(&ms, ev_main_start_, ev_main_stop_));
            main_ms_sum_ += ms;
            // algorithmic flops of the V products (the dominant work; the reported count
            // of the LODE main kernels is in flops): 2 NT 3 N n_k (columns of both parities)
            main_bytes_ += 2LL * NT * 3 * static_cast<long long>(N) * nk * (ncol_v[0] + ncol_v[1]);
            continue;
        }

        auto gemm_args = [&](int par) {
            LodeGemmArgs<T> g;
            g.n_atoms = N;
            g.n_k = nk;
            g.n_species = S;
            g.n_types = static_cast<int>(present.size());
            g.parity = par;
            g.ncol_v = ncol_v[par];
            g.n = n;
            g.types_present = d_present;
            g.type_index = d_tidx;
            g.type_rank = d_trank;
            g.below = d_below;
            g.cosv = d_cos;
            g.sinv = d_sin;
            g.cosf = d_cosf;
            g.sinf = d_sinf;
            g.sumc = d_sumc;
            g.sums = d_sums;
            g.bv = pa.bv[par];
            g.bg = pa.bg[par];
            g.col_l = d_col_l[par];
            g.col_in_l = d_col_in_l[par];
            g.values = d_values;
            g.value_base = d_value_base;
            g.value_row0 = d_vrow0;
            g.value_shift = d_shift;
            g.grads = d_grads;
            g.grad_base = d_grad_base;
            g.grad_row0 = d_grow0;
            g.rows_per_sample = d_rps;
            return g;
        };
        for (int par = 0; par < 2; par++) {
            if (ncol_v[par] == 0) continue;
            LodeGemmArgs<T> g = gemm_args(par);
            launch_lode_gemm<T>(stream, 0, g);
            if (grads.positions) launch_lode_gemm<T>(stream, 1, g);
        }
        if (grads.positions) {
            // the pair rows (i, j != i): the dominant GEMM, timed on the compute stream
            RA_HIP_CHECK(hipEventRecord(ev_main_start_, stream));
            for (int par = 0; par < 2; par++) {
                if (ncol_v[par] == 0) continue;
                LodeGemmArgs<T> g = gemm_args(par);
                launch_lode_gemm<T>(stream, 2, g);
            }
            RA_HIP_CHECK(hipEventRecord(ev_main_stop_, stream));
            main_bytes_ += 2LL * nk * 3 * (ncol_v[0] + ncol_v[1]) * static_cast<long long>(N) * (N - 1);
        }
        RA_HIP_CHECK(hipGetLastError());
        // the per-system scratch buffers are reused by the next system
        RA_HIP_CHECK(hipStreamSynchronize(stream));
        if (grads.positions) {
            float ms = 0.0f;
            RA_HIP_CHECK(hipEventElapsedTime(&ms, ev_main_start_, ev_main_stop_));
            main_ms_sum_ += ms;
        }
    }
    RA_HIP_CHECK(hipEventRecord(ev_stop_, stream));
    if (vjp_grad_) return nullptr;

    if (output) {
        // outputs stay in the device buffers; blocks in (center type, neighbour type)-major,
        // l-minor order, like the TensorMap's keys
        DeviceResult r;
        r.key_size = 4;
        for (int pr = 0; pr < S * S; pr++) {
            if (!pair_present[pr]) continue;
            const int ct = pr / S, nt = pr % S;
            for (int l = 0; l < L; l++) {
                r.keys.insert(r.keys.end(), {l, 1, sp.types[ct], sp.types[nt]});
                r.n_samples.push_back(n_samples[pr]);
                r.n_rows.push_back(grads.positions ? n_rows[pr] : 0);
                r.value_offsets.push_back(value_base[static_cast<size_t>(l) * S * S + pr]);
                r.gradient_offsets.push_back(grads.positions ? grad_base[static_cast<size_t>(l) * S * S + pr] : 0);
                r.block_properties.push_back((2LL * l + 1) * n);
            }
        }
        device_result_ = std::move(r);
        output->n_blocks = device_result_.n_samples.size();
        output->keys = device_result_.keys.data();
        output->key_size = device_result_.key_size;
        output->n_samples = device_result_.n_samples.data();
        output->n_gradient_rows = device_result_.n_rows.data();
        output->value_offsets = device_result_.value_offsets.data();
        output->gradient_offsets = device_result_.gradient_offsets.data();
        output->block_properties = device_result_.block_properties.data();
        output->values = d_values;
        output->gradients = d_grads;
        output->sample_labels = nullptr;
        output->gradient_labels = nullptr;
        output->sample_label_offsets = nullptr;
        output->gradient_label_offsets = nullptr;
        output->cell_gradients = nullptr;
        output->cell_gradient_offsets = nullptr;
        output->precision = sizeof(T) == 4 ? 32 : 64;
        return nullptr;
    }

    // ---- assemble the TensorMap; the blocks' float64 arrays are written by the staged
    // transfers at the end
    std::vector<HostSegment> seg_values, seg_grads;

    auto tm = std::make_unique<mts_tensormap_t>();
    std::vector<int32_t> keys;
    const LabelsData xyz = LabelsData::range("xyz", 3);
    const LabelsData properties = LabelsData::range("n", n);
    for (int pr = 0; pr < S * S; pr++) {
        if (!pair_present[pr]) continue;
        const int ct = pr / S, nt = pr % S;
        // samples (system, atom) and gradient samples (sample, system, atom)
        std::vector<int32_t> sl, gl;
        int32_t sample = 0;
        for (size_t s = 0; s < n_sys; s++) {
            if (count[s][nt] == 0) continue;
            const int64_t a0 = batch.offsets[s];
            const int N = static_cast<int>(batch.offsets[s + 1] - a0);
            for (int i = 0; i < N; i++) {
                if (sp.index[a0 + i] != ct) continue;
                sl.insert(sl.end(), {static_cast<int32_t>(s), i});
                for (int j = 0; j < N; j++) {
                    if (sp.index[a0 + j] == nt || j == i) gl.insert(gl.end(), {sample, static_cast<int32_t>(s), j});
                }
                sample++;
            }
        }
        const LabelsData samples({"system", "atom"}, sl);
        const LabelsData gsamples({"sample", "system", "atom"}, gl);
        for (int l = 0; l < L; l++) {
            keys.insert(keys.end(), {l, 1, sp.types[ct], sp.types[nt]});
            const long long ns = n_samples[pr], nr = n_rows[pr];
            const long long w = (2LL * l + 1) * n;
            auto block = make_block(samples, {mu_labels(l)}, properties,
                                    {static_cast<uintptr_t>(ns), static_cast<uintptr_t>(2 * l + 1), static_cast<uintptr_t>(n)});
            const long long vb = value_base[static_cast<size_t>(l) * S * S + pr];
            seg_values.push_back({vb, ns * w, block->array.data.data()});
            if (grads.positions) {
                auto g = make_block(gsamples, {xyz, mu_labels(l)}, properties,
                                    {static_cast<uintptr_t>(nr), 3, static_cast<uintptr_t>(2 * l + 1), static_cast<uintptr_t>(n)});
                const long long gb = grad_base[static_cast<size_t>(l) * S * S + pr];
                seg_grads.push_back({gb, nr * 3 * w, g->array.data.data()});
                block->add_gradient("positions", std::move(g));
            }
            tm->blocks.push_back(std::move(block));
        }
    }
    download_to_blocks<T>(d_values, std::move(seg_values), stream, staging_);
    if (d_grads) download_to_blocks<T>(d_grads, std::move(seg_grads), stream, staging_);
    RA_HIP_CHECK(hipStreamSynchronize(stream));
    tm->keys = LabelsData({"o3_lambda", "o3_sigma", "center_type", "neighbor_type"}, keys);
    return tm.release();
}

template mts_tensormap_t* Calculator::run_lode<float>(const HostBatch&, const Gradients&, hipStream_t,
                                                     featomic_amd_device_output_t*);
template mts_tensormap_t* Calculator::run_lode<double>(const HostBatch&, const Gradients&, hipStream_t,
                                                      featomic_amd_device_output_t*);

}  // namespace rascaline_amd
