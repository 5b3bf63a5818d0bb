#include "radial.hpp"

#include <algorithm>
#include <cmath>
#include <map>
#include <string>

#include "special.hpp"

namespace rascaline_amd {

namespace {

// Jacobi eigen-decomposition of a small symmetric matrix in extended precision;
// returns eigenvalues, `a` is replaced by the eigenvectors (columns).
std::vector<long double> jacobi_eigen(std::vector<long double>& a, int n) {
    std::vector<long double> v(n * n, 0.0L);
    for (int i = 0; i < n; i++) v[i * n + i] = 1.0L;
    for (int sweep = 0; sweep < 100; sweep++) {
        long double off = 0.0L;
        for (int p = 0; p < n; p++)
            for (int q = p + 1; q < n; q++) off += a[p * n + q] * a[p * n + q];
        if (off < 1e-4000L) break;
        for (int p = 0; p < n; p++) {
            for (int q = p + 1; q < n; q++) {
                const long double apq = a[p * n + q];
                if (fabsl(apq) < 1e-4000L) continue;
                const long double theta = (a[q * n + q] - a[p * n + p]) / (2.0L * apq);
                const long double t = (theta >= 0 ? 1.0L : -1.0L) / (fabsl(theta) + sqrtl(theta * theta + 1.0L));
                const long double c = 1.0L / sqrtl(t * t + 1.0L), sn = t * c;
                for (int k = 0; k < n; k++) {
                    const long double akp = a[k * n + p], akq = a[k * n + q];
                    a[k * n + p] = c * akp - sn * akq;
                    a[k * n + q] = sn * akp + c * akq;
                }
                for (int k = 0; k < n; k++) {
                    const long double apk = a[p * n + k], aqk = a[q * n + k];
                    a[p * n + k] = c * apk - sn * aqk;
                    a[q * n + k] = sn * apk + c * aqk;
                }
                for (int k = 0; k < n; k++) {
                    const long double vkp = v[k * n + p], vkq = v[k * n + q];
                    v[k * n + p] = c * vkp - sn * vkq;
                    v[k * n + q] = sn * vkp + c * vkq;
                }
            }
        }
    }
    std::vector<long double> w(n);
    for (int i = 0; i < n; i++) w[i] = a[i * n + i];
    a = v;
    return w;
}

void hermite(const SplineChannel& ch, int n, double x, double* values, double* grads) {
    const int K = static_cast<int>(ch.positions.size()) - 1;
    // last knot <= x; evaluating exactly at the last knot uses the previous interval
    int k = static_cast<int>(std::upper_bound(ch.positions.begin(), ch.positions.end(), x) - ch.positions.begin()) - 1;
    if (k < 0) k = 0;
    if (k >= K) k = K - 1;
    const double x0 = ch.positions[k], x1 = ch.positions[k + 1];
    const double delta = x1 - x0;
    const double t = (x - x0) / delta, t2 = t * t, t3 = t2 * t;
    const double h00 = 2.0 * t3 - 3.0 * t2 + 1.0, h10 = t3 - 2.0 * t2 + t;
    const double h01 = -2.0 * t3 + 3.0 * t2, h11 = t3 - t2;
    const double* p0 = &ch.values[k * n];
    const double* p1 = &ch.values[(k + 1) * n];
    const double* m0 = &ch.derivatives[k * n];
    const double* m1 = &ch.derivatives[(k + 1) * n];
    for (int i = 0; i < n; i++) values[i] = h00 * p0[i] + h10 * delta * m0[i] + h01 * p1[i] + h11 * delta * m1[i];
    if (grads) {
        const double d00 = 6.0 * (t2 - t), d10 = 3.0 * t2 - 4.0 * t + 1.0, d01 = -d00, d11 = 3.0 * t2 - 2.0 * t;
        for (int i = 0; i < n; i++) grads[i] = d00 * p0[i] / delta + d10 * m0[i] + d01 * p1[i] / delta + d11 * m1[i];
    }
}

SplineChannel build_channel(const RadialTables& t, int l, double accuracy) {
    const int n = t.n;
    SplineChannel ch;
    const double step = t.stop / 10.0;
    for (int k = 0; k <= 10; k++) {
        const double x = k < 10 ? k * step : t.stop;
        ch.positions.push_back(x);
        ch.values.resize((k + 1) * n);
        ch.derivatives.resize((k + 1) * n);
        t.exact(l, x, &ch.values[k * n], &ch.derivatives[k * n]);
    }
    std::vector<double> interp(n), nv, nd;
    for (;;) {
        const int count = static_cast<int>(ch.positions.size());
        const int n_new = count - 1;
        std::vector<double> npos(n_new);
        nv.assign(n_new * n, 0.0);
        nd.assign(n_new * n, 0.0);
        double mean_abs = 0.0, mean_rel = 0.0;
        for (int k = 0; k < n_new; k++) {
            const double x = (ch.positions[k] + ch.positions[k + 1]) / 2.0;
            npos[k] = x;
            t.exact(l, x, &nv[k * n], &nd[k * n]);
            hermite(ch, n, x, interp.data(), nullptr);
            for (int i = 0; i < n; i++) {
                const double v = nv[k * n + i];
                mean_abs += std::fabs(interp[i] - v);
                mean_rel += std::fabs((interp[i] - v) / v);
            }
        }
        mean_abs /= static_cast<double>(n_new * n);
        mean_rel /= static_cast<double>(n_new * n);
        if (mean_abs < accuracy || mean_rel < accuracy) break;
        if (count + n_new > 10000) {
            throw Error::internal("failed to reach requested accuracy (" + std::to_string(accuracy) +
                                  ") in spline interpolation, mean absolute error is " + std::to_string(mean_abs) +
                                  " and mean relative error is " + std::to_string(mean_rel));
        }
        SplineChannel next;
        next.positions.resize(count + n_new);
        next.values.resize((count + n_new) * n);
        next.derivatives.resize((count + n_new) * n);
        for (int k = 0; k < count; k++) {
            next.positions[2 * k] = ch.positions[k];
            std::copy(&ch.values[k * n], &ch.values[k * n] + n, &next.values[2 * k * n]);
            std::copy(&ch.derivatives[k * n], &ch.derivatives[k * n] + n, &next.derivatives[2 * k * n]);
            if (k < n_new) {
                next.positions[2 * k + 1] = npos[k];
                std::copy(&nv[k * n], &nv[k * n] + n, &next.values[(2 * k + 1) * n]);
                std::copy(&nd[k * n], &nd[k * n] + n, &next.derivatives[(2 * k + 1) * n]);
            }
        }
        ch = std::move(next);
    }
    return ch;
}

}  // namespace

std::vector<double> gto_orthonormalization(double radius, int n) {
    // The overlap matrix is ill-conditioned (condition number ~3e4 for 8 functions,
    // ~4e7 for 12), so S and S^{-1/2} are evaluated in extended precision.
    std::vector<long double> sigma(n);
    for (int i = 0; i < n; i++) {
        const long double s = sqrtl(static_cast<long double>(i));
        sigma[i] = static_cast<long double>(radius) * (s > 1.0L ? s : 1.0L) / static_cast<long double>(n);
    }
    std::vector<long double> S(n * n);
    for (int a = 0; a < n; a++) {
        for (int b = a; b < n; b++) {
            const long double sa = sigma[a], sb = sigma[b];
            const long double e = 0.5L * (3.0L + a + b);
            const long double v = powl(0.5L / (sa * sa) + 0.5L / (sb * sb), -e) / (powl(sa, a) * powl(sb, b)) * tgammal(e) /
                                  (powl(sa * sb, 1.5L) * sqrtl(tgammal(a + 1.5L) * tgammal(b + 1.5L)));
            S[a * n + b] = v;
            S[b * n + a] = v;
        }
    }
    std::vector<long double> w = jacobi_eigen(S, n);
    for (int i = 0; i < n; i++) {
        if (w[i] <= 2.220446049250313e-16L) {
            throw Error::internal("radial overlap matrix is singular, try with a lower max_radial (current value is " +
                                  std::to_string(n - 1) + ")");
        }
        w[i] = 1.0L / sqrtl(w[i]);
    }
    std::vector<double> O(n * n);
    for (int a = 0; a < n; a++) {
        for (int b = 0; b < n; b++) {
            const long double norm_b = sqrtl(2.0L / (powl(sigma[b], 2 * b + 3) * tgammal(b + 1.5L)));
            long double acc = 0.0L;
            for (int k = 0; k < n; k++) acc += S[a * n + k] * w[k] * S[b * n + k];
            O[a * n + b] = static_cast<double>(acc * norm_b);
        }
    }
    return O;
}

void RadialTables::exact(int l, double x, double* values, double* grads) const {
    double raw_v[64], raw_g[64];
    for (int i = 0; i < n; i++) {
        if (lode) {
            lode_gto_raw(l, i, x, gto[i], &raw_v[i], &raw_g[i]);
        } else {
            soap_gto_raw(l, i, x, sigma2_atomic, gto[i], &raw_v[i], &raw_g[i]);
        }
    }
    for (int j = 0; j < n; j++) {
        double v = 0.0, g = 0.0;
        for (int k = 0; k < n; k++) {
            v += raw_v[k] * ortho[j * n + k];
            g += raw_g[k] * ortho[j * n + k];
        }
        values[j] = v;
        if (grads) grads[j] = g;
    }
}

namespace {

// GTO x Gaussian tables with n radial functions (every channel 0..max_angular)
RadialTables soap_gto_tables(const SoapParams& p, int n, bool splined) {
    RadialTables t;
    t.lode = false;
    t.max_angular = p.max_angular;
    t.n = n;
    t.stop = p.cutoff;
    t.sigma2_atomic = p.gaussian_width * p.gaussian_width;
    t.gto.resize(n);
    for (int i = 0; i < n; i++) {
        const double s = gto_width(p.cutoff, i, n);
        t.gto[i] = 1.0 / (2.0 * s * s);
    }
    t.ortho = gto_orthonormalization(p.cutoff, n);
    t.use_spline = splined;
    return t;
}

// a channel of m radial functions widened to n (zero functions after the first m)
SplineChannel widen(const SplineChannel& ch, int m, int n) {
    if (m == n) return ch;
    SplineChannel out;
    out.positions = ch.positions;
    const size_t K = ch.positions.size();
    out.values.assign(K * n, 0.0);
    out.derivatives.assign(K * n, 0.0);
    for (size_t k = 0; k < K; k++) {
        for (int i = 0; i < m; i++) {
            out.values[k * n + i] = ch.values[k * m + i];
            out.derivatives[k * n + i] = ch.derivatives[k * m + i];
        }
    }
    return out;
}

}  // namespace

void evaluate_channel(const SplineChannel& ch, int n, double x, double* values, double* grads) {
    hermite(ch, n, x, values, grads);
}

RadialTables build_soap_radial(const SoapParams& p) {
    const bool one_gto = !p.padded() && !p.radial[0].tabulated;
    if (one_gto) {
        // one GTO basis for channels 0..max_angular: splined or exact
        RadialTables t = soap_gto_tables(p, p.n_radial, p.use_spline);
        if (t.use_spline) {
            for (int l = 0; l <= p.max_angular; l++) t.channels.push_back(build_channel(t, l, p.spline_accuracy));
        }
        return t;
    }
    // an Explicit or tabulated basis: one spline per channel, from that channel's own
    // radial basis (a tabulated one is used as given, SoapRadialIntegralSpline::
    // from_tabulated, soap/radial_integral/spline.rs:46-58), widened to the largest
    // radial size; a channel missing from the basis is zero
    RadialTables t;
    t.lode = false;
    t.max_angular = p.max_angular;
    t.n = p.n_radial;
    t.stop = p.cutoff;
    t.sigma2_atomic = p.gaussian_width * p.gaussian_width;
    t.use_spline = true;
    std::map<int, RadialTables> gto_by_size;
    for (int l = 0; l <= p.max_angular; l++) {
        const RadialSpec* spec = p.radial_of(l);
        SplineChannel ch;
        if (!spec) {
            ch.positions = {0.0, p.cutoff};
            ch.values.assign(2 * static_cast<size_t>(t.n), 0.0);
            ch.derivatives.assign(2 * static_cast<size_t>(t.n), 0.0);
        } else if (spec->tabulated) {
            SplineChannel given;
            given.positions = spec->positions;
            given.values = spec->values;
            given.derivatives = spec->derivatives;
            ch = widen(given, spec->n, t.n);
        } else {
            auto it = gto_by_size.find(spec->n);
            if (it == gto_by_size.end()) it = gto_by_size.emplace(spec->n, soap_gto_tables(p, spec->n, true)).first;
            ch = widen(build_channel(it->second, l, p.spline_accuracy), spec->n, t.n);
        }
        t.channels.push_back(std::move(ch));
    }
    return t;
}

RadialTables build_lode_radial(const LodeParams& p) {
    RadialTables t;
    t.lode = true;
    t.max_angular = p.max_angular;
    t.n = p.n_radial;
    t.stop = p.k_cutoff;
    t.gto.resize(t.n);
    for (int i = 0; i < t.n; i++) t.gto[i] = gto_width(p.radius, i, t.n);
    t.ortho = gto_orthonormalization(p.radius, t.n);
    t.use_spline = p.use_spline;
    if (t.use_spline) {
        for (int l = 0; l <= p.max_angular; l++) t.channels.push_back(build_channel(t, l, p.spline_accuracy));
    }
    return t;
}

std::vector<double> lode_center_contribution(const LodeParams& p, const RadialTables& t) {
    const int n = t.n;
    std::vector<double> contrib(n), out(n);
    const double s = p.smearing, e = static_cast<double>(p.exponent);
    for (int i = 0; i < n; i++) {
        const double n_eff = 0.5 * (3.0 + i);
        const double w = t.gto[i];
        if (p.exponent == 0) {
            const double factor = std::pow(RA_PI, -0.25) / std::pow(s * s, 0.75);
            const double alpha = 0.5 * (1.0 / (s * s) + 1.0 / (w * w));
            contrib[i] = factor * gamma_fn(n_eff) / std::pow(alpha, n_eff);
        } else {
            const double factor = 2.0 * std::sqrt(4.0 * RA_PI) / gamma_fn(e / 2.0) / e;
            const double r = s / w;
            const double arg = 1.0 / (1.0 + r * r);
            contrib[i] = factor * std::pow(2.0, (1.0 + i - e) / 2.0) * std::pow(s, 3.0 + i - e) * gamma_fn(n_eff) *
                         hyp2f1_series(1.0, n_eff, (e + 2.0) / 2.0, arg) * std::pow(arg, n_eff);
        }
    }
    for (int j = 0; j < n; j++) {
        double acc = 0.0;
        for (int k = 0; k < n; k++) acc += contrib[k] * t.ortho[j * n + k];
        out[j] = acc;
    }
    return out;
}

}  // namespace rascaline_amd
