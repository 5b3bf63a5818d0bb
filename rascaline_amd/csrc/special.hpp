// Special functions shared by host table builders and device kernels (FP64).
//
// Used for the exact GTO radial integrals (featomic/src/calculators/soap/radial_integral/gto.rs:115-205,
// featomic/src/calculators/lode/radial_integral/gto.rs:65-185), the LODE density in
// Fourier space (featomic/src/calculators/lode/spherical_expansion.rs:280-342) and the
// spline knot values. The gamma functions are the Lanczos approximation the reference
// uses, so that the integrals agree with the reference's to rounding: at 12 radial
// functions the orthonormalisation amplifies last-bit differences in Gamma(a) / Gamma(b)
// a million-fold.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>

#define RA_HD __host__ __device__ inline

namespace rascaline_amd {

constexpr double RA_PI = 3.141592653589793;

// Lanczos approximation with Pugh's coefficients (G. R. Pugh, "An Analysis of the
// Lanczos Gamma Approximation", 2004, p. 116: g = 10.900511, 11 terms), as in
// featomic/src/math/gamma.rs:20-92. `shift` is x - 1 for x >= 0.5 and -x below
// (reflection); the sum is accumulated in increasing k like the reference.
RA_HD double lanczos_sum(double x, bool reflected) {
    const double d[11] = {2.4857408913875355e-5,  1.0514237858172197,     -3.4568709722201625,
                          4.512277094668948,      -2.9828522532357664,    1.056397115771267,
                          -1.9542877319164587e-1, 1.709705434044412e-2,   -5.719261174043057e-4,
                          4.633994733599057e-6,   -2.7199490848860772e-9};
    double s = d[0];
    for (int k = 1; k < 11; k++) s += reflected ? d[k] / ((double)k - x) : d[k] / (x + (double)k - 1.0);
    return s;
}

constexpr double LANCZOS_G = 10.900511;
constexpr double TWO_SQRT_E_OVER_PI = 1.8603827342052657;
constexpr double LN_2_SQRT_E_OVER_PI = 0.6207822376352452;
constexpr double LN_PI = 1.1447298858494002;
constexpr double EULER_E = 2.718281828459045;

RA_HD double gamma_fn(double x) {
    if (x < 0.5) {
        const double s = lanczos_sum(x, true);
        return RA_PI / (sin(RA_PI * x) * s * TWO_SQRT_E_OVER_PI * pow((0.5 - x + LANCZOS_G) / EULER_E, 0.5 - x));
    }
    const double s = lanczos_sum(x, false);
    return s * TWO_SQRT_E_OVER_PI * pow((x - 0.5 + LANCZOS_G) / EULER_E, x - 0.5);
}

RA_HD double ln_gamma_fn(double x) {
    if (x < 0.5) {
        const double s = lanczos_sum(x, true);
        return LN_PI - log(sin(RA_PI * x)) - log(s) - LN_2_SQRT_E_OVER_PI -
               (0.5 - x) * log((0.5 - x + LANCZOS_G) / EULER_E);
    }
    const double s = lanczos_sum(x, false);
    return log(s) + LN_2_SQRT_E_OVER_PI + (x - 0.5) * log((x - 0.5 + LANCZOS_G) / EULER_E);
}

// 1F1(a; b; x) by its power series; valid when |x| is moderate compared to b.
RA_HD double hyp1f1_power_series(double a, double b, double x, bool* ok) {
    double term = 1.0, sum = 1.0;
    for (int k = 0; k < 1000; k++) {
        term *= (a + k) * x / (b + k) / (k + 1.0);
        sum += term;
        if (fabs(term) <= DBL_EPSILON * fabs(sum)) {
            *ok = true;
            return sum;
        }
    }
    *ok = false;
    return NAN;
}

// Confluent hypergeometric 1F1(a; b; x) for b > 0 and the a >= 0.5 used by the
// GTO integrals. Small arguments use the power series; larger ones use the
// Kummer transformation for x < 0 (DLMF 13.2.39), the asymptotic expansion
// (DLMF 13.7.2) for large |x|, and the forward recurrence in a (DLMF 13.3.1)
// to bring a into [0, 1] first, as in Zhang & Jin's CHGM routine that scipy and
// featomic (featomic/src/math/hyp1f1.rs:17-255) both use.
RA_HD double hyp1f1(double a, double b, double x) {
    if (a == 0.0 || x == 0.0) return 1.0;
    if (a == -1.0) return 1.0 - x / b;
    if (a == b) return exp(x);
    if (a - b == 1.0) return (1.0 + x / b) * exp(x);
    if ((fabs(a) + 1.0) * fabs(x) < 0.9 * b) {
        bool ok;
        return hyp1f1_power_series(a, b, x, &ok);
    }
    const double a_in = a, x_in = x;
    double aa = a, xx = x;
    if (xx < 0.0) {
        aa = b - aa;
        xx = -xx;
    }
    const double a_start = aa;
    int steps = 0;
    bool recur = false;
    if (aa >= 2.0) {
        recur = true;
        steps = (int)aa;
        aa = aa - (double)steps - 1.0;
    }
    double y0 = 0.0, y1 = 0.0, result = 0.0;
    const int passes = recur ? 2 : 1;
    for (int pass = 0; pass < passes; pass++) {
        if (recur) aa += 1.0;
        if (xx <= fabs(b) + 30.0 || aa < 0.0) {
            double sum = 1.0, term = 1.0;
            if (xx < 500.0) {
                for (int j = 1; j <= 500; j++) {
                    term *= (aa + j - 1.0) / (j * (b + j - 1.0)) * xx;
                    sum += term;
                    if (sum != 0.0 && fabs(term / sum) < 1e-15) break;
                }
                if (x_in < 0.0) sum *= exp(x_in);
            } else {
                for (int j = 1; j <= 50; j++) {
                    term *= (j - aa) * (b - aa + j - 1.0) / (j * xx);
                    sum += term;
                    if (sum != 0.0 && fabs(term / sum) < 1e-15) break;
                }
                sum *= pow(xx, aa - b) * gamma_fn(b) / gamma_fn(aa);
                if (x_in > 0.0) sum *= exp(xx);
            }
            result = sum;
        } else {
            const double lga = ln_gamma_fn(aa), lgb = ln_gamma_fn(b), lgba = ln_gamma_fn(b - aa);
            double s1 = 1.0, s2 = 1.0, t1 = 1.0, t2 = 1.0;
            for (int i = 1; i <= 30; i++) {
                t1 = -t1 * (aa + i - 1.0) * (aa - b + i) / (xx * i);
                t2 = -t2 * (b - aa + i - 1.0) * (aa - i) / (xx * i);
                s1 += t1;
                s2 += t2;
            }
            double h1, h2;
            if (x_in >= 0.0) {
                h1 = exp(lgb - lgba) * pow(xx, -aa) * cos(RA_PI * aa) * s1;
                h2 = exp(lgb - lga + xx) * pow(xx, aa - b) * s2;
            } else {
                h1 = exp(lgb - lgba + x_in) * pow(xx, -aa) * cos(RA_PI * aa) * s1;
                h2 = exp(lgb - lga) * pow(xx, aa - b) * s2;
            }
            result = h1 + h2;
        }
        if (pass == 0) y0 = result; else y1 = result;
    }
    if (recur) {
        for (int i = 1; i < steps; i++) {
            result = ((aa * 2.0 - b + xx) * y1 + (b - aa) * y0) / aa;
            y0 = y1;
            y1 = result;
            aa += 1.0;
        }
    }
    (void)a_start;
    (void)a_in;
    return result;
}

// 2F1(a, b; c; x) for 0 <= x < 1 by its power series.
RA_HD double hyp2f1_series(double a, double b, double c, double x) {
    if (x == 0.0 || a == 0.0 || b == 0.0) return 1.0;
    double term = 1.0, sum = 1.0;
    for (int k = 0; k < 200000; k++) {
        term *= (a + k) * (b + k) / ((c + k) * (k + 1.0)) * x;
        sum += term;
        if (fabs(term) <= 1e-17 * fabs(sum)) break;
    }
    return sum;
}

// Exponential integral E1(x), x > 0: series below 1, continued fraction above
// (Abramowitz & Stegun 5.1.11 and 5.1.22).
RA_HD double exp1(double x) {
    if (x == 0.0) return INFINITY;
    if (x <= 1.0) {
        double e1 = 1.0, r = 1.0;
        for (int k = 1; k <= 25; k++) {
            double d = k + 1.0;
            r = -r * k * x / (d * d);
            e1 += r;
            if (fabs(r) <= fabs(e1) * 1e-15) break;
        }
        return -0.5772156649015328 - log(x) + x * e1;
    }
    int m = (int)(80.0 / x) + 20;
    double t0 = 0.0;
    for (int k = m; k >= 1; k--) t0 = k / (k / (x + t0) + 1.0);
    return exp(-x) / (x + t0);
}

// Exponential integral Ei(x).
RA_HD double expi(double x) {
    if (x == 0.0) return -INFINITY;
    if (x < 0.0) return -exp1(-x);
    if (x <= 40.0) {
        double ei = 1.0, r = 1.0;
        for (int k = 1; k <= 100; k++) {
            double d = k + 1.0;
            r = r * k * x / (d * d);
            ei += r;
            if (fabs(r / ei) <= 1e-15) break;
        }
        return 0.5772156649015328 + log(x) + x * ei;
    }
    double ei = 1.0, r = 1.0;
    for (int k = 1; k <= 20; k++) {
        r = r * k / x;
        ei += r;
    }
    return exp(x) / x * ei;
}

// Radial basis widths sigma_n = radius * max(sqrt(n), 1) / n_max
// (featomic/src/calculators/shared/basis/radial/gto.rs:44-50).
RA_HD double gto_width(double radius, int n, int n_max) {
    double s = sqrt((double)n);
    return radius * (s > 1.0 ? s : 1.0) / (double)n_max;
}

// Raw (not orthonormalised) SOAP GTO x Gaussian radial integral and its
// derivative for one (l, n): featomic/src/calculators/soap/radial_integral/gto.rs:140-199.
RA_HD void soap_gto_raw(int l, int n, double r, double sigma2_atomic, double gto_constant,
                        double* value, double* grad) {
    const double global_factor = pow(RA_PI / sigma2_atomic, 0.75);
    const double c = 1.0 / (2.0 * sigma2_atomic);
    const double c_r = c * r;
    const double factor = global_factor * exp(-r * c_r) * pow(c_r, (double)l);
    const double z = c_r * c_r / (c + gto_constant);
    const double a = 0.5 * (n + l + 3), b = l + 1.5;
    const double ratio = gamma_fn(a) / gamma_fn(b);
    double v = ratio * hyp1f1(a, b, z);
    double g = ratio * (a / b * hyp1f1(a + 1.0, b + 1.0, z));
    const double c_dn = pow(c + gto_constant, -(0.5 * (n + l) + 1.5));
    v *= c_dn * factor;
    if (r == 0.0) {
        g = 0.0;
        if (l == 1) {
            const double aa = 0.5 * (n + l) + 1.5;
            g = gamma_fn(aa) / gamma_fn(2.5) * (global_factor * c * pow(c + gto_constant, -aa));
        }
    } else {
        g *= c_dn * factor * 2.0 * z / r;
        g += v * ((double)l / r - 2.0 * c_r);
    }
    *value = v;
    *grad = g;
}

// Raw LODE GTO radial integral in k-space and its derivative for one (l, n):
// featomic/src/calculators/lode/radial_integral/gto.rs:85-127.
RA_HD void lode_gto_raw(int l, int n, double k, double sigma_n, double* value, double* grad) {
    const double sqrt2 = 1.4142135623730951;
    const double global_factor = sqrt(RA_PI) / sqrt2;
    const double ks = k * sigma_n / sqrt2;
    const double factor = global_factor * pow(sigma_n, (double)(n + 3)) * pow(sqrt2, (double)n) * pow(ks, (double)l);
    const double kss = -k * sigma_n * sigma_n;
    const double z = 0.5 * k * kss;
    const double a = 0.5 * (n + l + 3), b = l + 1.5;
    const double ratio = gamma_fn(a) / gamma_fn(b);
    double v = ratio * hyp1f1(a, b, z) * factor;
    double g = 0.0;
    if (k == 0.0) {
        if (l == 1) {
            const double aa = 0.5 * (n + l) + 1.5;
            g = gamma_fn(aa) / gamma_fn(2.5) * (global_factor * pow(sigma_n, (double)(n + l + 3)) * pow(sqrt2, (double)(n - l)));
        }
    } else {
        g = ratio * (a / b * hyp1f1(a + 1.0, b + 1.0, z)) * kss * factor + (double)l / k * v;
    }
    *value = v;
    *grad = g;
}

}  // namespace rascaline_amd
