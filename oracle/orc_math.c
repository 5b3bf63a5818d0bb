/*
 * FP64 oracle: special functions, real spherical harmonics, eigen solver.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>

#include "oracle.h"

#define ORC_PI 3.141592653589793
#define ORC_E 2.718281828459045

/* Lanczos coefficients, featomic/src/math/gamma.rs:20-32 (Pugh 2004, p.116) */
static const double GAMMA_DK[11] = {
    2.4857408913875355e-5,
    1.0514237858172197,
    -3.4568709722201625,
    4.512277094668948,
    -2.9828522532357664,
    1.056397115771267,
    -1.9542877319164587e-1,
    1.709705434044412e-2,
    -5.719261174043057e-4,
    4.633994733599057e-6,
    -2.7199490848860772e-9,
};
static const double GAMMA_R = 10.900511;
static const double TWO_SQRT_E_OVER_PI = 1.8603827342052657;
static const double LN_2_SQRT_E_OVER_PI = 0.6207822376352452;
static const double LN_PI = 1.1447298858494002;

/* featomic/src/math/gamma.rs:41-63 */
double orc_gamma(double x) {
    if (x < 0.5) {
        double s = GAMMA_DK[0];
        for (int k = 1; k < 11; k++) s += GAMMA_DK[k] / ((double)k - x);
        return ORC_PI / (sin(ORC_PI * x) * s * TWO_SQRT_E_OVER_PI *
                         pow((0.5 - x + GAMMA_R) / ORC_E, 0.5 - x));
    }
    double s = GAMMA_DK[0];
    for (int k = 1; k < 11; k++) s += GAMMA_DK[k] / (x + (double)k - 1.0);
    return s * TWO_SQRT_E_OVER_PI * pow((x - 0.5 + GAMMA_R) / ORC_E, x - 0.5);
}

/* featomic/src/math/gamma.rs:68-92 */
double orc_ln_gamma(double x) {
    if (x < 0.5) {
        double s = GAMMA_DK[0];
        for (int k = 1; k < 11; k++) s += GAMMA_DK[k] / ((double)k - x);
        return LN_PI - log(sin(ORC_PI * x)) - log(s) - LN_2_SQRT_E_OVER_PI -
               (0.5 - x) * log((0.5 - x + GAMMA_R) / ORC_E);
    }
    double s = GAMMA_DK[0];
    for (int k = 1; k < 11; k++) s += GAMMA_DK[k] / (x + (double)k - 1.0);
    return log(s) + LN_2_SQRT_E_OVER_PI + (x - 0.5) * log((x - 0.5 + GAMMA_R) / ORC_E);
}

/* featomic/src/math/hyp1f1.rs:59-104 */
static double hyp1f1_series_track_convergence(double a, double b, double x) {
    double term = 1.0, result = 1.0, abssum = 1.0;
    int converged = 0;
    double n_steps = 0.0;
    for (int ki = 0; ki < 1000; ki++) {
        double k = (double)ki;
        double apk = a + k, bpk = b + k;
        if (bpk != 0.0) {
            term *= apk * x / bpk / (k + 1.0);
        } else if (apk == 0.0) {
            term = 0.0;
        } else {
            return NAN;
        }
        abssum += fabs(term);
        result += term;
        if (fabs(term) <= DBL_EPSILON * fabs(result)) {
            n_steps = k;
            converged = 1;
            break;
        }
    }
    if (!converged) return NAN;
    if (n_steps * DBL_EPSILON * abssum <= 1e-9 * fabs(result)) return result;
    return NAN;
}

/* featomic/src/math/hyp1f1.rs:107-124 */
static double hyp1f1_series(double a, double b, double x) {
    double term = 1.0, result = 1.0;
    int converged = 0;
    for (int ki = 0; ki < 1000; ki++) {
        double k = (double)ki;
        term *= (a + k) * x / (b + k) / (k + 1.0);
        result += term;
        if (fabs(term) <= DBL_EPSILON * fabs(result)) {
            converged = 1;
            break;
        }
    }
    return converged ? result : NAN;
}

/* scipy CHGM as restated in featomic/src/math/hyp1f1.rs:132-255 */
static double chgm(double a, double b, double x) {
    double a0 = a, x0 = x, result = 0.0, y0 = 0.0, y1 = 0.0;
    int nl = 0, la = 0;
    if (x < 0.0) {
        a = b - a;
        a0 = a;
        x = fabs(x);
    }
    if (a >= 2.0) {
        nl = 1;
        la = (int)a;
        a = a - (double)la - 1.0;
    }
    for (int n = 0; n <= nl; n++) {
        if (a0 >= 2.0) a += 1.0;
        if (x <= fabs(b) + 30.0 || a < 0.0) {
            result = 1.0;
            double term = 1.0;
            if (x < 500.0) {
                for (int ji = 1; ji <= 500; ji++) {
                    double j = (double)ji;
                    term *= (a + j - 1.0) / (j * (b + j - 1.0)) * x;
                    result += term;
                    if (result != 0.0 && fabs(term / result) < 1e-15) break;
                }
                if (x0 < 0.0) result *= exp(x0);
            } else {
                for (int ji = 1; ji <= 50; ji++) {
                    double j = (double)ji;
                    term *= (j - a) * (b - a + j - 1.0) / (j * x);
                    result += term;
                    if (result != 0.0 && fabs(term / result) < 1e-15) break;
                }
                result *= pow(x, a - b) * orc_gamma(b) / orc_gamma(a);
                if (x0 > 0.0) result *= exp(x);
            }
        } else {
            double lga = orc_ln_gamma(a), lgb = orc_ln_gamma(b), lgba = orc_ln_gamma(b - a);
            double s1 = 1.0, s2 = 1.0, t1 = 1.0, t2 = 1.0;
            for (int ii = 1; ii <= 30; ii++) {
                double i = (double)ii;
                t1 = -t1 * (a + i - 1.0) * (a - b + i) / (x * i);
                t2 = -t2 * (b - a + i - 1.0) * (a - i) / (x * i);
                s1 += t1;
                s2 += t2;
            }
            double hg1, hg2;
            if (x0 >= 0.0) {
                hg1 = exp(lgb - lgba) * pow(x, -a) * cos(ORC_PI * a) * s1;
                hg2 = exp(lgb - lga + x) * pow(x, a - b) * s2;
            } else {
                hg1 = exp(lgb - lgba + x0) * pow(x, -a) * cos(ORC_PI * a) * s1;
                hg2 = exp(lgb - lga) * pow(x, a - b) * s2;
            }
            result = hg1 + hg2;
        }
        if (n == 0) y0 = result;
        if (n == 1) y1 = result;
    }
    if (a0 >= 2.0) {
        for (int i = 1; i < la; i++) {
            result = ((a * 2.0 - b + x) * y1 + (b - a) * y0) / a;
            y0 = y1;
            y1 = result;
            a += 1.0;
        }
    }
    return result;
}

/* featomic/src/math/hyp1f1.rs:17-53 */
double orc_hyp1f1(double a, double b, double x) {
    if (a == 0.0 || x == 0.0) return 1.0;
    if (a == -1.0) return 1.0 - x / b;
    if (a == b) return exp(x);
    if (a - b == 1.0) return (1.0 + x / b) * exp(x);
    if (a <= 0.0 && a == floor(a)) return hyp1f1_series_track_convergence(a, b, x);
    if (b > 0.0 && (fabs(a) + 1.0) * fabs(x) < 0.9 * b) return hyp1f1_series(a, b, x);
    return chgm(a, b, x);
}

/*
 * Gauss hypergeometric 2F1(a, b; c; x) for 0 <= x < 1 by its defining power
 * series, summed until the terms stop changing the result. Used only for the
 * LODE centre contribution (featomic/src/calculators/lode/radial_integral/gto.rs:170-180,
 * reference implementation math/hyp2f1.rs:24), where a = 1 and x <= ~0.95.
 */
double orc_hyp2f1(double a, double b, double c, double x) {
    if (x == 0.0 || a == 0.0 || b == 0.0) return 1.0;
    double term = 1.0, sum = 1.0;
    for (int k = 0; k < 200000; k++) {
        term *= (a + k) * (b + k) / ((c + k) * (k + 1.0)) * x;
        sum += term;
        if (fabs(term) <= 1e-17 * fabs(sum)) break;
    }
    return sum;
}

/* featomic/src/math/exp.rs:5-33 (scipy specfun E1) */
static double orc_exp1(double x) {
    if (x == 0.0) return INFINITY;
    if (x <= 1.0) {
        double e1 = 1.0, r = 1.0;
        for (int k = 1; k <= 25; k++) {
            double d1 = (double)k + 1.0;
            r = -r * (double)k * x / (d1 * d1);
            e1 += r;
            if (fabs(r) <= fabs(e1) * 1e-15) break;
        }
        return -0.5772156649015328 - log(x) + x * e1;
    }
    int m = (int)(80.0 / x) + 20;
    double t0 = 0.0;
    for (int k = m; k >= 1; k--) t0 = (double)k / ((double)k / (x + t0) + 1.0);
    return exp(-x) * (1.0 / (x + t0));
}

/* featomic/src/math/exp.rs:36-66 */
double orc_expi(double x) {
    if (x == 0.0) return -INFINITY;
    if (x < 0.0) return -orc_exp1(-x);
    if (fabs(x) <= 40.0) {
        double ei = 1.0, r = 1.0;
        for (int k = 1; k <= 100; k++) {
            double d = (double)k + 1.0;
            r = r * (double)k * x / (d * d);
            ei += r;
            if (fabs(r / ei) <= 1e-15) break;
        }
        return 0.5772156649015328 + log(x) + x * ei;
    }
    double ei = 1.0, r = 1.0;
    for (int k = 1; k <= 20; k++) {
        r = r * (double)k / x;
        ei += r;
    }
    return exp(x) / x * ei;
}

/* The reference ships its own erfc (math/erf.rs:44); the C library's erfc is
 * correctly rounded to ~1 ulp and serves as the oracle for it. */
double orc_erfc(double x) { return erfc(x); }

/*
 * Real spherical harmonics and their gradients, restating
 * featomic/src/math/spherical_harmonics.rs:203-424 (arXiv:1410.1748 recursion,
 * Wikipedia real convention with the Condon-Shortley phase cancelled).
 * Index of (l, m) is l*l + l + m. Gradients are r * dY/dr (tangential part).
 */
#define LIDX(l, m) ((l) * ((l) + 1) / 2 + (m))
void orc_spherical_harmonics(int lmax, const double d[3], double* values, double* grads) {
    const double SQRT_1_OVER_2PI = 0.3989422804014327;
    const double SQRT_3 = 1.7320508075688772;
    const double SQRT_3_OVER_2 = 1.224744871391589;
    const double SQRT_2 = 1.4142135623730951;
    int nleg = (lmax + 1) * (lmax + 2) / 2;
    double* A = calloc((size_t)nleg, sizeof(double));
    double* B = calloc((size_t)nleg, sizeof(double));
    double* P = calloc((size_t)nleg, sizeof(double));
    double* DP = calloc((size_t)nleg, sizeof(double));
    double* PT = calloc((size_t)nleg, sizeof(double));

    for (int l = 2; l <= lmax; l++) {
        double ls = (double)(l * l), lm1s = (double)((l - 1) * (l - 1));
        for (int m = 0; m < l - 1; m++) {
            double ms = (double)(m * m);
            A[LIDX(l, m)] = sqrt((4.0 * ls - 1.0) / (ls - ms));
            B[LIDX(l, m)] = -sqrt((lm1s - ms) / (4.0 * lm1s - 1.0));
        }
    }

    double sqrt_xy = hypot(d[0], d[1]);
    double cos_theta = d[2], sin_theta = sqrt_xy;
    double cos_phi = 1.0, sin_phi = 0.0;
    if (sqrt_xy > DBL_EPSILON) {
        cos_phi = d[0] / sqrt_xy;
        sin_phi = d[1] / sqrt_xy;
    }

    /* Legendre recursion, spherical_harmonics.rs:228-251 */
    double value = SQRT_1_OVER_2PI;
    P[LIDX(0, 0)] = value;
    if (lmax > 0) {
        P[LIDX(1, 0)] = cos_theta * SQRT_3 * value;
        value *= -SQRT_3_OVER_2 * sin_theta;
        P[LIDX(1, 1)] = value;
        for (int l = 2; l <= lmax; l++) {
            for (int m = 0; m < l - 1; m++) {
                P[LIDX(l, m)] = A[LIDX(l, m)] * (cos_theta * P[LIDX(l - 1, m)] + B[LIDX(l, m)] * P[LIDX(l - 2, m)]);
            }
            P[LIDX(l, l - 1)] = cos_theta * sqrt(2.0 * l + 1.0) * value;
            value *= -sqrt(1.0 + 0.5 / l) * sin_theta;
            P[LIDX(l, l)] = value;
        }
    }

    if (grads) {
        /* spherical_harmonics.rs:256-298 */
        DP[LIDX(0, 0)] = 0.0;
        for (int l = 1; l <= lmax; l++) {
            for (int m = 0; m <= l; m++) {
                double pm1, pp1;
                if (m == 0) {
                    pm1 = -1.0 / (double)(l * l + l) * P[LIDX(l, 1)];
                } else {
                    pm1 = P[LIDX(l, m - 1)];
                }
                pp1 = (m < l) ? P[LIDX(l, m + 1)] : 0.0;
                DP[LIDX(l, m)] = sqrt((double)((l + m) * (l - m + 1))) * pm1 -
                                 sqrt((double)((l - m) * (l + m + 1))) * pp1;
            }
        }
        for (int l = 0; l <= lmax; l++) {
            for (int m = 0; m <= l; m++) {
                if (sin_theta > 0.1) {
                    PT[LIDX(l, m)] = (double)m / sin_theta * P[LIDX(l, m)];
                } else {
                    PT[LIDX(l, m)] = -0.5 / cos_theta * DP[LIDX(l, m)];
                }
            }
        }
    }

    int L2 = (lmax + 1) * (lmax + 1);
    double* gx = grads;
    double* gy = grads ? grads + L2 : NULL;
    double* gz = grads ? grads + 2 * L2 : NULL;

    for (int l = 0; l <= lmax; l++) values[l * l + l] = P[LIDX(l, 0)] / SQRT_2;
    if (grads) {
        gx[0] = gy[0] = gz[0] = 0.0;
        for (int l = 1; l <= lmax; l++) {
            double lf = sqrt(0.5 * (double)(l * (l + 1))) * P[LIDX(l, 1)];
            gx[l * l + l] = cos_phi * cos_theta * lf;
            gy[l * l + l] = sin_phi * cos_theta * lf;
            gz[l * l + l] = -sin_theta * lf;
        }
    }

    /* sin/cos(m phi) recursion, spherical_harmonics.rs:368-420 */
    double cos_1 = 1.0, sin_1 = 0.0, cos_2 = -cos_phi, sin_2 = sin_phi;
    double minus_two_cos = -2.0 * cos_phi;
    for (int m = 1; m <= lmax; m++) {
        double sin_m = minus_two_cos * sin_1 - sin_2;
        double cos_m = minus_two_cos * cos_1 - cos_2;
        sin_2 = sin_1;
        sin_1 = sin_m;
        cos_2 = cos_1;
        cos_1 = cos_m;
        for (int l = m; l <= lmax; l++) {
            double p = P[LIDX(l, m)];
            values[l * l + l + m] = p * cos_m;
            values[l * l + l - m] = p * sin_m;
        }
        if (grads) {
            for (int l = m; l <= lmax; l++) {
                double dp = DP[LIDX(l, m)];
                double s_dp = sin_m * dp, c_dp = cos_m * dp;
                double pt = PT[LIDX(l, m)];
                gx[l * l + l + m] = sin_phi * pt * sin_m - 0.5 * cos_theta * cos_phi * c_dp;
                gx[l * l + l - m] = -sin_phi * pt * cos_m - 0.5 * cos_theta * cos_phi * s_dp;
                gy[l * l + l + m] = -cos_phi * pt * sin_m - 0.5 * cos_theta * sin_phi * c_dp;
                gy[l * l + l - m] = cos_phi * pt * cos_m - 0.5 * cos_theta * sin_phi * s_dp;
                gz[l * l + l + m] = 0.5 * sin_theta * c_dp;
                gz[l * l + l - m] = 0.5 * sin_theta * s_dp;
            }
        }
    }
    free(A);
    free(B);
    free(P);
    free(DP);
    free(PT);
}

/*
 * Symmetric eigen-decomposition by cyclic Jacobi rotations. The reference uses
 * the JAMA Householder + QL solver (featomic/src/math/eigen.rs:23-270); both
 * give the same spectral decomposition, and its only use here is
 * S^{-1/2} = V diag(lambda^{-1/2}) V^T, which is unique for an SPD matrix.
 * On exit a holds the eigenvectors as columns and w the eigenvalues.
 */
void orc_symmetric_eigen(int n, double* a, double* w) {
    double* v = calloc((size_t)(n * n), sizeof(double));
    for (int i = 0; i < n; i++) v[i * n + i] = 1.0;
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0.0;
        for (int p = 0; p < n; p++)
            for (int q = p + 1; q < n; q++) off += a[p * n + q] * a[p * n + q];
        if (off < 1e-300) break;
        for (int p = 0; p < n; p++) {
            for (int q = p + 1; q < n; q++) {
                double apq = a[p * n + q];
                if (fabs(apq) < 1e-300) continue;
                double app = a[p * n + p], aqq = a[q * n + q];
                double theta = (aqq - app) / (2.0 * apq);
                double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; k++) {
                    double akp = a[k * n + p], akq = a[k * n + q];
                    a[k * n + p] = c * akp - s * akq;
                    a[k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; k++) {
                    double apk = a[p * n + k], aqk = a[q * n + k];
                    a[p * n + k] = c * apk - s * aqk;
                    a[q * n + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; k++) {
                    double vkp = v[k * n + p], vkq = v[k * n + q];
                    v[k * n + p] = c * vkp - s * vkq;
                    v[k * n + q] = s * vkp + c * vkq;
                }
            }
        }
    }
    for (int i = 0; i < n; i++) w[i] = a[i * n + i];
    memcpy(a, v, sizeof(double) * (size_t)(n * n));
    free(v);
}

void orc_free(void* ptr) { free(ptr); }
