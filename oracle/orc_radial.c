/*
 * FP64 oracle: GTO radial integrals (SOAP, LODE) and their Hermite splines.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define ORC_PI 3.141592653589793
#define ORC_SQRT2 1.4142135623730951

struct orc_radial {
    int lode;
    int max_angular;
    int n;
    double stop;
    /* SOAP */
    double sigma2, c;
    /* per n: SOAP 1/(2 sigma_n^2); LODE sigma_n */
    double* gto;
    /* orthonormalisation O (row-major): values <- values . O^T */
    double* ortho;
    /* spline tables per l (a tabulated integral: one table, used for every l) */
    int tabulated;
    int spline;
    int* n_knots;
    double** pos;
    double** val;
    double** der;
};

/* featomic/src/calculators/shared/basis/radial/gto.rs:44-50 */
static void gaussian_widths(double radius, int n, double* out) {
    for (int i = 0; i < n; i++) {
        double s = sqrt((double)i);
        out[i] = radius * (s > 1.0 ? s : 1.0) / (double)n;
    }
}

/* featomic/src/calculators/shared/basis/radial/gto.rs:17-41,53-77.
 * Deviation from the reference: S and S^{-1/2} are evaluated in long double.
 * The overlap is ill-conditioned (condition number ~3e4 for 8 functions, ~4e7
 * for 12), so an FP64 eigen-solver (the reference's JAMA Householder + QL) leaves
 * noise of order cond(S) * 1e-16 in the orthonormalisation; extended precision
 * removes it, so that this oracle pins the mathematical value. */
static void jacobi_ld(int n, long double* a, long double* w) {
    long double* v = calloc((size_t)(n * n), sizeof(long double));
    for (int i = 0; i < n; i++) v[i * n + i] = 1.0L;
    for (int sweep = 0; sweep < 100; sweep++) {
        long double off = 0.0L;
        for (int p = 0; p < n; p++)
            for (int q = p + 1; q < n; q++) off += a[p * n + q] * a[p * n + q];
        if (off < 1e-4000L) break;
        for (int p = 0; p < n; p++) {
            for (int q = p + 1; q < n; q++) {
                long double apq = a[p * n + q];
                if (fabsl(apq) < 1e-4000L) continue;
                long double theta = (a[q * n + q] - a[p * n + p]) / (2.0L * apq);
                long double t = (theta >= 0 ? 1.0L : -1.0L) / (fabsl(theta) + sqrtl(theta * theta + 1.0L));
                long double c = 1.0L / sqrtl(t * t + 1.0L), s = t * c;
                for (int k = 0; k < n; k++) {
                    long double akp = a[k * n + p], akq = a[k * n + q];
                    a[k * n + p] = c * akp - s * akq;
                    a[k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; k++) {
                    long double apk = a[p * n + k], aqk = a[q * n + k];
                    a[p * n + k] = c * apk - s * aqk;
                    a[q * n + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; k++) {
                    long double vkp = v[k * n + p], vkq = v[k * n + q];
                    v[k * n + p] = c * vkp - s * vkq;
                    v[k * n + q] = s * vkp + c * vkq;
                }
            }
        }
    }
    for (int i = 0; i < n; i++) w[i] = a[i * n + i];
    memcpy(a, v, sizeof(long double) * (size_t)(n * n));
    free(v);
}

static int orthonormalization(double radius, int n, double* out) {
    long double* sigma = malloc(sizeof(long double) * n);
    for (int i = 0; i < n; i++) {
        long double s = sqrtl((long double)i);
        sigma[i] = (long double)radius * (s > 1.0L ? s : 1.0L) / (long double)n;
    }
    long double* S = malloc(sizeof(long double) * n * n);
    for (int n1 = 0; n1 < n; n1++) {
        long double s1 = sigma[n1], s1sq = s1 * s1;
        for (int n2 = n1; n2 < n; n2++) {
            long double s2 = sigma[n2], s2sq = s2 * s2;
            long double e = 0.5L * (3.0L + n1 + n2);
            long double v = powl(0.5L / s1sq + 0.5L / s2sq, -e) / (powl(s1, n1) * powl(s2, n2)) * tgammal(e) /
                            (powl(s1 * s2, 1.5L) * sqrtl(tgammal(n1 + 1.5L) * tgammal(n2 + 1.5L)));
            S[n2 * n + n1] = v;
            S[n1 * n + n2] = v;
        }
    }
    long double* w = malloc(sizeof(long double) * n);
    jacobi_ld(n, S, w);
    int status = 0;
    for (int i = 0; i < n; i++) {
        if (w[i] <= 2.220446049250313e-16L) status = -1;
        w[i] = 1.0L / sqrtl(w[i]);
    }
    for (int a = 0; a < n; a++) {
        for (int b = 0; b < n; b++) {
            long double norm_b = sqrtl(2.0L / (powl(sigma[b], 2 * b + 3) * tgammal(b + 1.5L)));
            long double acc = 0.0L;
            for (int k = 0; k < n; k++) acc += S[a * n + k] * w[k] * S[b * n + k];
            out[a * n + b] = (double)(acc * norm_b);
        }
    }
    free(sigma);
    free(S);
    free(w);
    return status;
}

static void apply_ortho(const orc_radial* r, double* v) {
    double tmp[64];
    for (int j = 0; j < r->n; j++) {
        double acc = 0.0;
        for (int k = 0; k < r->n; k++) acc += v[k] * r->ortho[j * r->n + k];
        tmp[j] = acc;
    }
    memcpy(v, tmp, sizeof(double) * r->n);
}

/* featomic/src/calculators/soap/radial_integral/gto.rs:115-205 */
static void soap_gto_exact(const orc_radial* r, int l, double distance, double* values, double* grads) {
    double global_factor = pow(ORC_PI / r->sigma2, 0.75);
    double c = r->c;
    double c_rij = c * distance;
    double c_rij_l = pow(c_rij, l);
    double exp_c_rij = exp(-distance * c_rij);
    double factor = global_factor * exp_c_rij * c_rij_l;
    for (int n = 0; n < r->n; n++) {
        double gto_c = r->gto[n];
        double z = c_rij * c_rij / (c + gto_c);
        double a = 0.5 * (double)(n + l + 3), b = (double)l + 1.5;
        double ratio = orc_gamma(a) / orc_gamma(b);
        values[n] = ratio * orc_hyp1f1(a, b, z);
        double g = ratio * (a / b * orc_hyp1f1(a + 1.0, b + 1.0, z));
        double c_dn = pow(c + gto_c, -(0.5 * (double)(n + l) + 1.5));
        values[n] *= c_dn * factor;
        if (grads) {
            g *= c_dn * factor * 2.0 * z / distance;
            g += values[n] * ((double)l / distance - 2.0 * c_rij);
            grads[n] = g;
        }
    }
    if (distance == 0.0 && grads) {
        for (int n = 0; n < r->n; n++) {
            if (l == 1) {
                double a = 0.5 * (double)(n + l) + 1.5, b = 2.5;
                double c_dn = pow(c + r->gto[n], -a);
                grads[n] = orc_gamma(a) / orc_gamma(b) * (global_factor * c * c_dn);
            } else {
                grads[n] = 0.0;
            }
        }
    }
    apply_ortho(r, values);
    if (grads) apply_ortho(r, grads);
}

/* featomic/src/calculators/lode/radial_integral/gto.rs:65-134 */
static void lode_gto_exact(const orc_radial* r, int l, double k, double* values, double* grads) {
    double global_factor = sqrt(ORC_PI) / ORC_SQRT2;
    for (int n = 0; n < r->n; n++) {
        double sn = r->gto[n];
        double ks = k * sn / ORC_SQRT2;
        double factor = global_factor * pow(sn, n + 3) * pow(ORC_SQRT2, n) * pow(ks, l);
        double kss = -k * sn * sn;
        double z = 0.5 * k * kss;
        double a = 0.5 * (double)(n + l + 3), b = (double)l + 1.5;
        double ratio = orc_gamma(a) / orc_gamma(b);
        values[n] = ratio * orc_hyp1f1(a, b, z);
        double g = ratio * (a / b * orc_hyp1f1(a + 1.0, b + 1.0, z));
        values[n] *= factor;
        if (grads) {
            g *= kss * factor;
            g += (double)l / k * values[n];
            grads[n] = g;
        }
    }
    if (k == 0.0 && grads) {
        for (int n = 0; n < r->n; n++) {
            grads[n] = 0.0;
            if (l == 1) {
                double sn = r->gto[n];
                double a = 0.5 * (double)(n + l) + 1.5, b = 2.5;
                double f = global_factor * pow(sn, n + l + 3) * pow(ORC_SQRT2, n - l);
                grads[n] = orc_gamma(a) / orc_gamma(b) * f;
            }
        }
    }
    apply_ortho(r, values);
    if (grads) apply_ortho(r, grads);
}

static void exact(const orc_radial* r, int l, double x, double* values, double* grads) {
    if (r->lode) {
        lode_gto_exact(r, l, x, values, grads);
    } else {
        soap_gto_exact(r, l, x, values, grads);
    }
}

/* featomic/src/math/splines.rs:225-289 */
static void spline_eval(int n_knots, const double* pos, const double* val, const double* der, int n,
                        double x, double* values, double* grads) {
    /* binary search for the last knot <= x */
    int lo = 0, hi = n_knots - 1;
    int k;
    if (x == pos[hi]) {
        k = hi;
    } else {
        while (hi - lo > 1) {
            int mid = (lo + hi) / 2;
            if (pos[mid] <= x) lo = mid; else hi = mid;
        }
        k = lo;
        if (pos[hi] == x) k = hi;
    }
    if (k == n_knots - 1) k -= 1;
    double xk = pos[k], xk1 = pos[k + 1];
    double delta = xk1 - xk;
    double t = (x - xk) / delta;
    double t2 = t * t, t3 = t2 * t;
    double h00 = 2.0 * t3 - 3.0 * t2 + 1.0;
    double h10 = t3 - 2.0 * t2 + t;
    double h01 = -2.0 * t3 + 3.0 * t2;
    double h11 = t3 - t2;
    const double* pk = val + (size_t)k * n;
    const double* pk1 = val + (size_t)(k + 1) * n;
    const double* mk = der + (size_t)k * n;
    const double* mk1 = der + (size_t)(k + 1) * n;
    for (int i = 0; i < n; i++) {
        values[i] = h00 * pk[i] + h10 * delta * mk[i] + h01 * pk1[i] + h11 * delta * mk1[i];
    }
    if (grads) {
        double d00 = 6.0 * (t2 - t);
        double d10 = 3.0 * t2 - 4.0 * t + 1.0;
        double d01 = -d00;
        double d11 = 3.0 * t2 - 2.0 * t;
        double dx = 1.0 / delta;
        for (int i = 0; i < n; i++) {
            grads[i] = d00 * pk[i] * dx + d10 * mk[i] + d01 * pk1[i] * dx + d11 * mk1[i];
        }
    }
}

/* featomic/src/math/splines.rs:79-194 */
static int build_spline(orc_radial* r, int l, double accuracy) {
    int n = r->n;
    int cap = 16;
    int count = 0;
    double* pos = malloc(sizeof(double) * cap);
    double* val = malloc(sizeof(double) * cap * n);
    double* der = malloc(sizeof(double) * cap * n);
    double step = r->stop / 10.0;
    for (int k = 0; k <= 10; k++) {
        double x = (k < 10) ? (double)k * step : r->stop;
        pos[count] = x;
        exact(r, l, x, val + (size_t)count * n, der + (size_t)count * n);
        count++;
    }
    double* interp = malloc(sizeof(double) * n);
    for (;;) {
        int n_new = count - 1;
        double* npos = malloc(sizeof(double) * n_new);
        double* nval = malloc(sizeof(double) * n_new * n);
        double* nder = malloc(sizeof(double) * n_new * n);
        double mean_abs = 0.0, mean_rel = 0.0;
        long err_count = 0;
        for (int k = 0; k < n_new; k++) {
            double x = (pos[k] + pos[k + 1]) / 2.0;
            npos[k] = x;
            exact(r, l, x, nval + (size_t)k * n, nder + (size_t)k * n);
            spline_eval(count, pos, val, der, n, x, interp, NULL);
            for (int i = 0; i < n; i++) {
                double v = nval[(size_t)k * n + i];
                double ae = fabs(interp[i] - v);
                mean_abs += ae;
                mean_rel += fabs((interp[i] - v) / v);
                err_count++;
            }
        }
        mean_abs /= (double)err_count;
        mean_rel /= (double)err_count;
        if (mean_abs < accuracy || mean_rel < accuracy) {
            free(npos);
            free(nval);
            free(nder);
            break;
        }
        if (count + n_new > 10000) {
            fprintf(stderr, "oracle spline: failed to reach accuracy %g\n", accuracy);
            free(npos);
            free(nval);
            free(nder);
            free(interp);
            return -1;
        }
        /* interleave old and new knots */
        int total = count + n_new;
        double* p2 = malloc(sizeof(double) * total);
        double* v2 = malloc(sizeof(double) * total * n);
        double* d2 = malloc(sizeof(double) * total * n);
        for (int k = 0; k < count; k++) {
            p2[2 * k] = pos[k];
            memcpy(v2 + (size_t)(2 * k) * n, val + (size_t)k * n, sizeof(double) * n);
            memcpy(d2 + (size_t)(2 * k) * n, der + (size_t)k * n, sizeof(double) * n);
            if (k < n_new) {
                p2[2 * k + 1] = npos[k];
                memcpy(v2 + (size_t)(2 * k + 1) * n, nval + (size_t)k * n, sizeof(double) * n);
                memcpy(d2 + (size_t)(2 * k + 1) * n, nder + (size_t)k * n, sizeof(double) * n);
            }
        }
        free(pos);
        free(val);
        free(der);
        free(npos);
        free(nval);
        free(nder);
        pos = p2;
        val = v2;
        der = d2;
        count = total;
    }
    free(interp);
    r->n_knots[l] = count;
    r->pos[l] = pos;
    r->val[l] = val;
    r->der[l] = der;
    return 0;
}

static orc_radial* radial_alloc(int lode, double radius, int max_angular, int n, double stop,
                                double spline_accuracy) {
    if (n > 64) return NULL;
    orc_radial* r = calloc(1, sizeof(orc_radial));
    r->lode = lode;
    r->max_angular = max_angular;
    r->n = n;
    r->stop = stop;
    r->gto = malloc(sizeof(double) * n);
    r->ortho = malloc(sizeof(double) * n * n);
    if (orthonormalization(radius, n, r->ortho) != 0) {
        orc_radial_free(r);
        return NULL;
    }
    return r;
}

static int radial_finish(orc_radial* r, double spline_accuracy) {
    if (spline_accuracy > 0.0) {
        int L = r->max_angular + 1;
        r->spline = 1;
        r->n_knots = calloc(L, sizeof(int));
        r->pos = calloc(L, sizeof(double*));
        r->val = calloc(L, sizeof(double*));
        r->der = calloc(L, sizeof(double*));
        for (int l = 0; l < L; l++) {
            if (build_spline(r, l, spline_accuracy) != 0) return -1;
        }
    }
    return 0;
}

orc_radial* orc_soap_radial_new(double cutoff, double gaussian_width, int max_angular, int n_radial,
                                double spline_accuracy) {
    orc_radial* r = radial_alloc(0, cutoff, max_angular, n_radial, cutoff, spline_accuracy);
    if (!r) return NULL;
    double* sigma = malloc(sizeof(double) * n_radial);
    gaussian_widths(cutoff, n_radial, sigma);
    for (int i = 0; i < n_radial; i++) r->gto[i] = 1.0 / (2.0 * sigma[i] * sigma[i]);
    free(sigma);
    r->sigma2 = gaussian_width * gaussian_width;
    r->c = 1.0 / (2.0 * r->sigma2);
    if (radial_finish(r, spline_accuracy) != 0) {
        orc_radial_free(r);
        return NULL;
    }
    return r;
}

orc_radial* orc_lode_radial_new(double radius, int max_angular, int n_radial, double k_cutoff,
                                double spline_accuracy) {
    orc_radial* r = radial_alloc(1, radius, max_angular, n_radial, k_cutoff, spline_accuracy);
    if (!r) return NULL;
    gaussian_widths(radius, n_radial, r->gto);
    if (radial_finish(r, spline_accuracy) != 0) {
        orc_radial_free(r);
        return NULL;
    }
    return r;
}

orc_radial* orc_radial_tabulated(int n, int n_points, const double* positions, const double* values,
                                 const double* derivatives) {
    if (n < 1 || n_points < 2) return NULL;
    orc_radial* r = calloc(1, sizeof(orc_radial));
    r->n = n;
    r->stop = positions[n_points - 1];
    r->tabulated = 1;
    r->spline = 1;
    r->n_knots = calloc(1, sizeof(int));
    r->pos = calloc(1, sizeof(double*));
    r->val = calloc(1, sizeof(double*));
    r->der = calloc(1, sizeof(double*));
    r->n_knots[0] = n_points;
    r->pos[0] = malloc(sizeof(double) * n_points);
    r->val[0] = malloc(sizeof(double) * n_points * n);
    r->der[0] = malloc(sizeof(double) * n_points * n);
    memcpy(r->pos[0], positions, sizeof(double) * n_points);
    memcpy(r->val[0], values, sizeof(double) * n_points * n);
    memcpy(r->der[0], derivatives, sizeof(double) * n_points * n);
    return r;
}

void orc_radial_compute(const orc_radial* r, int l, double x, double* values, double* grads) {
    if (r->tabulated) l = 0;
    if (r->spline) {
        spline_eval(r->n_knots[l], r->pos[l], r->val[l], r->der[l], r->n, x, values, grads);
    } else {
        exact(r, l, x, values, grads);
    }
}

int orc_radial_n_knots(const orc_radial* r, int l) { return r->spline ? r->n_knots[l] : 0; }

void orc_radial_knots(const orc_radial* r, int l, double* positions, double* values, double* derivs) {
    int k = r->n_knots[l];
    memcpy(positions, r->pos[l], sizeof(double) * k);
    memcpy(values, r->val[l], sizeof(double) * k * r->n);
    memcpy(derivs, r->der[l], sizeof(double) * k * r->n);
}

void orc_radial_orthonormalization(const orc_radial* r, double* out) {
    memcpy(out, r->ortho, sizeof(double) * r->n * r->n);
}

/* gto widths, exposed for the LODE centre contribution */
const double* orc_radial_gto(const orc_radial* r) { return r->gto; }
int orc_radial_size(const orc_radial* r) { return r->n; }

void orc_radial_free(orc_radial* r) {
    if (!r) return;
    if (r->spline) {
        for (int l = 0; l <= r->max_angular; l++) {
            free(r->pos[l]);
            free(r->val[l]);
            free(r->der[l]);
        }
        free(r->n_knots);
        free(r->pos);
        free(r->val);
        free(r->der);
    }
    free(r->gto);
    free(r->ortho);
    free(r);
}
