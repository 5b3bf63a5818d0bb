/*
 * FP64 oracle: LODE spherical expansion (k-space).
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define ORC_PI 3.141592653589793

const double* orc_radial_gto(const orc_radial* r);
int orc_radial_size(const orc_radial* r);

struct orc_lode {
    orc_lode_hypers h;
    double k_cutoff;
    orc_radial* radial;
    double center[64];
};

/* featomic/src/calculators/lode/radial_integral/gto.rs:136-185 */
static void center_contribution(const orc_lode* lode, double* out) {
    int n = lode->h.n_radial;
    const double* widths = orc_radial_gto(lode->radial);
    double smearing = lode->h.smearing, p = (double)lode->h.exponent;
    double contrib[64];
    for (int i = 0; i < n; i++) {
        double n_eff = 0.5 * (3.0 + i);
        if (lode->h.exponent == 0) {
            double factor = pow(ORC_PI, -0.25) / pow(smearing * smearing, 0.75);
            double alpha = 0.5 * (1.0 / (smearing * smearing) + 1.0 / (widths[i] * widths[i]));
            contrib[i] = factor * orc_gamma(n_eff) / pow(alpha, n_eff);
        } else {
            double factor = 2.0 * sqrt(4.0 * ORC_PI) / orc_gamma(p / 2.0) / p;
            double s = smearing / widths[i];
            double hyparg = 1.0 / (1.0 + s * s);
            contrib[i] = factor * pow(2.0, (1.0 + i - p) / 2.0) * pow(smearing, 3 + i - lode->h.exponent) *
                         orc_gamma(n_eff) * orc_hyp2f1(1.0, n_eff, (p + 2.0) / 2.0, hyparg) * pow(hyparg, n_eff);
        }
    }
    /* contrib . O^T */
    double* O = malloc(sizeof(double) * n * n);
    orc_radial_orthonormalization(lode->radial, O);
    for (int j = 0; j < n; j++) {
        double acc = 0.0;
        for (int k = 0; k < n; k++) acc += contrib[k] * O[j * n + k];
        out[j] = acc;
    }
    free(O);
}

orc_lode* orc_lode_new(const orc_lode_hypers* hypers) {
    orc_lode* lode = calloc(1, sizeof(orc_lode));
    lode->h = *hypers;
    /* lode/spherical_expansion.rs:53-60 */
    lode->k_cutoff = hypers->k_cutoff > 0.0 ? hypers->k_cutoff : 1.2 * ORC_PI / hypers->smearing;
    lode->radial = orc_lode_radial_new(hypers->radius, hypers->max_angular, hypers->n_radial, lode->k_cutoff,
                                       hypers->spline_accuracy);
    if (!lode->radial) {
        free(lode);
        return NULL;
    }
    center_contribution(lode, lode->center);
    return lode;
}

void orc_lode_free(orc_lode* lode) {
    if (!lode) return;
    orc_radial_free(lode->radial);
    free(lode);
}

static double det3(const double m[9]) {
    return m[0] * (m[4] * m[8] - m[7] * m[5]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
           m[2] * (m[3] * m[7] - m[4] * m[6]);
}

static void inverse3(const double m[9], double r[9]) {
    double inv = 1.0 / det3(m);
    r[0] = (m[4] * m[8] - m[7] * m[5]) * inv;
    r[1] = (m[2] * m[7] - m[1] * m[8]) * inv;
    r[2] = (m[1] * m[5] - m[2] * m[4]) * inv;
    r[3] = (m[5] * m[6] - m[3] * m[8]) * inv;
    r[4] = (m[0] * m[8] - m[2] * m[6]) * inv;
    r[5] = (m[3] * m[2] - m[0] * m[5]) * inv;
    r[6] = (m[3] * m[7] - m[6] * m[4]) * inv;
    r[7] = (m[6] * m[1] - m[0] * m[7]) * inv;
    r[8] = (m[0] * m[4] - m[3] * m[1]) * inv;
}

/* featomic/src/math/k_vectors.rs:20-83; out: [n][4] = direction xyz, norm */
int orc_k_vectors(const double* cell, double k_cutoff, double* out) {
    double t[9] = {cell[0], cell[3], cell[6], cell[1], cell[4], cell[7], cell[2], cell[5], cell[8]};
    double inv[9], rc[9];
    inverse3(t, inv);
    for (int k = 0; k < 9; k++) rc[k] = 2.0 * ORC_PI * inv[k];
    double cutoff2 = k_cutoff * k_cutoff;
    /* m = rc . rc^T */
    double m[9];
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) m[a * 3 + b] = rc[a * 3 + 0] * rc[b * 3 + 0] + rc[a * 3 + 1] * rc[b * 3 + 1] + rc[a * 3 + 2] * rc[b * 3 + 2];
    double kv = det3(rc);
    int n1max = (int)(sqrt(m[4] * m[8] - m[5] * m[5]) / kv * k_cutoff);
    int n2max = (int)(sqrt(m[8] * m[0] - m[6] * m[6]) / kv * k_cutoff);
    int n3max = (int)(sqrt(m[0] * m[4] - m[1] * m[1]) / kv * k_cutoff);
    const double* b1 = rc;
    const double* b2 = rc + 3;
    const double* b3 = rc + 6;
    int count = 0;
#define ORC_PUSH(kx, ky, kz)                                       \
    do {                                                           \
        double n2_ = kx * kx + ky * ky + kz * kz;                  \
        if (n2_ < cutoff2) {                                       \
            double nrm = sqrt(n2_);                                \
            if (out) {                                             \
                out[4 * count + 0] = kx / nrm;                     \
                out[4 * count + 1] = ky / nrm;                     \
                out[4 * count + 2] = kz / nrm;                     \
                out[4 * count + 3] = nrm;                          \
            }                                                      \
            count++;                                               \
        }                                                          \
    } while (0)
    for (int n3 = 1; n3 <= n3max; n3++) {
        double kx = (double)n3 * b3[0], ky = (double)n3 * b3[1], kz = (double)n3 * b3[2];
        ORC_PUSH(kx, ky, kz);
    }
    for (int n2 = 1; n2 <= n2max; n2++) {
        for (int n3 = -n3max; n3 <= n3max; n3++) {
            double kx = (double)n2 * b2[0] + (double)n3 * b3[0];
            double ky = (double)n2 * b2[1] + (double)n3 * b3[1];
            double kz = (double)n2 * b2[2] + (double)n3 * b3[2];
            ORC_PUSH(kx, ky, kz);
        }
    }
    for (int n1 = 1; n1 <= n1max; n1++) {
        for (int n2 = -n2max; n2 <= n2max; n2++) {
            for (int n3 = -n3max; n3 <= n3max; n3++) {
                double kx = (double)n1 * b1[0] + (double)n2 * b2[0] + (double)n3 * b3[0];
                double ky = (double)n1 * b1[1] + (double)n2 * b2[1] + (double)n3 * b3[1];
                double kz = (double)n1 * b1[2] + (double)n2 * b2[2] + (double)n3 * b3[2];
                ORC_PUSH(kx, ky, kz);
            }
        }
    }
#undef ORC_PUSH
    return count;
}

/* lode/spherical_expansion.rs:280-342 */
static double density_fourier(const orc_lode* lode, double k) {
    double p = (double)lode->h.exponent;
    double s2 = lode->h.smearing * lode->h.smearing;
    if (lode->h.exponent == 0) {
        return pow(4.0 * ORC_PI * s2, 0.75) * exp(-0.5 * k * k * s2);
    } else if (lode->h.exponent == 1) {
        double k2 = k * k;
        return 4.0 * ORC_PI * (exp(-0.5 * k2 * s2) / k2);
    }
    double p_eff = 3.0 - p;
    double factor = pow(ORC_PI, 1.5) * pow(2.0 * s2, 0.5 * p_eff) / orc_gamma(0.5 * p);
    double x = 0.5 * k * k * s2;
    double v;
    switch (lode->h.exponent) {
    case 2: v = sqrt(ORC_PI / x) * orc_erfc(sqrt(x)); break;
    case 3: v = -orc_expi(-x); break;
    case 4: v = 2.0 * (exp(-x) - sqrt(ORC_PI * x) * orc_erfc(sqrt(x))); break;
    case 5: v = exp(-x) + x * orc_expi(-x); break;
    case 6: v = ((2.0 - 4.0 * x) * exp(-x) + 4.0 * sqrt(ORC_PI) * pow(x, 1.5) * orc_erfc(sqrt(x))) / 3.0; break;
    case 7: v = (1.0 - x) * exp(-x) / 2.0 - x * x / 2.0 * orc_expi(-x); break;
    case 8: v = -2.0 / 15.0 * ((-3.0 + 2.0 * x - 4.0 * x * x) * exp(-x) + 4.0 * sqrt(ORC_PI) * pow(x, 2.5) * orc_erfc(sqrt(x))); break;
    case 9: v = (x * x - x + 2.0) * exp(-x) / 6.0 + x * x * x / 6.0 * orc_expi(-x); break;
    default: v = NAN;
    }
    return factor * v;
}

double orc_lode_density_fourier(const orc_lode* lode, double k) { return density_fourier(lode, k); }

void orc_lode_center_contribution(const orc_lode* lode, double* out) {
    for (int n = 0; n < lode->h.n_radial; n++) out[n] = lode->center[n];
}

/* compute_k0_contributions, lode/spherical_expansion.rs:347-390: factor(p, smearing) times the
 * l = 0 radial integral at k = 0; zero for 0 < p <= 3. Returns 1 when the term is non-zero. */
int orc_lode_k0_contribution(const orc_lode* lode, double* out) {
    int p = lode->h.exponent, N = lode->h.n_radial;
    double smearing = lode->h.smearing;
    double factor;
    if (p == 0) {
        double s2 = smearing * smearing;
        factor = pow(2.0 * ORC_PI * s2, 1.5) / pow(ORC_PI * s2, 0.75) / sqrt(4.0 * ORC_PI);
    } else if (p > 3) {
        double p_eff = 3.0 - p;
        factor = 0.5 * ORC_PI * pow(2.0, p_eff) / orc_gamma(0.5 * p) * pow(2.0, (p - 1.0) / 2.0) / -p_eff *
                 pow(smearing, -p_eff) / pow(smearing, 2.0 * p - 6.0);
    } else {
        for (int n = 0; n < N; n++) out[n] = 0.0;
        return 0;
    }
    double ri0[64];
    orc_radial_compute(lode->radial, 0, 0.0, ri0, NULL);
    for (int n = 0; n < N; n++) out[n] = factor * ri0[n];
    return 1;
}

static int species_index(int n_species, const int32_t* species, int32_t t) {
    for (int s = 0; s < n_species; s++) {
        if (species[s] == t) return s;
    }
    return -1;
}

/* centers == NULL: gradient rows of every center; else of centers[0 .. n_centers), stored at the
 * center's position in that list */
static int lode_system(const orc_lode* lode, size_t n_atoms, const double* positions, const int32_t* types,
                       const double* cell, int n_species, const int32_t* species, double* values, size_t n_centers,
                       const int32_t* centers, double* grads, double* self_grads) {
    int lmax = lode->h.max_angular, N = lode->h.n_radial;
    int L2 = (lmax + 1) * (lmax + 1);
    size_t blk = (size_t)L2 * N;
    size_t S = (size_t)n_species;
    int nk = orc_k_vectors(cell, lode->k_cutoff, NULL);
    if (nk == 0) return -2;
    double* kv = malloc(sizeof(double) * 4 * nk);
    orc_k_vectors(cell, lode->k_cutoff, kv);

    /* project_k_to_nlm, lode/spherical_expansion.rs:227-277: P[lm][n][k] */
    double* P = malloc(sizeof(double) * blk * nk);
    double* y = malloc(sizeof(double) * L2);
    double* ri = malloc(sizeof(double) * N);
    for (int k = 0; k < nk; k++) {
        orc_spherical_harmonics(lmax, kv + 4 * k, y, NULL);
        for (int l = 0; l <= lmax; l++) {
            orc_radial_compute(lode->radial, l, kv[4 * k + 3], ri, NULL);
            for (int lm = l * l; lm < (l + 1) * (l + 1); lm++)
                for (int n = 0; n < N; n++) P[((size_t)lm * N + n) * nk + k] = ri[n] * y[lm];
        }
    }

    /* compute_structure_factors, lode/spherical_expansion.rs:95-157 */
    double* cosv = malloc(sizeof(double) * n_atoms * nk);
    double* sinv = malloc(sizeof(double) * n_atoms * nk);
    for (size_t i = 0; i < n_atoms; i++) {
        const double* r = positions + 3 * i;
        for (int k = 0; k < nk; k++) {
            double kx = kv[4 * k + 3] * kv[4 * k + 0], ky = kv[4 * k + 3] * kv[4 * k + 1], kz = kv[4 * k + 3] * kv[4 * k + 2];
            double sarg = kx * r[0] + ky * r[1] + kz * r[2];
            cosv[i * nk + k] = cos(sarg);
            sinv[i * nk + k] = sin(sarg);
        }
    }
    double* sum_cos = calloc(S * nk, sizeof(double));
    double* sum_sin = calloc(S * nk, sizeof(double));
    for (size_t j = 0; j < n_atoms; j++) {
        int t = species_index(n_species, species, types[j]);
        for (int k = 0; k < nk; k++) {
            sum_cos[t * nk + k] += cosv[j * nk + k];
            sum_sin[t * nk + k] += sinv[j * nk + k];
        }
    }
    /* real/imag per center and neighbour type: [S][n_atoms][nk] */
    double* sf_real = calloc(S * n_atoms * nk, sizeof(double));
    double* sf_imag = calloc(S * n_atoms * nk, sizeof(double));
    for (size_t i = 0; i < n_atoms; i++) {
        for (size_t t = 0; t < S; t++) {
            for (int k = 0; k < nk; k++) {
                double re = cosv[i * nk + k] * sum_cos[t * nk + k] + sinv[i * nk + k] * sum_sin[t * nk + k];
                double im = sinv[i * nk + k] * sum_cos[t * nk + k] - cosv[i * nk + k] * sum_sin[t * nk + k];
                sf_real[(t * n_atoms + i) * nk + k] += 2.0 * re;
                sf_imag[(t * n_atoms + i) * nk + k] += 2.0 * im;
            }
        }
    }

    double* df = malloc(sizeof(double) * nk);
    for (int k = 0; k < nk; k++) df[k] = density_fourier(lode, kv[4 * k + 3]);
    double volume = fabs(det3(cell));
    double global_factor = 4.0 * ORC_PI / volume;

    memset(values, 0, sizeof(double) * n_atoms * S * blk);
    if (!centers) n_centers = n_atoms;
    if (grads) memset(grads, 0, sizeof(double) * n_centers * n_atoms * 3 * blk);
    if (self_grads) memset(self_grads, 0, sizeof(double) * n_centers * S * 3 * blk);

    /* centre contribution, lode/spherical_expansion.rs:397-449 (l=0, nt == ct) */
    double w = lode->h.center_atom_weight;
    for (size_t i = 0; i < n_atoms; i++) {
        int t = species_index(n_species, species, types[i]);
        for (int n = 0; n < N; n++) values[((i * S + t) * L2 + 0) * N + n] -= (1.0 - w) * lode->center[n];
    }

    /* k = 0 term, lode/spherical_expansion.rs:347-390,658-687 */
    double k0[64];
    if (orc_lode_k0_contribution(lode, k0)) {
        /* for each neighbour atom (counted with multiplicity), every center gets the k0 term in
         * block (ct, type_neighbor) */
        for (size_t j = 0; j < n_atoms; j++) {
            int nt = species_index(n_species, species, types[j]);
            for (size_t i = 0; i < n_atoms; i++) {
                for (int n = 0; n < N; n++) values[((i * S + nt) * L2 + 0) * N + n] += global_factor * k0[n];
            }
        }
    }

    double* kdep = malloc(sizeof(double) * nk);
    for (int l = 0; l <= lmax; l++) {
        double phase = (l % 2 == 0) ? ((l / 2) % 2 == 0 ? 1.0 : -1.0) : (((l + 1) / 2) % 2 == 0 ? 1.0 : -1.0);
        for (size_t t = 0; t < S; t++) {
            const double* sf = (l % 2 == 0) ? sf_real + t * n_atoms * nk : sf_imag + t * n_atoms * nk;
            for (size_t i = 0; i < n_atoms; i++) {
                for (int k = 0; k < nk; k++) kdep[k] = global_factor * phase * df[k] * sf[i * nk + k];
                for (int lm = l * l; lm < (l + 1) * (l + 1); lm++) {
                    for (int n = 0; n < N; n++) {
                        double v = 0.0;
                        const double* pp = P + ((size_t)lm * N + n) * nk;
                        for (int k = 0; k < nk; k++) v += kdep[k] * pp[k];
                        values[((i * S + t) * L2 + lm) * N + n] += v;
                    }
                }
            }
        }
        if (!grads && !self_grads) continue;
        /* gradients, lode/spherical_expansion.rs:771-871 */
        for (size_t c = 0; c < n_centers; c++) {
            size_t i = centers ? (size_t)centers[c] : c;
            int ct = species_index(n_species, species, types[i]);
            for (size_t j = 0; j < n_atoms; j++) {
                int nt_j = species_index(n_species, species, types[j]);
                for (size_t t = 0; t < S; t++) {
                    /* rows (i, j != i) exist only in the block of type_j; (i, i) in every block */
                    if (j != i && (int)t != nt_j) continue;
                    for (int k = 0; k < nk; k++) {
                        if (l % 2 == 0) {
                            kdep[k] = 2.0 * (sinv[i * nk + k] * cosv[j * nk + k] - cosv[i * nk + k] * sinv[j * nk + k]);
                        } else {
                            kdep[k] = -2.0 * (cosv[i * nk + k] * cosv[j * nk + k] + sinv[i * nk + k] * sinv[j * nk + k]);
                        }
                    }
                    if (i == j) {
                        double sf_phase = (l % 2 == 0) ? 1.0 : -1.0;
                        const double* sfg = (l % 2 == 0) ? sf_imag + t * n_atoms * nk : sf_real + t * n_atoms * nk;
                        for (int k = 0; k < nk; k++) {
                            if ((int)t == ct) {
                                kdep[k] = -(sf_phase * sfg[i * nk + k] - kdep[k]);
                            } else {
                                kdep[k] = -sf_phase * sfg[i * nk + k];
                            }
                        }
                    }
                    for (int k = 0; k < nk; k++) kdep[k] *= global_factor * phase * df[k];
                    double* out = (i == j) ? (self_grads ? self_grads + (c * S + t) * 3 * blk : NULL)
                                           : (grads ? grads + (c * n_atoms + j) * 3 * blk : NULL);
                    if (!out) continue;
                    for (int lm = l * l; lm < (l + 1) * (l + 1); lm++) {
                        for (int n = 0; n < N; n++) {
                            const double* pp = P + ((size_t)lm * N + n) * nk;
                            double g[3] = {0.0, 0.0, 0.0};
                            for (int k = 0; k < nk; k++) {
                                double f = kdep[k] * pp[k] * kv[4 * k + 3];
                                g[0] += f * kv[4 * k + 0];
                                g[1] += f * kv[4 * k + 1];
                                g[2] += f * kv[4 * k + 2];
                            }
                            for (int x = 0; x < 3; x++) out[(x * L2 + lm) * N + n] = g[x];
                        }
                    }
                }
            }
        }
    }

    free(kdep);
    free(df);
    free(sf_real);
    free(sf_imag);
    free(sum_cos);
    free(sum_sin);
    free(cosv);
    free(sinv);
    free(P);
    free(y);
    free(ri);
    free(kv);
    return 0;
}

int orc_lode_system(const orc_lode* lode, size_t n_atoms, const double* positions, const int32_t* types,
                    const double* cell, int n_species, const int32_t* species, double* values, double* grads,
                    double* self_grads) {
    return lode_system(lode, n_atoms, positions, types, cell, n_species, species, values, n_atoms, NULL, grads,
                       self_grads);
}

int orc_lode_system_centers(const orc_lode* lode, size_t n_atoms, const double* positions, const int32_t* types,
                            const double* cell, int n_species, const int32_t* species, double* values,
                            size_t n_centers, const int32_t* centers, double* grads, double* self_grads) {
    return lode_system(lode, n_atoms, positions, types, cell, n_species, species, values, n_centers, centers, grads,
                       self_grads);
}
