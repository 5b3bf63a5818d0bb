/*
 * FP64 oracle: SOAP spherical expansion and power spectrum.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define ORC_PI 3.141592653589793

struct orc_soap {
    orc_soap_hypers h;
    orc_radial* radial;   /* TensorProduct: one GTO basis for every channel */
    orc_radial** by_l;    /* per channel l: its radial integral, or NULL (not a channel) */
    int* n_l;             /* per channel l: its radial size */
};

void orc_soap_free(orc_soap* s) {
    if (!s) return;
    orc_radial_free(s->radial);
    if (s->by_l) {
        for (int l = 0; l <= s->h.max_angular; l++) orc_radial_free(s->by_l[l]);
        free(s->by_l);
    }
    free(s->n_l);
    free(s);
}

orc_soap* orc_soap_new(const orc_soap_hypers* hypers) {
    orc_soap* s = calloc(1, sizeof(orc_soap));
    s->h = *hypers;
    if (!hypers->channel_n) {
        s->radial = orc_soap_radial_new(hypers->cutoff, hypers->gaussian_width, hypers->max_angular,
                                        hypers->n_radial, hypers->spline_accuracy);
        if (!s->radial) {
            free(s);
            return NULL;
        }
        return s;
    }
    /* SoapRadialIntegralCacheByAngular::new, soap/radial_integral/mod.rs:138-183 */
    int L = hypers->max_angular + 1;
    s->by_l = calloc(L, sizeof(orc_radial*));
    s->n_l = calloc(L, sizeof(int));
    for (int l = 0; l < L; l++) {
        int n = hypers->channel_n[l];
        s->n_l[l] = n;
        if (n == 0) continue;
        if (hypers->channel_points[l] > 0) {
            s->by_l[l] = orc_radial_tabulated(n, hypers->channel_points[l], hypers->channel_positions[l],
                                              hypers->channel_values[l], hypers->channel_derivatives[l]);
        } else {
            s->by_l[l] = orc_soap_radial_new(hypers->cutoff, hypers->gaussian_width, l, n, hypers->spline_accuracy);
        }
        if (!s->by_l[l]) {
            orc_soap_free(s);
            return NULL;
        }
    }
    return s;
}

/* radial integral of channel l at r, in the [n_radial] layout (zero beyond the
 * channel's own size, and for a channel outside the basis) */
static void soap_radial(const orc_soap* s, int l, double r, double* values, double* grads) {
    if (!s->by_l) {
        orc_radial_compute(s->radial, l, r, values, grads);
        return;
    }
    int N = s->h.n_radial, n = s->n_l[l];
    double tv[64], tg[64];
    for (int i = 0; i < N; i++) {
        values[i] = 0.0;
        if (grads) grads[i] = 0.0;
    }
    if (n == 0) return;
    orc_radial_compute(s->by_l[l], l, r, tv, grads ? tg : NULL);
    for (int i = 0; i < n; i++) {
        values[i] = tv[i];
        if (grads) grads[i] = tg[i];
    }
}

/* featomic/src/calculators/soap/cutoff.rs:48-64 */
double orc_cutoff_smoothing(const orc_soap* s, double r) {
    double rc = s->h.cutoff;
    if (s->h.smoothing == 0) return r >= rc ? 0.0 : 1.0;
    double w = s->h.smoothing_width;
    if (r <= rc - w) return 1.0;
    if (r >= rc) return 0.0;
    double x = ORC_PI * (r - rc + w) / w;
    return 0.5 * (1.0 + cos(x));
}

/* featomic/src/calculators/soap/cutoff.rs:67-79 */
double orc_cutoff_smoothing_gradient(const orc_soap* s, double r) {
    double rc = s->h.cutoff;
    if (s->h.smoothing == 0) return 0.0;
    double w = s->h.smoothing_width;
    if (r <= rc - w || r >= rc) return 0.0;
    double x = ORC_PI * (r - rc + w) / w;
    return -0.5 * ORC_PI * sin(x) / w;
}

/* featomic/src/calculators/shared/density.rs:108-140 and
 * soap/spherical_expansion_pair.rs:189-210 */
static double scaling_functions(const orc_soap* s, double r) {
    double scaling = 1.0;
    if (s->h.has_scaling) {
        scaling = s->h.scaling_rate / (s->h.scaling_rate + pow(r / s->h.scaling_scale, s->h.scaling_exponent));
    }
    return scaling * orc_cutoff_smoothing(s, r);
}

static double scaling_functions_gradient(const orc_soap* s, double r) {
    double scaling = 1.0, scaling_grad = 0.0;
    if (s->h.has_scaling) {
        double rate = s->h.scaling_rate, scale = s->h.scaling_scale, ex = s->h.scaling_exponent;
        scaling = rate / (rate + pow(r / scale, ex));
        double rs = r / scale;
        double rs_m1 = pow(rs, ex - 1.0);
        double rs_m = rs * rs_m1;
        double factor = -rate * ex / scale;
        scaling_grad = factor * rs_m1 / ((rate + rs_m) * (rate + rs_m));
    }
    return orc_cutoff_smoothing_gradient(s, r) * scaling + orc_cutoff_smoothing(s, r) * scaling_grad;
}

/* soap/spherical_expansion_pair.rs:222-252 */
void orc_soap_self_contribution(const orc_soap* s, double* out) {
    int n = s->h.n_radial;
    double* ri = malloc(sizeof(double) * n);
    soap_radial(s, 0, 0.0, ri, NULL);
    double z[3] = {0.0, 0.0, 1.0};
    double y[1024];
    orc_spherical_harmonics(s->h.max_angular, z, y, NULL);
    double factor = s->h.center_atom_weight * scaling_functions(s, 0.0) * y[0];
    for (int i = 0; i < n; i++) out[i] = factor * ri[i];
    free(ri);
}

/*
 * Per-pair contribution, soap/spherical_expansion_pair.rs:323-422:
 *   values [L2][N]      f * Y_lm * I_nl
 *   grads  [3][L2][N]   d/d r_ij of the values
 */
static void pair_contribution(const orc_soap* s, double distance, const double vector[3], int do_grad,
                              double* values, double* grads, double* work) {
    int lmax = s->h.max_angular, N = s->h.n_radial;
    int L2 = (lmax + 1) * (lmax + 1);
    double dir[3];
    if (distance < 1e-6) {
        dir[0] = 0.0;
        dir[1] = 0.0;
        dir[2] = 1.0;
    } else {
        dir[0] = vector[0] / distance;
        dir[1] = vector[1] / distance;
        dir[2] = vector[2] / distance;
    }
    double* y = work;                 /* L2 */
    double* dy = y + L2;              /* 3 L2 */
    double* ri = dy + 3 * L2;         /* (lmax+1) N */
    double* dri = ri + (lmax + 1) * N;
    for (int l = 0; l <= lmax; l++) {
        soap_radial(s, l, distance, ri + l * N, do_grad ? dri + l * N : NULL);
    }
    orc_spherical_harmonics(lmax, dir, y, do_grad ? dy : NULL);
    double f = scaling_functions(s, distance);
    double fg = scaling_functions_gradient(s, distance);
    for (int l = 0; l <= lmax; l++) {
        for (int m = 0; m < 2 * l + 1; m++) {
            int lm = l * l + m;
            for (int n = 0; n < N; n++) values[lm * N + n] = f * y[lm] * ri[l * N + n];
        }
        if (do_grad) {
            for (int m = 0; m < 2 * l + 1; m++) {
                int lm = l * l + m;
                double sv = y[lm];
                for (int n = 0; n < N; n++) {
                    double r = ri[l * N + n], rg = dri[l * N + n];
                    for (int x = 0; x < 3; x++) {
                        grads[(x * L2 + lm) * N + n] = fg * dir[x] * r * sv + f * rg * dir[x] * sv +
                                                       f * r * dy[x * L2 + lm] / distance;
                    }
                }
            }
        }
    }
}

void orc_soap_pair_contribution(const orc_soap* s, double distance, const double vector[3], int do_grad,
                                double* values, double* grads) {
    int lmax = s->h.max_angular, N = s->h.n_radial;
    int L2 = (lmax + 1) * (lmax + 1);
    double* work = malloc(sizeof(double) * (4 * L2 + 2 * (lmax + 1) * N));
    pair_contribution(s, distance, vector, do_grad, values, grads, work);
    free(work);
}

static int species_index(int n_species, const int32_t* species, int32_t t) {
    for (int s = 0; s < n_species; s++) {
        if (species[s] == t) return s;
    }
    return -1;
}

int orc_soap_se_system(const orc_soap* s, size_t n_atoms, const int32_t* types, size_t n_pairs,
                       const orc_pair* pairs, int n_species, const int32_t* species, double* values,
                       double* pair_grads, double* self_grads, double* strain, double* cell_grads) {
    int lmax = s->h.max_angular, N = s->h.n_radial;
    int L2 = (lmax + 1) * (lmax + 1);
    size_t blk = (size_t)L2 * N;
    int do_grad = pair_grads || self_grads || strain || cell_grads;
    size_t S = (size_t)n_species;
    memset(values, 0, sizeof(double) * n_atoms * S * blk);
    if (self_grads) memset(self_grads, 0, sizeof(double) * n_atoms * S * 3 * blk);
    if (strain) memset(strain, 0, sizeof(double) * n_atoms * S * 9 * blk);
    if (cell_grads) memset(cell_grads, 0, sizeof(double) * n_atoms * S * 9 * blk);

    /* self contribution, soap/spherical_expansion.rs:48-96 */
    double self_c[64];
    orc_soap_self_contribution(s, self_c);
    for (size_t i = 0; i < n_atoms; i++) {
        int t = species_index(n_species, species, types[i]);
        if (t < 0) return -1;
        for (int n = 0; n < N; n++) values[((i * S + t) * L2 + 0) * N + n] += self_c[n];
    }

    double* cv = malloc(sizeof(double) * blk);
    double* cg = malloc(sizeof(double) * 3 * blk);
    double* work = malloc(sizeof(double) * (4 * L2 + 2 * (lmax + 1) * N));

    /* accumulate_all_pairs, soap/spherical_expansion.rs:188-355 */
    for (size_t p = 0; p < n_pairs; p++) {
        const orc_pair* pair = &pairs[p];
        pair_contribution(s, pair->distance, pair->vector, do_grad, cv, cg, work);
        for (int side = 0; side < 2; side++) {
            size_t center = side == 0 ? pair->first : pair->second;
            size_t neighbor = side == 0 ? pair->second : pair->first;
            double sign_v[32], sign_g[32];
            for (int l = 0; l <= lmax; l++) {
                double m1l = (l % 2 == 0) ? 1.0 : -1.0;
                sign_v[l] = side == 0 ? 1.0 : m1l;
                sign_g[l] = side == 0 ? 1.0 : -m1l;
            }
            double vsign = side == 0 ? 1.0 : -1.0; /* pair vector / shift sign */
            int nt = species_index(n_species, species, types[neighbor]);
            double* out = values + (center * S + nt) * blk;
            for (int l = 0; l <= lmax; l++) {
                for (int lm = l * l; lm < (l + 1) * (l + 1); lm++) {
                    for (int n = 0; n < N; n++) out[lm * N + n] += sign_v[l] * cv[lm * N + n];
                }
            }
            if (!do_grad) continue;
            if (side == 0 && pair_grads) memcpy(pair_grads + p * 3 * blk, cg, sizeof(double) * 3 * blk);
            if (pair->first != pair->second && self_grads) {
                double* sg = self_grads + (center * S + nt) * 3 * blk;
                for (int x = 0; x < 3; x++)
                    for (int l = 0; l <= lmax; l++)
                        for (int lm = l * l; lm < (l + 1) * (l + 1); lm++)
                            for (int n = 0; n < N; n++)
                                sg[(x * L2 + lm) * N + n] -= sign_g[l] * cg[(x * L2 + lm) * N + n];
            }
            if (cell_grads) {
                double* cgo = cell_grads + (center * S + nt) * 9 * blk;
                for (int a = 0; a < 3; a++) {
                    double shift = vsign * (double)pair->shift[a];
                    for (int x = 0; x < 3; x++)
                        for (int l = 0; l <= lmax; l++)
                            for (int lm = l * l; lm < (l + 1) * (l + 1); lm++)
                                for (int n = 0; n < N; n++)
                                    cgo[((a * 3 + x) * L2 + lm) * N + n] +=
                                        shift * (sign_g[l] * cg[(x * L2 + lm) * N + n]);
                }
            }
            if (strain) {
                double* so = strain + (center * S + nt) * 9 * blk;
                for (int a = 0; a < 3; a++) {
                    double v = vsign * pair->vector[a];
                    for (int x = 0; x < 3; x++)
                        for (int l = 0; l <= lmax; l++)
                            for (int lm = l * l; lm < (l + 1) * (l + 1); lm++)
                                for (int n = 0; n < N; n++)
                                    so[((a * 3 + x) * L2 + lm) * N + n] += v * (sign_g[l] * cg[(x * L2 + lm) * N + n]);
                }
            }
        }
    }
    free(cv);
    free(cg);
    free(work);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* CPU baseline: SE + PS (values + position gradients) in the reference loop
 * structure (soap/power_spectrum.rs:584-678), one system per thread. */

typedef struct {
    size_t atom;      /* neighbour atom index */
    size_t first, count; /* range into the per-center pair index list */
} nb_group;

static int cmp_size_t(const void* a, const void* b) {
    size_t x = *(const size_t*)a, y = *(const size_t*)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}

static double soap_ps_system(const orc_soap* s, size_t n_atoms, const double* positions, const int32_t* types,
                             const double* cell, int n_species, const int32_t* species, int64_t* n_rows) {
    int lmax = s->h.max_angular, N = s->h.n_radial;
    int L2 = (lmax + 1) * (lmax + 1), L = lmax + 1;
    size_t blk = (size_t)L2 * N;
    size_t S = (size_t)n_species;
    orc_pair* pairs;
    size_t n_pairs;
    orc_neighbors(positions, n_atoms, cell, s->h.cutoff, &pairs, &n_pairs);

    double* values = malloc(sizeof(double) * n_atoms * S * blk);
    double* pair_grads = malloc(sizeof(double) * (n_pairs ? n_pairs : 1) * 3 * blk);
    double* self_grads = malloc(sizeof(double) * n_atoms * S * 3 * blk);
    orc_soap_se_system(s, n_atoms, types, n_pairs, pairs, n_species, species, values, pair_grads, self_grads,
                       NULL, NULL);

    /* neighbour types present around each atom (incl. itself) */
    unsigned* mask = calloc(n_atoms, sizeof(unsigned));
    size_t* deg = calloc(n_atoms + 1, sizeof(size_t));
    for (size_t i = 0; i < n_atoms; i++) mask[i] = 1u << species_index(n_species, species, types[i]);
    for (size_t p = 0; p < n_pairs; p++) {
        mask[pairs[p].first] |= 1u << species_index(n_species, species, types[pairs[p].second]);
        mask[pairs[p].second] |= 1u << species_index(n_species, species, types[pairs[p].first]);
        deg[pairs[p].first]++;
        deg[pairs[p].second]++;
    }
    size_t* off = calloc(n_atoms + 1, sizeof(size_t));
    for (size_t i = 0; i < n_atoms; i++) off[i + 1] = off[i] + deg[i];
    /* entries: (neighbour << 32 | pair index) would overflow for big systems;
     * store pair indices and sort them by neighbour */
    size_t* ent = malloc(sizeof(size_t) * (off[n_atoms] ? off[n_atoms] : 1) * 2);
    size_t* fill = calloc(n_atoms, sizeof(size_t));
    for (size_t p = 0; p < n_pairs; p++) {
        size_t a = pairs[p].first, b = pairs[p].second;
        size_t ia = off[a] + fill[a]++;
        ent[2 * ia] = b;
        ent[2 * ia + 1] = p;
        size_t ib = off[b] + fill[b]++;
        ent[2 * ib] = a;
        ent[2 * ib + 1] = p;
    }
    for (size_t i = 0; i < n_atoms; i++) {
        qsort(ent + 2 * off[i], off[i + 1] - off[i], 2 * sizeof(size_t), cmp_size_t);
    }

    double checksum = 0.0;
    int64_t rows = 0;
    double* dc = malloc(sizeof(double) * 3 * blk);
    double norm[32];
    for (int l = 0; l <= lmax; l++) norm[l] = (l % 2 == 0 ? 1.0 : -1.0) * sqrt(2.0 * l + 1.0);
    for (size_t i = 0; i < n_atoms; i++) {
        int ct = species_index(n_species, species, types[i]);
        (void)ct;
        for (int t1 = 0; t1 < n_species; t1++) {
            for (int t2 = t1; t2 < n_species; t2++) {
                if (!((mask[i] >> t1) & 1u) || !((mask[i] >> t2) & 1u)) continue;
                double kappa = (t1 != t2) ? 1.4142135623730951 : 1.0;
                const double* c1 = values + (i * S + t1) * blk;
                const double* c2 = values + (i * S + t2) * blk;
                /* values row */
                for (int l = 0; l < L; l++)
                    for (int n1 = 0; n1 < N; n1++)
                        for (int n2 = 0; n2 < N; n2++) {
                            double sum = 0.0;
                            for (int m = l * l; m < (l + 1) * (l + 1); m++) sum += c1[m * N + n1] * c2[m * N + n2];
                            checksum += kappa * sum / norm[l];
                        }
                /* gradient rows: the center itself, then unique neighbours of type t1 or t2 */
                size_t e = off[i];
                size_t prev = (size_t)-1;
                for (;;) {
                    size_t j;
                    if (prev == (size_t)-1) {
                        j = i;
                    } else {
                        while (e < off[i + 1] && (ent[2 * e] == prev || ent[2 * e] == i)) e++;
                        if (e >= off[i + 1]) break;
                        j = ent[2 * e];
                    }
                    prev = j;
                    int tj = species_index(n_species, species, types[j]);
                    int has1, has2;
                    const double* g1 = NULL;
                    const double* g2 = NULL;
                    if (j == i) {
                        has1 = has2 = 1;
                        g1 = self_grads + (i * S + t1) * 3 * blk;
                        g2 = self_grads + (i * S + t2) * 3 * blk;
                    } else {
                        if (tj != t1 && tj != t2) continue;
                        memset(dc, 0, sizeof(double) * 3 * blk);
                        for (size_t q = e; q < off[i + 1] && ent[2 * q] == j; q++) {
                            const orc_pair* pr = &pairs[ent[2 * q + 1]];
                            const double* pg = pair_grads + ent[2 * q + 1] * 3 * blk;
                            for (int x = 0; x < 3; x++)
                                for (int l = 0; l < L; l++) {
                                    double f = (pr->first == i) ? 1.0 : -((l % 2 == 0) ? 1.0 : -1.0);
                                    for (int lm = l * l; lm < (l + 1) * (l + 1); lm++)
                                        for (int n = 0; n < N; n++)
                                            dc[(x * L2 + lm) * N + n] += f * pg[(x * L2 + lm) * N + n];
                                }
                        }
                        has1 = tj == t1;
                        has2 = tj == t2;
                        g1 = dc;
                        g2 = dc;
                    }
                    rows++;
                    for (int l = 0; l < L; l++)
                        for (int n1 = 0; n1 < N; n1++)
                            for (int n2 = 0; n2 < N; n2++) {
                                double sum[3] = {0.0, 0.0, 0.0};
                                if (has1) {
                                    for (int m = l * l; m < (l + 1) * (l + 1); m++) {
                                        double v2 = c2[m * N + n2];
                                        for (int x = 0; x < 3; x++) sum[x] += v2 * g1[(x * L2 + m) * N + n1];
                                    }
                                }
                                if (has2) {
                                    for (int m = l * l; m < (l + 1) * (l + 1); m++) {
                                        double v1 = c1[m * N + n1];
                                        for (int x = 0; x < 3; x++) sum[x] += v1 * g2[(x * L2 + m) * N + n2];
                                    }
                                }
                                checksum += kappa * (sum[0] + sum[1] + sum[2]) / norm[l];
                            }
                }
            }
        }
    }
    free(dc);
    free(fill);
    free(ent);
    free(off);
    free(deg);
    free(mask);
    free(values);
    free(pair_grads);
    free(self_grads);
    free(pairs);
    *n_rows = rows;
    return checksum;
}

int64_t orc_soap_ps_batch(const orc_soap* s, int n_systems, const int64_t* atom_offsets, const double* positions,
                          const int32_t* types, const double* cells, int n_species, const int32_t* species,
                          int n_threads, double* checksum) {
    int64_t total_rows = 0;
    double total = 0.0;
#pragma omp parallel for schedule(dynamic, 1) num_threads(n_threads) reduction(+ : total_rows, total)
    for (int sys = 0; sys < n_systems; sys++) {
        int64_t a0 = atom_offsets[sys], a1 = atom_offsets[sys + 1];
        int64_t rows = 0;
        total += soap_ps_system(s, (size_t)(a1 - a0), positions + 3 * a0, types + a0, cells + 9 * sys, n_species,
                                species, &rows);
        total_rows += rows;
    }
    *checksum = total;
    return total_rows;
}

/* Power-spectrum position-gradient rows of one block (soap/power_spectrum.rs:621-678):
 * row r of center c = center[r] holds, per l,
 *   kappa / ((-1)^l sqrt(2l+1)) * (sum_m g1[x][m][n1] c2[m][n2] + sum_m c1[m][n1] g2[x][m][n2])
 * with c1, c2 the center's coefficients of the block's two neighbour types and g1, g2 the
 * spherical-expansion gradient rows table[idx1[r]], table[idx2[r]] (table entry 0 is zero when
 * the row has no such term). values [n_atoms][S][L2][N], table [entries][3][L2][N],
 * out [n_rows][3][L N N]. */
void orc_ps_gradient_rows(int64_t n_rows, int L, int N, int S, int i1, int i2, double kappa, const double* values,
                          const int64_t* center, const int64_t* idx1, const int64_t* idx2, const double* table,
                          double* out) {
    const int L2 = L * L;
    const int P = L * N * N;
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < n_rows; r++) {
        const double* c1 = values + ((size_t)center[r] * S + i1) * L2 * N;
        const double* c2 = values + ((size_t)center[r] * S + i2) * L2 * N;
        const double* g1 = table + (size_t)idx1[r] * 3 * L2 * N;
        const double* g2 = table + (size_t)idx2[r] * 3 * L2 * N;
        double* o = out + (size_t)r * 3 * P;
        for (int x = 0; x < 3; x++) {
            for (int l = 0; l < L; l++) {
                const double norm = (l % 2 == 0 ? 1.0 : -1.0) * sqrt(2.0 * l + 1.0);
                for (int a = 0; a < N; a++) {
                    for (int b = 0; b < N; b++) {
                        double acc = 0.0;
                        for (int m = l * l; m < (l + 1) * (l + 1); m++) {
                            acc += g1[((size_t)x * L2 + m) * N + a] * c2[m * N + b];
                            acc += c1[m * N + a] * g2[((size_t)x * L2 + m) * N + b];
                        }
                        o[(size_t)x * P + (l * N + a) * N + b] = kappa * acc / norm;
                    }
                }
            }
        }
    }
}
