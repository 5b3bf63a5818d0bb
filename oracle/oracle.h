/*
 * FP64 CPU oracle for the featomic SOAP / LODE hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is a plain-C restatement of the
 * reference algorithms (featomic 0.6.0-rc2, ex-rascaline), used by the test
 * suite and by the CPU-baseline leg of bench.py as a *checker*. The shipped
 * product (rascaline_amd) never links, loads or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the reference repository root).
 */
#ifndef RASCALINE_AMD_ORACLE_H
#define RASCALINE_AMD_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same memory layout as featomic_pair_t (featomic/include/featomic.h:114-139). */
typedef struct orc_pair {
    size_t first;
    size_t second;
    double distance;
    double vector[3];
    int32_t shift[3];
} orc_pair;

/* ---- math (featomic/src/math) ---- */
double orc_gamma(double x);
double orc_ln_gamma(double x);
double orc_hyp1f1(double a, double b, double x);
double orc_hyp2f1(double a, double b, double c, double x);
double orc_expi(double x);
double orc_erfc(double x);
/* values: (lmax+1)^2, grads: 3*(lmax+1)^2 (x, y, z planes) or NULL */
void orc_spherical_harmonics(int lmax, const double direction[3], double* values, double* grads);
/* symmetric eigen-decomposition in place: a (n*n, row major) -> eigenvectors
 * in columns, w -> eigenvalues */
void orc_symmetric_eigen(int n, double* a, double* w);

/* ---- neighbour list (featomic/src/systems/neighbors.rs) ---- */
/* cell: 9 doubles, rows are the cell vectors; all zero = infinite cell.
 * Returns 0 on success. *pairs must be released with orc_free. */
int orc_neighbors(const double* positions, size_t n_atoms, const double* cell,
                  double cutoff, orc_pair** pairs, size_t* n_pairs);
void orc_free(void* ptr);

/* ---- radial integrals ---- */
typedef struct orc_radial orc_radial;
/* SOAP GTO x Gaussian density; spline_accuracy <= 0 means the exact path */
orc_radial* orc_soap_radial_new(double cutoff, double gaussian_width, int max_angular,
                                int n_radial, double spline_accuracy);
/* LODE GTO, k-space; spline_accuracy <= 0 means exact */
orc_radial* orc_lode_radial_new(double radius, int max_angular, int n_radial,
                                double k_cutoff, double spline_accuracy);
/* a tabulated radial integral: the Hermite spline through the given points, the same
 * for every l (shared/basis/radial/tabulated.rs:151-166); points sorted by position */
orc_radial* orc_radial_tabulated(int n, int n_points, const double* positions, const double* values,
                                 const double* derivatives);
void orc_radial_compute(const orc_radial* radial, int l, double r, double* values, double* grads);
int orc_radial_n_knots(const orc_radial* radial, int l);
/* knot positions, values, derivatives ([n_knots], [n_knots][n], [n_knots][n]) */
void orc_radial_knots(const orc_radial* radial, int l, double* positions, double* values, double* derivs);
/* GTO orthonormalisation matrix O (n x n, row-major), values <- values . O^T */
void orc_radial_orthonormalization(const orc_radial* radial, double* out);
void orc_radial_free(orc_radial* radial);

/* ---- SOAP hyper-parameters ---- */
typedef struct orc_soap_hypers {
    double cutoff;
    int smoothing;          /* 0: Step, 1: ShiftedCosine */
    double smoothing_width;
    double gaussian_width;
    int has_scaling;        /* Willatt2018 */
    double scaling_scale, scaling_rate, scaling_exponent;
    double center_atom_weight;
    int max_angular;
    int n_radial;
    double spline_accuracy; /* <= 0: exact */
    /* Per angular channel l = 0..max_angular, its own radial integral
     * (SoapRadialIntegralCacheByAngular, soap/radial_integral/mod.rs:138-183):
     * channel_n[l] radial functions (0: l is not a channel of the Explicit basis),
     * GTO x Gaussian when channel_points[l] == 0, else the tabulated spline points of
     * a DiracDelta density (spline.rs:46-58). channel_n == NULL: TensorProduct with one
     * GTO basis of n_radial functions. n_radial is the largest channel_n. */
    const int* channel_n;
    const int* channel_points;
    const double* const* channel_positions;   /* [points] */
    const double* const* channel_values;      /* [points][channel_n[l]] */
    const double* const* channel_derivatives; /* [points][channel_n[l]] */
} orc_soap_hypers;

typedef struct orc_soap orc_soap;
orc_soap* orc_soap_new(const orc_soap_hypers* hypers);
void orc_soap_free(orc_soap* soap);
/* self contribution (n values, l = 0) */
void orc_soap_self_contribution(const orc_soap* soap, double* out);
double orc_cutoff_smoothing(const orc_soap* soap, double r);
/* one pair's contribution (soap/spherical_expansion_pair.rs:323-422):
 * values [L2][N] = f * Y_lm * I_nl, grads [3][L2][N] = d values / d pair vector */
void orc_soap_pair_contribution(const orc_soap* soap, double distance, const double vector[3], int do_grad,
                                double* values, double* grads);
double orc_cutoff_smoothing_gradient(const orc_soap* soap, double r);

/*
 * Spherical expansion of one system, dense layout. S = n_species (the global
 * sorted list `species`), L2 = (lmax+1)^2, N = n_radial.
 *   values      [n_atoms][S][L2][N]             (self contribution included)
 *   pair_grads  [n_pairs][3][L2][N]   d c_first / d r_second, or NULL
 *   self_grads  [n_atoms][S][3][L2][N] or NULL
 *   strain      [n_atoms][S][3][3][L2][N] or NULL
 *   cell_grads  [n_atoms][S][3][3][L2][N] or NULL
 * Follows featomic/src/calculators/soap/spherical_expansion.rs:101-358.
 */
int orc_soap_se_system(const orc_soap* soap, size_t n_atoms, const int32_t* types,
                       size_t n_pairs, const orc_pair* pairs,
                       int n_species, const int32_t* species,
                       double* values, double* pair_grads, double* self_grads,
                       double* strain, double* cell_grads);

/*
 * CPU baseline: full SOAP power spectrum + position gradients for a batch of
 * systems, in the reference's loop structure (SE accumulate, then per PS row,
 * per property, sum over m). Outputs are folded into a checksum so the timing
 * does not depend on memory for the (large) outputs. Returns the number of
 * PS gradient rows computed; *checksum receives the sum of all outputs.
 */
int64_t orc_soap_ps_batch(const orc_soap* soap, int n_systems, const int64_t* atom_offsets,
                          const double* positions, const int32_t* types, const double* cells,
                          int n_species, const int32_t* species, int n_threads,
                          double* checksum);

/* power-spectrum position-gradient rows of one block from the spherical-expansion
 * gradient rows (table) and coefficients (values), see orc_soap.c */
void orc_ps_gradient_rows(int64_t n_rows, int L, int N, int S, int i1, int i2, double kappa, const double* values,
                          const int64_t* center, const int64_t* idx1, const int64_t* idx2, const double* table,
                          double* out);

/* ---- LODE ---- */
typedef struct orc_lode_hypers {
    double smearing;
    int exponent;
    double center_atom_weight;
    int max_angular;
    int n_radial;
    double radius;
    double k_cutoff;        /* <= 0: default 1.2 pi / smearing */
    double spline_accuracy; /* <= 0: exact */
} orc_lode_hypers;

typedef struct orc_lode orc_lode;
orc_lode* orc_lode_new(const orc_lode_hypers* hypers);
void orc_lode_free(orc_lode* lode);
/* number of k-vectors for a cell (math/k_vectors.rs:20-83) */
int orc_k_vectors(const double* cell, double k_cutoff, double* out_directions_norms /* [n][4] or NULL */);
/* rho(k) of the smeared power-law density (lode/spherical_expansion.rs:280-342) */
double orc_lode_density_fourier(const orc_lode* lode, double k);
/* central-atom term of the l = 0 radial basis, [n_radial] (lode/radial_integral/gto.rs:136-185) */
void orc_lode_center_contribution(const orc_lode* lode, double* out);
/* k = 0 term, [n_radial]; returns 0 (and zeros) when it vanishes (lode/spherical_expansion.rs:347-390) */
int orc_lode_k0_contribution(const orc_lode* lode, double* out);
/*
 * LODE spherical expansion of one (periodic) system, dense layout
 *   values [n_atoms][S][L2][N]
 *   grads  [n_atoms (center)][n_atoms (neighbor)][3][L2][N] or NULL; the row
 *          (i, j) holds d c_i^{type_j} / d r_j for j != i and, for j == i, the
 *          self gradient is stored in self_grads [n_atoms][S][3][L2][N]
 * Follows featomic/src/calculators/lode/spherical_expansion.rs:602-880.
 */
int orc_lode_system(const orc_lode* lode, size_t n_atoms, const double* positions,
                    const int32_t* types, const double* cell, int n_species,
                    const int32_t* species, double* values, double* grads, double* self_grads);

/* the same with the gradient rows of the listed centers only: grads [n_centers][n_atoms][3][L2][N],
 * self_grads [n_centers][S][3][L2][N] (values of every atom) */
int orc_lode_system_centers(const orc_lode* lode, size_t n_atoms, const double* positions,
                            const int32_t* types, const double* cell, int n_species, const int32_t* species,
                            double* values, size_t n_centers, const int32_t* centers, double* grads,
                            double* self_grads);

/* radial-integral surrogate network (orc_mlp.c): 1 -> 64 -> 64 -> n_out, SiLU */
#define ORC_MLP_HIDDEN 64
int64_t orc_mlp_n_params(int n_out);
void orc_mlp_forward(int n_out, int64_t n_samples, const double* params, const double* inputs, double* outputs);
/* returns the mean squared error; gradient has orc_mlp_n_params(n_out) entries */
double orc_mlp_loss_gradient(int n_out, int64_t n_samples, const double* params, const double* inputs,
                             const double* targets, double* gradient);

#ifdef __cplusplus
}
#endif

#endif
