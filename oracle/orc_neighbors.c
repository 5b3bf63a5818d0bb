/*
 * FP64 oracle: cell-list half neighbour list with periodic images.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restates featomic/src/systems/neighbors.rs:101-403, cell.rs and
 * types/matrix.rs with the same floating-point operation order, so that pair
 * vectors and distances are bit-identical. The reference sorts with
 * `sort_unstable_by_key((first, second))` (neighbors.rs:393), which leaves the
 * order of several images of one (first, second) pair unspecified; this oracle
 * breaks such ties by the cell shift, lexicographically.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

typedef struct {
    size_t index;
    int32_t shift[3];
} atom_data;

typedef struct {
    atom_data* atoms;
    size_t count, capacity;
} cell_bin;

static void bin_push(cell_bin* bin, atom_data atom) {
    if (bin->count == bin->capacity) {
        bin->capacity = bin->capacity ? 2 * bin->capacity : 8;
        bin->atoms = realloc(bin->atoms, bin->capacity * sizeof(atom_data));
    }
    bin->atoms[bin->count++] = atom;
}

/* python-style divmod, neighbors.rs:316-327 */
static void divmod(int32_t a, size_t b, int32_t* q, size_t* r) {
    int32_t bi = (int32_t)b;
    int32_t quotient = a / bi, remainder = a % bi;
    if (remainder < 0) {
        remainder += bi;
        quotient -= 1;
    }
    *q = quotient;
    *r = (size_t)remainder;
}

/* types/matrix.rs:243-247 */
static double det3(const double m[9]) {
    return m[0] * (m[4] * m[8] - m[7] * m[5]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
           m[2] * (m[3] * m[7] - m[4] * m[6]);
}

/* types/matrix.rs:210-225 */
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

static void cross(const double a[3], const double b[3], double c[3]) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

static double dot(const double a[3], const double b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* cell.rs:154-169 */
static void face_distances(const double m[9], double out[3]) {
    const double* a = m;
    const double* b = m + 3;
    const double* c = m + 6;
    double na[3], nb[3], nc[3];
    cross(b, c, na);
    cross(c, a, nb);
    cross(a, b, nc);
    double la = sqrt(dot(na, na)), lb = sqrt(dot(nb, nb)), lc = sqrt(dot(nc, nc));
    for (int k = 0; k < 3; k++) {
        na[k] /= la;
        nb[k] /= lb;
        nc[k] /= lc;
    }
    out[0] = fabs(dot(na, a));
    out[1] = fabs(dot(nb, b));
    out[2] = fabs(dot(nc, c));
}

static int cmp_pair(const void* pa, const void* pb) {
    const orc_pair* a = pa;
    const orc_pair* b = pb;
    if (a->first != b->first) return a->first < b->first ? -1 : 1;
    if (a->second != b->second) return a->second < b->second ? -1 : 1;
    for (int k = 0; k < 3; k++) {
        if (a->shift[k] != b->shift[k]) return a->shift[k] < b->shift[k] ? -1 : 1;
    }
    return 0;
}

int orc_neighbors(const double* positions, size_t n_atoms, const double* cell, double cutoff,
                  orc_pair** out_pairs, size_t* out_count) {
    int infinite = 1;
    for (int k = 0; k < 9; k++) {
        if (cell[k] != 0.0) infinite = 0;
    }
    double inverse[9] = {0};
    double faces[3];
    if (infinite) {
        faces[0] = faces[1] = faces[2] = 1.0;
    } else {
        inverse3(cell, inverse);
        face_distances(cell, faces);
    }

    /* CellList::new, neighbors.rs:101-163 */
    double n_cells_f[3];
    for (int k = 0; k < 3; k++) {
        double v = trunc(faces[k] / cutoff);
        n_cells_f[k] = v < 1.0 ? 1.0 : v;
    }
    double total = n_cells_f[0] * n_cells_f[1] * n_cells_f[2];
    if (total > 1e5) {
        double rxy = n_cells_f[0] / n_cells_f[1];
        double ryz = n_cells_f[1] / n_cells_f[2];
        n_cells_f[2] = trunc(cbrt(1e5 / (rxy * ryz * ryz)));
        n_cells_f[1] = trunc(ryz * n_cells_f[2]);
        n_cells_f[0] = trunc(rxy * n_cells_f[1]);
    }
    int32_t n_search[3];
    size_t n_cells[3];
    for (int k = 0; k < 3; k++) {
        n_search[k] = (int32_t)ceil(cutoff * n_cells_f[k] / faces[k]);
        n_cells[k] = (size_t)n_cells_f[k];
    }
    for (int k = 0; k < 3; k++) {
        if (n_search[k] < 1) n_search[k] = 1;
        if (n_cells[k] == 1 && infinite) n_search[k] = 0;
    }
    size_t n_bins = n_cells[0] * n_cells[1] * n_cells[2];
    cell_bin* bins = calloc(n_bins, sizeof(cell_bin));

    /* add_atom, neighbors.rs:167-201 */
    for (size_t i = 0; i < n_atoms; i++) {
        const double* p = positions + 3 * i;
        double frac[3];
        if (infinite) {
            frac[0] = p[0];
            frac[1] = p[1];
            frac[2] = p[2];
        } else {
            /* Vector3D * Matrix3, matrix.rs:400-407 */
            frac[0] = p[0] * inverse[0] + p[1] * inverse[3] + p[2] * inverse[6];
            frac[1] = p[0] * inverse[1] + p[1] * inverse[4] + p[2] * inverse[7];
            frac[2] = p[0] * inverse[2] + p[1] * inverse[5] + p[2] * inverse[8];
        }
        int32_t idx[3];
        for (int k = 0; k < 3; k++) idx[k] = (int32_t)floor(frac[k] * (double)n_cells[k]);
        atom_data atom;
        atom.index = i;
        size_t bin[3];
        if (infinite) {
            for (int k = 0; k < 3; k++) {
                /* `cell_index as usize` wraps a negative i32 to a huge value,
                 * which the clamp then sends to the last bin */
                size_t v = idx[k] < 0 ? (size_t)-1 : (size_t)idx[k];
                bin[k] = v > n_cells[k] - 1 ? n_cells[k] - 1 : v;
                atom.shift[k] = 0;
            }
        } else {
            for (int k = 0; k < 3; k++) divmod(idx[k], n_cells[k], &atom.shift[k], &bin[k]);
        }
        bin_push(&bins[(bin[0] * n_cells[1] + bin[1]) * n_cells[2] + bin[2]], atom);
    }

    size_t cap = 64, count = 0;
    orc_pair* pairs = malloc(cap * sizeof(orc_pair));
    double cutoff2 = cutoff * cutoff;

    /* CellList::pairs, neighbors.rs:217-309, then the distance filter :364-389 */
    for (size_t cx = 0; cx < n_cells[0]; cx++)
    for (size_t cy = 0; cy < n_cells[1]; cy++)
    for (size_t cz = 0; cz < n_cells[2]; cz++) {
        cell_bin* current = &bins[(cx * n_cells[1] + cy) * n_cells[2] + cz];
        for (int32_t dx = -n_search[0]; dx <= n_search[0]; dx++)
        for (int32_t dy = -n_search[1]; dy <= n_search[1]; dy++)
        for (int32_t dz = -n_search[2]; dz <= n_search[2]; dz++) {
            int32_t ci[3] = {(int32_t)cx + dx, (int32_t)cy + dy, (int32_t)cz + dz};
            int32_t cshift[3];
            size_t nb[3];
            for (int k = 0; k < 3; k++) divmod(ci[k], n_cells[k], &cshift[k], &nb[k]);
            cell_bin* neighbor = &bins[(nb[0] * n_cells[1] + nb[1]) * n_cells[2] + nb[2]];
            for (size_t a = 0; a < current->count; a++) {
                atom_data ai = current->atoms[a];
                for (size_t b = 0; b < neighbor->count; b++) {
                    atom_data aj = neighbor->atoms[b];
                    if (ai.index > aj.index) continue;
                    int32_t s[3];
                    for (int k = 0; k < 3; k++) s[k] = cshift[k] + ai.shift[k] - aj.shift[k];
                    int zero = s[0] == 0 && s[1] == 0 && s[2] == 0;
                    if (ai.index == aj.index) {
                        if (zero) continue;
                        if (s[0] + s[1] + s[2] < 0) continue;
                        if ((s[0] + s[1] + s[2] == 0) && (s[2] < 0 || (s[2] == 0 && s[1] < 0))) continue;
                    }
                    if (infinite && !zero) continue;

                    const double* pi = positions + 3 * ai.index;
                    const double* pj = positions + 3 * aj.index;
                    double v[3];
                    for (int k = 0; k < 3; k++) v[k] = pj[k] - pi[k];
                    /* CellShift::cartesian, neighbors.rs:52-57 */
                    double sx = cell[0] * (double)s[0] + cell[3] * (double)s[1] + cell[6] * (double)s[2];
                    double sy = cell[1] * (double)s[0] + cell[4] * (double)s[1] + cell[7] * (double)s[2];
                    double sz = cell[2] * (double)s[0] + cell[5] * (double)s[1] + cell[8] * (double)s[2];
                    v[0] += sx;
                    v[1] += sy;
                    v[2] += sz;
                    double d2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
                    if (d2 < cutoff2) {
                        if (count == cap) {
                            cap *= 2;
                            pairs = realloc(pairs, cap * sizeof(orc_pair));
                        }
                        orc_pair* pp = &pairs[count++];
                        pp->first = ai.index;
                        pp->second = aj.index;
                        pp->distance = sqrt(d2);
                        pp->vector[0] = v[0];
                        pp->vector[1] = v[1];
                        pp->vector[2] = v[2];
                        pp->shift[0] = s[0];
                        pp->shift[1] = s[1];
                        pp->shift[2] = s[2];
                    }
                }
            }
        }
    }

    for (size_t b = 0; b < n_bins; b++) free(bins[b].atoms);
    free(bins);
    qsort(pairs, count, sizeof(orc_pair), cmp_pair);
    *out_pairs = pairs;
    *out_count = count;
    return 0;
}
