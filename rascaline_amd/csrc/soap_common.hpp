// Per-pair device math shared by the SOAP kernels: real spherical harmonics
// with Cartesian gradients, radial integral (spline or exact GTO), and the
// cutoff / radial scaling functions.
#pragma once

#include <hip/hip_runtime.h>

#include "device.hpp"
#include "special.hpp"

namespace rascaline_amd {

// Legendre storage index for (l, m >= 0)
__device__ __forceinline__ int leg_index(int l, int m) { return l * (l + 1) / 2 + m; }

// Real spherical harmonics Y_lm(d) for l < L (index l*l + l + m) and, when
// `grads` is not null, r * dY/dr_x (3 planes of L*L values).
// Follows featomic/src/math/spherical_harmonics.rs:203-424 (arXiv:1410.1748
// recursion, Condon-Shortley phase cancelled, the `sin(theta) > 0.1` switch for
// m P / sin(theta)). `P`, `DP`, `PT` are scratch arrays of L(L+1)/2 values;
// `A`, `B` the recursion coefficients indexed like P.
template <typename T>
__device__ void spherical_harmonics(int L, const T* d, T* Y, T* grads, T* P, T* DP, T* PT, const T* A, const T* B) {
    const T SQRT_1_OVER_2PI = (T)0.3989422804014327;
    const T SQRT_3 = (T)1.7320508075688772;
    const T SQRT_3_OVER_2 = (T)1.224744871391589;
    const T SQRT_2 = (T)1.4142135623730951;
    const T x = d[0], y = d[1], z = d[2];
    const T sqrt_xy = (T)hypot((double)x, (double)y);
    const T cos_theta = z, sin_theta = sqrt_xy;
    T cos_phi = 1, sin_phi = 0;
    if ((double)sqrt_xy > (sizeof(T) == 8 ? 2.220446049250313e-16 : 1.1920929e-07)) {
        cos_phi = x / sqrt_xy;
        sin_phi = y / sqrt_xy;
    }
    T value = SQRT_1_OVER_2PI;
    P[0] = value;
    if (L > 1) {
        P[leg_index(1, 0)] = cos_theta * SQRT_3 * value;
        value *= -SQRT_3_OVER_2 * sin_theta;
        P[leg_index(1, 1)] = value;
        for (int l = 2; l < L; l++) {
            for (int m = 0; m < l - 1; m++) {
                const int k = leg_index(l, m);
                P[k] = A[k] * (cos_theta * P[leg_index(l - 1, m)] + B[k] * P[leg_index(l - 2, m)]);
            }
            P[leg_index(l, l - 1)] = cos_theta * (T)sqrt(2.0 * l + 1.0) * value;
            value *= -(T)sqrt(1.0 + 0.5 / l) * sin_theta;
            P[leg_index(l, l)] = value;
        }
    }
    const int L2 = L * L;
    if (grads) {
        DP[0] = 0;
        for (int l = 1; l < L; l++) {
            for (int m = 0; m <= l; m++) {
                const T pm1 = (m == 0) ? -(T)1 / (T)(l * l + l) * P[leg_index(l, 1)] : P[leg_index(l, m - 1)];
                const T pp1 = (m < l) ? P[leg_index(l, m + 1)] : (T)0;
                DP[leg_index(l, m)] = (T)sqrt((double)((l + m) * (l - m + 1))) * pm1 -
                                      (T)sqrt((double)((l - m) * (l + m + 1))) * pp1;
            }
        }
        for (int l = 0; l < L; l++) {
            for (int m = 0; m <= l; m++) {
                const int k = leg_index(l, m);
                PT[k] = (sin_theta > (T)0.1) ? (T)m / sin_theta * P[k] : -(T)0.5 / cos_theta * DP[k];
            }
        }
    }
    T* gx = grads;
    T* gy = grads ? grads + L2 : nullptr;
    T* gz = grads ? grads + 2 * L2 : nullptr;
    for (int l = 0; l < L; l++) Y[l * l + l] = P[leg_index(l, 0)] / SQRT_2;
    if (grads) {
        gx[0] = gy[0] = gz[0] = 0;
        for (int l = 1; l < L; l++) {
            const T lf = (T)sqrt(0.5 * (double)(l * (l + 1))) * P[leg_index(l, 1)];
            gx[l * l + l] = cos_phi * cos_theta * lf;
            gy[l * l + l] = sin_phi * cos_theta * lf;
            gz[l * l + l] = -sin_theta * lf;
        }
    }
    T cos_1 = 1, sin_1 = 0, cos_2 = -cos_phi, sin_2 = sin_phi;
    const T minus_two_cos = -2 * cos_phi;
    for (int m = 1; m < L; m++) {
        const T sin_m = minus_two_cos * sin_1 - sin_2;
        const T cos_m = minus_two_cos * cos_1 - cos_2;
        sin_2 = sin_1;
        sin_1 = sin_m;
        cos_2 = cos_1;
        cos_1 = cos_m;
        for (int l = m; l < L; l++) {
            const T p = P[leg_index(l, m)];
            Y[l * l + l + m] = p * cos_m;
            Y[l * l + l - m] = p * sin_m;
        }
        if (grads) {
            for (int l = m; l < L; l++) {
                const int k = leg_index(l, m);
                const T s_dp = sin_m * DP[k], c_dp = cos_m * DP[k];
                const T pt = PT[k];
                gx[l * l + l + m] = sin_phi * pt * sin_m - (T)0.5 * cos_theta * cos_phi * c_dp;
                gx[l * l + l - m] = -sin_phi * pt * cos_m - (T)0.5 * cos_theta * cos_phi * s_dp;
                gy[l * l + l + m] = -cos_phi * pt * sin_m - (T)0.5 * cos_theta * sin_phi * c_dp;
                gy[l * l + l - m] = cos_phi * pt * cos_m - (T)0.5 * cos_theta * sin_phi * s_dp;
                gz[l * l + l + m] = (T)0.5 * sin_theta * c_dp;
                gz[l * l + l - m] = (T)0.5 * sin_theta * s_dp;
            }
        }
    }
}

// Radial integral of channel l for neighbour entry e at distance r: values /
// grads of size n, from the spline tables or the per-entry exact values.
template <typename T>
__device__ void radial_integral(const DeviceRadial<T>& rad, int l, long long e, double r, T* values, T* grads) {
    const int n = rad.n;
    if (rad.use_spline) {
        const int K = rad.n_knots[l] - 1;
        const T* pos = rad.knot_pos + rad.knot_offset[l];
        // uniform grid: direct index, then adjust against the stored knot positions
        // so the interval is the one the reference's binary search picks (a
        // tabulated channel may have its own, non-uniform knots)
        int k = (int)((r - pos[0]) * K / ((double)pos[K] - pos[0]));
        if (k < 0) k = 0;
        if (k > K - 1) k = K - 1;
        while (k > 0 && (T)r < pos[k]) k--;
        while (k < K - 1 && (T)r >= pos[k + 1]) k++;
        const T x0 = pos[k], x1 = pos[k + 1];
        const T delta = x1 - x0;
        const T t = ((T)r - x0) / delta, t2 = t * t, t3 = t2 * t;
        const T h00 = 2 * t3 - 3 * t2 + 1, h10 = t3 - 2 * t2 + t, h01 = -2 * t3 + 3 * t2, h11 = t3 - t2;
        const T d00 = 6 * (t2 - t), d10 = 3 * t2 - 4 * t + 1, d01 = -d00, d11 = 3 * t2 - 2 * t;
        const T* v0 = rad.knot_val + (long long)(rad.knot_offset[l] + k) * n;
        const T* m0 = rad.knot_der + (long long)(rad.knot_offset[l] + k) * n;
        const T* v1 = v0 + n;
        const T* m1 = m0 + n;
        for (int i = 0; i < n; i++) {
            values[i] = h00 * v0[i] + h10 * delta * m0[i] + h01 * v1[i] + h11 * delta * m1[i];
            if (grads) grads[i] = d00 * v0[i] / delta + d10 * m0[i] + d01 * v1[i] / delta + d11 * m1[i];
        }
    } else {
        const long long base = (e * (rad.max_angular + 1) + l) * n;
        for (int i = 0; i < n; i++) {
            values[i] = rad.exact_val[base + i];
            if (grads) grads[i] = rad.exact_der[base + i];
        }
    }
}

// Cutoff smoothing x radial scaling and its derivative
// (featomic/src/calculators/soap/cutoff.rs:48-79, shared/density.rs:108-140,
//  soap/spherical_expansion_pair.rs:189-210).
__device__ inline void scaling_functions(const DeviceSoapScalars& s, double r, double* f, double* df) {
    double fc = 1.0, dfc = 0.0;
    if (s.smoothing == 0) {
        fc = r >= s.cutoff ? 0.0 : 1.0;
    } else {
        if (r <= s.cutoff - s.width) {
            fc = 1.0;
        } else if (r >= s.cutoff) {
            fc = 0.0;
        } else {
            const double x = RA_PI * (r - s.cutoff + s.width) / s.width;
            fc = 0.5 * (1.0 + cos(x));
            dfc = -0.5 * RA_PI * sin(x) / s.width;
        }
    }
    double sc = 1.0, dsc = 0.0;
    if (s.has_scaling) {
        sc = s.rate / (s.rate + pow(r / s.scale, s.exponent));
        const double rs = r / s.scale;
        const double rs_m1 = pow(rs, s.exponent - 1.0);
        const double rs_m = rs * rs_m1;
        dsc = -s.rate * s.exponent / s.scale * rs_m1 / ((s.rate + rs_m) * (s.rate + rs_m));
    }
    *f = sc * fc;
    *df = dfc * sc + fc * dsc;
}

}  // namespace rascaline_amd
