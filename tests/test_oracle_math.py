"""Pin the FP64 oracle's building blocks with known answers the reference's unit
tests hold, closed forms from scipy, quadrature and finite differences."""

import ctypes
import math

import numpy as np
import pytest
import scipy.integrate
import scipy.special

import featomic_oracle as orc

L = orc.lib()

HYP1F1_Z = [0.0, 1.3, 42.0, 200.0, 660.0, -1.0, -40.0, -200.0, -500.0, 3e-1, 2e-4, -4e-1, -8e-4]
# mpmath values held by featomic/src/math/hyp1f1.rs:267-274, keyed by (n, l)
HYP1F1_REF = {
    (0, 0): [1.0, 3.66929666761924, 1.7392749415205e+18, 7.22597376812575e+86, 4.30881706558659e+286, 0.367879441171442, 4.24835425529159e-18, 1.38389652673674e-87, 7.12457640674129e-218, 1.349858807576, 1.00020002000133, 0.670320046035639, 0.999200319914684],
    (5, 6): [1.0, 3.38401720456627, 6.51794558745076e+17, 1.3083980191885e+86, 4.33929373469791e+285, 0.394775712369039, 7.11266970817605e-9, 8.39682564330044e-14, 1.36092469650155e-16, 1.32355791582708, 1.0001866842364, 0.688849034868493, 0.999253614360372],
    (13, 12): [1.0, 3.8421649506219, 3.54519102852153e+18, 2.89704505824636e+87, 3.07058123688245e+287, 0.354011483088197, -3.5964781936063e-14, -3.28638636385084e-24, -8.25083453006346e-30, 1.36478040300447, 1.00020742886482, 0.660319994010777, 0.999170713570911],
    (12, 15): [1.0, 3.27284968979622, 2.44964775706581e+17, 1.37335603673345e+85, 1.46589177252824e+284, 0.403871683240152, 4.2979422347514e-12, 1.71877602004918e-22, 1.88996873825758e-28, 1.31381868447175, 1.00018183480621, 0.695409738903967, 0.99927299318159],
}


@pytest.mark.parametrize("nl", list(HYP1F1_REF))
def test_hyp1f1_mpmath_table(nl):
    n, l = nl
    a, b = 0.5 * (n + l + 3), l + 1.5
    for z, expected in zip(HYP1F1_Z, HYP1F1_REF[nl]):
        assert L.orc_hyp1f1(a, b, z) == pytest.approx(expected, rel=1e-13)


def test_hyp1f1_scipy_gto_range():
    # arguments used by the GTO radial integrals (z up to ~300, negative for LODE)
    for n in range(12):
        for l in range(9):
            a, b = 0.5 * (n + l + 3), l + 1.5
            for z in [0.01, 0.5, 3.0, 17.0, 45.0, 120.0, 250.0, -0.5, -7.0, -60.0]:
                assert L.orc_hyp1f1(a, b, z) == pytest.approx(scipy.special.hyp1f1(a, b, z), rel=1e-9)


def test_gamma_known_values():
    # featomic/src/math/gamma.rs:330-360
    for x, e in [(1.000001e-5, 99999.32279432558), (-4.8, -0.06242336135475955), (-1.5, 2.363271801207355),
                 (0.1, 9.513507698668732), (1.5, 0.886226925452758), (2.5, 1.329340388179137),
                 (5.5, 52.34277778455352), (10.1, 454760.75144158595)]:
        assert L.orc_gamma(x) == pytest.approx(e, rel=1e-13)
    for x in np.linspace(0.6, 30, 40):
        assert L.orc_ln_gamma(x) == pytest.approx(math.lgamma(x), rel=1e-13, abs=1e-14)


def test_expi_known_values():
    # featomic/src/math/exp.rs:83-94
    for x, e in [(-42.0, -1.3377908810011776e-20), (-1.0, -0.21938393439552062), (1e-05, -10.935699800043697),
                 (0.33, -0.1720950921354428), (1.0, 1.8951178163559368), (2.5, 7.073765894578603),
                 (43.0, 1.1263482901669605e17)]:
        assert L.orc_expi(x) == pytest.approx(e, rel=1e-14)


def test_hyp2f1_scipy():
    for b in [1.5, 2.0, 3.5, 7.0]:
        for c in [1.5, 2.0, 3.0, 4.5]:
            for x in [0.1, 0.5, 0.8, 0.92]:
                assert L.orc_hyp2f1(1.0, b, c, x) == pytest.approx(scipy.special.hyp2f1(1.0, b, c, x), rel=1e-12)


def _real_sph_scipy(l, m, d):
    theta = math.acos(max(-1.0, min(1.0, d[2])))
    phi = math.atan2(d[1], d[0])
    # scipy's sph_harm(m, l, azimuth, polar) includes the Condon-Shortley phase; the
    # real harmonics of the reference cancel it (spherical_harmonics.rs:366-392)
    y = scipy.special.sph_harm_y(l, abs(m), theta, phi)
    cs = (-1) ** abs(m)
    if m > 0:
        return math.sqrt(2) * cs * y.real
    if m < 0:
        return math.sqrt(2) * cs * y.imag
    return y.real


def test_spherical_harmonics_closed_form():
    rng = np.random.default_rng(1)
    for _ in range(10):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        values, _ = orc.spherical_harmonics(8, d)
        for l in range(9):
            for m in range(-l, l + 1):
                assert values[l * l + l + m] == pytest.approx(_real_sph_scipy(l, m, d), rel=1e-10, abs=1e-12)


def test_spherical_harmonics_gradients_finite_differences():
    # featomic/src/math/spherical_harmonics.rs:519-589: r * grad Y vs finite differences
    rng = np.random.default_rng(2)
    # directions with 0 < sin(theta) < 0.1 are covered by the next test
    dirs = [np.array([1.0, 0, 0]), np.array([0, 0, 1.0]), np.array([0.1, 0.2, 0.9])] + [rng.normal(size=3) for _ in range(5)]
    for d in dirs:
        d = d / np.linalg.norm(d)
        values, grads = orc.spherical_harmonics(12, d)
        for x in range(3):
            delta = 1e-6
            dp = d.copy()
            dp[x] += delta
            vp, _ = orc.spherical_harmonics(12, dp / np.linalg.norm(dp))
            dm = d.copy()
            dm[x] -= delta
            vm, _ = orc.spherical_harmonics(12, dm / np.linalg.norm(dm))
            fd = (vp - vm) / (2 * delta)
            np.testing.assert_allclose(grads[x], fd, rtol=1e-5, atol=1e-6)


def test_spherical_harmonics_small_theta_branch_is_the_reference_formula():
    # For 0 < sin(theta) <= 0.1 the reference replaces m P_l^m / sin(theta) by
    # -1 / (2 cos(theta)) * Delta P_l^m with Delta P = a P^{m-1} - b P^{m+1}
    # (featomic/src/math/spherical_harmonics.rs:285-297). The exact identity has
    # a + sign, so in that cone the m != 0 gradients are not the derivative of the
    # values. Parity follows the reference: the oracle reproduces the formula, which
    # shows up as a jump of the gradient across sin(theta) = 0.1, while the values
    # stay continuous.
    s_hi, s_lo = 0.1 + 1e-9, 0.1 - 1e-9
    d_hi = np.array([0.6 * s_hi, 0.8 * s_hi, math.sqrt(1 - s_hi ** 2)])
    d_lo = np.array([0.6 * s_lo, 0.8 * s_lo, math.sqrt(1 - s_lo ** 2)])
    v_hi, g_hi = orc.spherical_harmonics(4, d_hi)
    v_lo, g_lo = orc.spherical_harmonics(4, d_lo)
    np.testing.assert_allclose(v_hi, v_lo, atol=1e-8)
    assert abs(g_hi[0][21] - g_lo[0][21]) > 1e-2   # l=4, m=1 jumps
    np.testing.assert_allclose(g_hi[2], g_lo[2], atol=1e-7)  # d/dz does not use the branch


def test_cutoff_known_values():
    # featomic/src/calculators/soap/cutoff.rs:97-112
    h = orc.SoapHypers(cutoff=4.0, smoothing=1, smoothing_width=0.5, gaussian_width=0.3, center_atom_weight=1.0,
                       max_angular=0, n_radial=1, spline_accuracy=-1.0)
    s = L.orc_soap_new(ctypes.byref(h))
    assert L.orc_cutoff_smoothing(s, 3.5) == 1.0
    assert L.orc_cutoff_smoothing(s, 3.8) == 0.34549150281252683
    assert L.orc_cutoff_smoothing(s, 4.0) == 0.0
    assert L.orc_cutoff_smoothing_gradient(s, 3.8) == -2.987832164741557
    assert L.orc_cutoff_smoothing_gradient(s, 5.0) == 0.0
    L.orc_soap_free(s)


def _gto_basis_funcs(radius, n):
    sig = [radius * max(math.sqrt(i), 1.0) / n for i in range(n)]
    return sig


@pytest.mark.parametrize("n_radial", [4, 8, 12])
def test_orthonormalization_by_quadrature(n_radial):
    # psi_j(r) = sum_n O[j, n] r^n exp(-r^2 / 2 sigma_n^2) must be orthonormal with weight r^2
    radius = 5.0
    rad = orc.Radial.soap(radius, 0.3, 0, n_radial)
    O = rad.orthonormalization()
    sig = _gto_basis_funcs(radius, n_radial)

    def psi(j, r):
        return sum(O[j, k] * r ** k * math.exp(-r * r / (2 * sig[k] ** 2)) for k in range(n_radial))

    for i in range(n_radial):
        for j in range(i, n_radial):
            val, _ = scipy.integrate.quad(lambda r: psi(i, r) * psi(j, r) * r * r, 0, 60, limit=400)
            assert val == pytest.approx(1.0 if i == j else 0.0, abs=1e-7)


@pytest.mark.parametrize("l", [0, 1, 3, 6])
def test_soap_gto_radial_by_quadrature(l):
    # I_nl(r) = 4 pi / (pi sigma^2)^{3/4} e^{-r^2/2s^2} int r'^2 R_nl(r') e^{-r'^2/2s^2} i_l(r r'/s^2) dr'
    # (featomic/src/calculators/soap/radial_integral/mod.rs:30-41)
    radius, width, n_radial = 5.0, 0.5, 6
    rad = orc.Radial.soap(radius, width, l, n_radial)
    O = rad.orthonormalization()
    sig = _gto_basis_funcs(radius, n_radial)
    for r in [0.3, 1.7, 3.9]:
        values, grads = rad.compute(l, r)
        for j in range(n_radial):
            def integrand(rp):
                R = sum(O[j, k] * rp ** k * math.exp(-rp * rp / (2 * sig[k] ** 2)) for k in range(n_radial))
                x = r * rp / width ** 2
                # i_l(x) e^{-(r - r')^2 / 2 s^2} combined to avoid overflow
                il = scipy.special.ive(l + 0.5, x) * math.sqrt(math.pi / (2 * x)) if x > 0 else (1.0 if l == 0 else 0.0)
                return rp * rp * R * il * math.exp(-(r - rp) ** 2 / (2 * width ** 2))
            val, _ = scipy.integrate.quad(integrand, 0, 40, limit=400, epsabs=1e-14)
            val *= 4 * math.pi / (math.pi * width ** 2) ** 0.75
            assert values[j] == pytest.approx(val, rel=1e-7, abs=1e-12)
        # derivative by finite differences (gto.rs:248-291)
        vp, _ = rad.compute(l, r + 1e-6)
        vm, _ = rad.compute(l, r - 1e-6)
        np.testing.assert_allclose(grads, (vp - vm) / 2e-6, rtol=1e-6, atol=1e-10)


def test_soap_radial_gradient_near_zero():
    # featomic/src/calculators/soap/radial_integral/gto.rs:248-265
    for l in range(4):
        rad = orc.Radial.soap(3.4, 0.5, l, 8)
        _, g0 = rad.compute(l, 0.0)
        _, g1 = rad.compute(l, 1e-12)
        np.testing.assert_allclose(g0, g1, rtol=1e-6, atol=1e-11)


def test_spline_accuracy_and_grid():
    # featomic/src/math/splines.rs:79-194: uniform grids of 10 * 2^k intervals per l,
    # spline values within the requested accuracy of the exact integral
    rad_exact = orc.Radial.soap(5.0, 0.3, 6, 8)
    rad = orc.Radial.soap(5.0, 0.3, 6, 8, spline_accuracy=1e-8)
    for l in range(7):
        pos, val, der = rad.knots(l)
        k = len(pos) - 1
        assert k >= 10 and (k // 10) & (k // 10 - 1) == 0 and k % 10 == 0
        np.testing.assert_allclose(np.diff(pos), 5.0 / k, rtol=1e-9)
        errs = []
        for r in np.linspace(0.01, 4.99, 97):
            v, g = rad.compute(l, r)
            e, _ = rad_exact.compute(l, r)
            errs.append(np.abs(v - e))
        assert np.mean(errs) < 1e-7
        # the spline is exact at the knots
        v, g = rad.compute(l, pos[3])
        e, eg = rad_exact.compute(l, pos[3])
        np.testing.assert_allclose(v, e, rtol=1e-13, atol=1e-16)
        np.testing.assert_allclose(g, eg, rtol=1e-12, atol=1e-14)


def test_k_vectors_counts():
    # featomic/src/math/k_vectors.rs:93-139
    sqrt3, sqrt5, eps = math.sqrt(3), math.sqrt(5), 1e-2
    recips = [np.eye(3), np.array([[1.0, 0, 0], [5.0, 1.0, 0], [3.0, 4.0, 1.0]]),
              np.array([[-0.10740572, -0.73747025, -0.66678455], [0.30874359, -3.40258247, -3.7851169],
                        [3.58349326, -1.68574423, -3.21198419]])]
    cutoffs = [1 + eps, math.sqrt(2) - eps, math.sqrt(2) + eps, sqrt3 - eps, sqrt3 + eps, 2 - eps, sqrt5 - eps]
    expected = [3, 3, 9, 9, 13, 13, 16]
    for rc in recips:
        cell = np.linalg.inv(rc / (2 * math.pi)).T
        for kc, e in zip(cutoffs, expected):
            kv = orc.k_vectors(cell, kc)
            assert len(kv) == e
            assert np.all(kv[:, 3] < kc)
