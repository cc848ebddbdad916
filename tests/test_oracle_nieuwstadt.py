"""Pins of the oracle's Nieuwstadt inflow profile (src/inflow.c:2732-2811) and of the complex gamma,
digamma and hypergeometric functions it is built on (src/include/inline.h:372-525), against facts
outside the oracle's code:
  - gamma: exact factorials at positive integers, the library gamma on the real axis and SciPy's
    complex gamma (both branches, Re z >= 0 and the reflection formula), 1e20 at the poles;
  - digamma: SciPy's complex psi on both branches, the recurrence psi(z+1) = psi(z) + 1/z;
  - 2F1: the closed forms 2F1(1,1;2;z) = -log(1-z)/z and 2F1(a,b;b;z) = (1-z)^-a for complex a,
    2F1(a,b;c;0) = 1, and Euler's integral by quadrature at the profile's own parameters;
  - the profile: above the inversion height it is the geostrophic wind Wg u_tau, whose digamma
    constant comes from SciPy; below it W = Wg - dW with dW assembled from SciPy's gamma and the
    quadrature 2F1; both then rotated by the reference's angle (its normalisation of V_r uses the
    already normalised U_r, kept).
"""
import cmath
import math

import numpy as np
import pytest
import scipy.special as sp
mpmath = pytest.importorskip("mpmath")

from oracle import oracle as O
from tosca_amd import case as C


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def test_gamma_integers_reals_and_complex():
    for n in range(1, 15):
        assert O.cgamma(float(n)) == complex(math.factorial(n - 1), 0.0)
    for x in np.linspace(0.05, 25.3, 61):
        if x != round(x):
            assert rel(O.cgamma(x), math.gamma(x)) < 2e-13, x
    for z in (1 + 1j, 0.3 + 2.5j, 4.7 - 1.2j, 1.55 + 0.8j, 3.1 + 0.77j, -0.4 + 0.3j, -2.7 - 1.1j, -5.5 + 0.5j):
        assert rel(O.cgamma(z), complex(sp.gamma(z))) < 1e-12, z
    for n in (0, -1, -4):
        assert O.cgamma(float(n)) == complex(1e20, 0.0)


def test_digamma_against_scipy_and_recurrence():
    for z in (1.0, 0.5, 2.5 + 0.5j, 1.05 + 0.9j, 0.2 + 0.1j, -1.3 + 0.6j, 7.0 - 3.0j, 0.45 + 0.4j):
        assert abs(O.cdigamma(z) - complex(sp.psi(z))) < 1e-12 * max(1.0, abs(sp.psi(z))), z
    for z in (0.7 + 0.3j, 2.2 - 1.4j, 0.3 + 0.05j):
        assert abs(O.cdigamma(z + 1) - (O.cdigamma(z) + 1 / z)) < 1e-12 * abs(O.cdigamma(z + 1))
    assert O.cdigamma(-2.0) == complex(1e20, 0.0)


def test_hypergeometric_closed_forms():
    for z in (-0.9, -0.3, 0.1, 0.5, 0.8, 0.95):
        assert rel(O.hypergeom(1, 1, 2, z), -math.log(1 - z) / z) < 1e-13, z
    for a, b in ((0.3 + 0.7j, 1.4 - 0.2j), (2.1 + 1.5j, 0.6 + 0.3j)):
        for z in (-0.5, 0.2, 0.6, 0.9):
            assert rel(O.hypergeom(a, b, b, z), (1 - z) ** (-a)) < 1e-12, (a, b, z)
    assert O.hypergeom(0.3 + 1j, 2.0, 1.7 - 0.2j, 0.0) == 1.0
    assert O.hypergeom(1.0, 2.0, -3.0, 0.5) == complex(1e20, 0.0)


def mp_2f1(a, b, c, z):
    with mpmath.workdps(40):
        return complex(mpmath.hyp2f1(mpmath.mpc(a), mpmath.mpc(b), mpmath.mpc(c), mpmath.mpf(z)))


INLET = dict(type=C.INLET_NIEUWSTADT, grnd_level=0.0, u_tau=0.42, roughness=0.05, h_inv=600.0, h_ref=90.0,
             fc=1.0e-4, udir=(math.cos(0.3), math.sin(0.3), 0.0))


def alpha_and_c(inlet):
    Cc = inlet["fc"] * inlet["h_inv"] / (0.4 * inlet["u_tau"])
    return 0.5 + 0.5 * cmath.sqrt(complex(1.0, 4.0 * Cc)), Cc


def test_hypergeometric_at_profile_parameters():
    alpha, _ = alpha_and_c(INLET)
    for z in (0.2, 0.55, 0.85, 0.99, 0.999):
        got = O.hypergeom(alpha + 1, alpha - 1, 2 * alpha, z)
        ref = mp_2f1(alpha + 1, alpha - 1, 2 * alpha, z)
        assert rel(got, ref) < 1e-12, (z, got, ref)


def w_independent(inlet, eta):
    """W(eta) from mpmath's gamma, digamma and 2F1."""
    alpha, Cc = alpha_and_c(inlet)
    with mpmath.workdps(40):
        a = mpmath.mpc(alpha)
        const = mpmath.digamma(a + 1) + mpmath.digamma(a - 1) - 2 * mpmath.digamma(1)
        wg = (mpmath.log(mpmath.mpf(inlet["h_inv"]) / inlet["roughness"]) - const) / mpmath.mpf("0.4")
        if eta >= 1.0:
            return complex(wg)
        dw = (-1j / mpmath.mpf("0.4")) * a ** 2 * mpmath.gamma(a) ** 2 / (Cc * mpmath.gamma(2 * a)) \
            * mpmath.power(1 - mpmath.mpf(eta), a - 1) * mpmath.hyp2f1(a + 1, a - 1, 2 * a, 1 - mpmath.mpf(eta))
        return complex(wg - dw)


def rotate(inlet, wp, wr):
    up, vp = wp.real * inlet["u_tau"], wp.imag * inlet["u_tau"]
    ur, vr = wr.real * inlet["u_tau"], wr.imag * inlet["u_tau"]
    ur = ur / math.sqrt(ur * ur + vr * vr)
    vr = vr / math.sqrt(ur * ur + vr * vr)
    th = math.acos(inlet["udir"][0] * ur + inlet["udir"][1] * vr)
    return np.array([math.cos(th) * up - math.sin(th) * vp, math.sin(th) * up + math.cos(th) * vp, 0.0])


@pytest.mark.parametrize("h", [700.0, 600.0, 420.0, 150.0, 35.0])
def test_profile_against_independent_special_functions(h):
    wr = w_independent(INLET, INLET["h_ref"] / INLET["h_inv"])
    wp = w_independent(INLET, min(h / INLET["h_inv"], 1.0))
    ref = rotate(INLET, wp, wr)
    got = O.nieuwstadt(INLET, h)
    assert np.abs(got - ref).max() < 1e-12 * np.abs(ref).max(), (h, got, ref)


def test_profile_shape():
    """Speed grows with height below the inversion, is constant above it, and the wind turns with
    height (veer); the ground-level ghost of the lowest cell is finite."""
    hs = [5.0, 20.0, 80.0, 200.0, 400.0, 599.0]
    u = np.array([O.nieuwstadt(INLET, h) for h in hs])
    speed = np.hypot(u[:, 0], u[:, 1])
    assert np.all(np.isfinite(u))
    assert np.all(np.diff(speed[:4]) > 0)
    above = [O.nieuwstadt(INLET, h) for h in (600.0, 800.0, 5000.0)]
    assert np.array_equal(above[0], above[1]) and np.array_equal(above[1], above[2])
    angle = np.arctan2(u[:, 1], u[:, 0])
    assert abs(angle[0] - angle[-1]) > 1e-3
