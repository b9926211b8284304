"""The energy contour's quadratures and mesh (include/gfa.h: gfa_gauss_legendre, gfa_gauss_fermi_dirac,
gfa_energy_mesh) on the CPU.

Pins: (1) Gauss-Legendre rules equal numpy's (an independent Golub-Welsch implementation) to 2e-15 for every
point count up to 112, with the reference's rounding of the count; (2) the Fermi-Dirac rule integrates
x^k / (1 + exp(30 x)) on [-1, infinity) exactly for k <= 2n - 1, against scipy's adaptive quadrature; (3) both
rules and the energy meshes of five parameter sets equal the reference's own functions (golden fixture made
from the reference build, and the build itself when present) to 1e-13 (the 73-80 point table to 5e-14 in
the nodes and 2e-12 in the weights, its own accuracy), except the reference's Gauss-Legendre
tables above 80 points, which are not Gauss rules (their weights do not even sum to 2) and are not followed."""
import ctypes as C
import os

import numpy as np
import pytest
from scipy.integrate import quad

from angstromcube_amd import contour as K

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = np.load(os.path.join(HERE, "golden", "contour", "quadrature.npz"))
REF_SO = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "libref_green.so")


@pytest.mark.parametrize("n", list(range(1, 113)))
def test_gauss_legendre_is_the_gauss_rule(n):
    x, w = K.gauss_legendre(n)
    m = n if n <= 32 else (((n - 1) // 4 + 1) * 4 if n <= 64 else ((n - 1) // 8 + 1) * 8)
    assert len(x) == m
    xr, wr = np.polynomial.legendre.leggauss(m)
    assert np.abs(x - xr).max() < 2e-15 and np.abs(w - wr).max() < 1e-14
    assert np.all(np.diff(x) > 0)


def test_gauss_legendre_limits():
    assert K.gauss_legendre(0)[0].size == 0
    with pytest.raises(Exception, match="112"):
        K.gauss_legendre(113)


@pytest.mark.parametrize("n", list(range(1, 17)))
def test_fermi_dirac_rule_integrates_polynomials_exactly(n):
    x, w = K.gauss_fermi_dirac(n)
    f = lambda t, k: t ** k / (1 + np.exp(min(30 * t, 700)))  # noqa: E731
    for k in range(min(2 * n, 12)):
        exact = quad(f, -1, 0, args=(k,), epsabs=1e-16, epsrel=1e-13, limit=200)[0] + \
            quad(f, 0, 4, args=(k,), epsabs=1e-16, epsrel=1e-13, limit=200)[0]
        assert abs(np.sum(w * x ** k) - exact) <= 1e-12 * max(1.0, abs(exact)), (n, k)


def test_fermi_dirac_limits():
    assert K.gauss_fermi_dirac(0)[0].size == 0
    with pytest.raises(Exception, match="16"):
        K.gauss_fermi_dirac(17)


@pytest.mark.parametrize("n", list(range(1, 81)))
def test_gauss_legendre_equals_reference_tables(n):
    x, w = K.gauss_legendre(n)
    g = GOLD[f"gl_{n}"]
    assert g.shape[1] == len(x)
    assert np.abs(x - g[0]).max() < 5e-14 and np.abs(w - g[1]).max() < 2e-12


@pytest.mark.parametrize("n", list(range(1, 17)))
def test_fermi_dirac_equals_reference_tables(n):
    x, w = K.gauss_fermi_dirac(n)
    g = GOLD[f"fd_{n}"]
    assert g.shape[1] == len(x)
    assert np.abs(x - g[0]).max() < 1e-13 and np.abs(w - g[1]).max() < 1e-13


@pytest.mark.parametrize("i", range(5))
def test_energy_mesh_equals_reference(i):
    kBT, eBot, nBot, nPar, nFer, nPol = GOLD["meshes"][i]
    E, w = K.energy_mesh(kBT, eBot, int(nBot), int(nPar), int(nFer), int(nPol))
    g = GOLD[f"mesh_{i}"]
    assert len(E) == g.shape[0]
    assert np.abs(E - (g[:, 0] + 1j * g[:, 1])).max() < 1e-13
    assert np.abs(w - (g[:, 2] + 1j * g[:, 3])).max() < 1e-13


@pytest.mark.skipif(not os.path.exists(REF_SO), reason="reference build oracle/_ref absent")
def test_quadratures_equal_the_reference_build():
    R = C.CDLL(REF_SO)
    p = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    for n in (1, 7, 16, 33, 64, 80):
        xr, wr = np.zeros(128), np.zeros(128)
        m = R.ref_gauss_legendre(n, p(xr), p(wr))
        x, w = K.gauss_legendre(n)
        assert m == len(x) and np.abs(x - xr[:m]).max() < 5e-14 and np.abs(w - wr[:m]).max() < 2e-12
    for n in (1, 9, 16):
        xr, wr = np.zeros(128), np.zeros(128)
        m = R.ref_gauss_fermi_dirac(n, p(xr), p(wr))
        x, w = K.gauss_fermi_dirac(n)
        assert m == len(x) and np.abs(x - xr[:m]).max() < 1e-13 and np.abs(w - wr[:m]).max() < 1e-13


def test_default_mesh_is_a_closed_contour_rule():
    """The default open contour (bottom, parallel, Fermi-Dirac, 3 Matsubara poles) integrates a function analytic
    in the upper half plane and decaying fast enough times the Fermi function; a check on the weights: the bottom
    and top pieces are purely imaginary, the parallel and Fermi-Dirac pieces real, the poles -2 pi i kT."""
    E, w = K.energy_mesh()
    assert len(E) == 5 + 15 + 9 + 3
    assert np.all(w[:5].real == 0) and np.all(w[5:29].imag == 0)
    assert np.allclose(w[29:], -2j * np.pi * 1e-2)
    assert np.allclose(E[29:].imag, np.array([5, 3, 1]) * np.pi * 1e-2)
