"""Pins of the oracle's 1D finite-element ingredients (closed forms, exactness)."""
import numpy as np
import pytest

import oracle as O


def test_gll_closed_forms():
    assert np.allclose(O.gll_points(1), [0, 1], atol=1e-15)
    assert np.allclose(O.gll_points(2), [0, 0.5, 1], atol=1e-15)
    s5 = np.sqrt(5.0)
    assert np.allclose(O.gll_points(3), [0, 0.5 - 0.5 / s5, 0.5 + 0.5 / s5, 1], atol=1e-15)
    s37 = np.sqrt(3.0 / 7.0)
    assert np.allclose(O.gll_points(4), [0, 0.5 - 0.5 * s37, 0.5, 0.5 + 0.5 * s37, 1], atol=1e-15)


@pytest.mark.parametrize("p", range(2, 8))
def test_gll_are_roots_of_legendre_derivative(p):
    x = O.gll_points(p)
    t = 2 * x[1:-1] - 1
    dP = np.polynomial.legendre.Legendre.basis(p).deriv()(t)
    assert np.abs(dP).max() < 1e-12
    assert np.all(np.diff(x) > 0)


def test_gauss_closed_forms():
    x, w = O.gauss_rule(2)
    assert np.allclose(x, [0.5 - 0.5 / np.sqrt(3), 0.5 + 0.5 / np.sqrt(3)], atol=1e-15)
    assert np.allclose(w, [0.5, 0.5], atol=1e-15)
    x, w = O.gauss_rule(3)
    assert np.allclose(x, [0.5 - 0.5 * np.sqrt(0.6), 0.5, 0.5 + 0.5 * np.sqrt(0.6)], atol=1e-15)
    assert np.allclose(w, [5 / 18, 8 / 18, 5 / 18], atol=1e-15)


@pytest.mark.parametrize("n", range(1, 9))
def test_gauss_exactness(n):
    x, w = O.gauss_rule(n)
    for k in range(2 * n):
        assert abs(np.dot(w, x ** k) - 1.0 / (k + 1)) < 1e-14
    # and not exact one degree higher (the rule really has n points)
    assert abs(np.dot(w, x ** (2 * n)) - 1.0 / (2 * n + 1)) > 1e-12


@pytest.mark.parametrize("p", range(1, 7))
def test_lagrange_partition_unity_and_derivative(p):
    z = O.gll_points(p)
    x = np.linspace(0.013, 0.987, 11)
    val, der = O.lagrange(z, x)
    assert np.allclose(val.sum(axis=1), 1.0, atol=1e-13)
    assert np.allclose(der.sum(axis=1), 0.0, atol=1e-11)
    # reproduction of x^p and its derivative
    assert np.allclose(val @ z ** p, x ** p, atol=1e-13)
    assert np.allclose(der @ z ** p, p * x ** (p - 1), atol=1e-11)
    # derivative vs central differences
    h = 1e-6
    vp, _ = O.lagrange(z, x + h)
    vm, _ = O.lagrange(z, x - h)
    assert np.allclose((vp - vm) / (2 * h), der, atol=1e-7)
    # Kronecker property at the nodes
    vn, _ = O.lagrange(z, z)
    assert np.allclose(vn, np.eye(p + 1), atol=1e-14)
