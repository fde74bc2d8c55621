"""Pins the oracle's stencils (deriv.f90 der_main / der2_main / derij_main)
against what the mathematics fixes: the exact symbols of the 6th-order
central operators on periodic sine waves, exactness on polynomials, and the
bidiagonal structure of the mixed derivative."""
import math

import numpy as np
import pytest

import oracle as O


def grid(n, L=2 * math.pi):
    h = L / n
    x = -L / 2 + h * np.arange(n)
    return h, x[None, None, :], x[None, :, None], x[:, None, None]


def k1_sym(k, h):
    # D1 sin(kx) = k1 cos(kx): (90 sin kh - 18 sin 2kh + 2 sin 3kh)/(60h)
    return (90 * math.sin(k * h) - 18 * math.sin(2 * k * h) + 2 * math.sin(3 * k * h)) / (60 * h)


def k2_sym(k, h):
    # D2 sin(kx) = -k2 sin(kx): (490 - 540 cos kh + 54 cos 2kh - 4 cos 3kh)/(180 h^2)
    return (490 - 540 * math.cos(k * h) + 54 * math.cos(2 * k * h) - 4 * math.cos(3 * k * h)) / (180 * h * h)


def kxy_sym(a, b, hx, hy):
    # bidiagonal Dxy sin(ax+by) = -kxy sin(ax+by) with
    # kxy = -(2/(720 hx hy)) sum_k c_k [cos(k(a hx + b hy)) - cos(k(a hx - b hy))]
    c = (270.0, -27.0, 2.0)
    s = sum(ck * (math.cos((q + 1) * (a * hx + b * hy)) - math.cos((q + 1) * (a * hx - b * hy)))
            for q, ck in enumerate(c))
    return -2 * s / (720 * hx * hy)


@pytest.mark.parametrize("n,k", [(16, 1), (16, 3), (24, 5), (32, 7)])
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_first_and_second_derivative_symbols(n, k, axis):
    h, X, Y, Z = grid(n)
    C = (X, Y, Z)[axis]
    p = O.make_params(n)
    fi = np.zeros((4, n, n, n))
    fi[1] = np.sin(k * C + 0.3)
    f = O.with_ghosts(fi)
    d1 = O.der(f, p, 1, axis)
    d2 = O.der2(f, p, 1, axis)
    np.testing.assert_allclose(d1, k1_sym(k, h) * np.cos(k * C + 0.3) * np.ones_like(d1), rtol=0, atol=1e-13 * k)
    np.testing.assert_allclose(d2, -k2_sym(k, h) * np.sin(k * C + 0.3) * np.ones_like(d2), rtol=0,
                               atol=1e-12 * k * k)
    # derivative along another axis of this field vanishes identically
    other = (axis + 1) % 3
    assert np.abs(O.der(f, p, 1, other)).max() < 1e-13


def test_symbols_converge_to_exact_derivatives():
    # 6th order: k1 -> k and k2 -> k^2 with error O((kh)^6)
    for n in (32, 64):
        h = 2 * math.pi / n
        assert abs(k1_sym(1, h) - 1) < 2 * (h ** 6)
        assert abs(k2_sym(1, h) - 1) < 2 * (h ** 6)


@pytest.mark.parametrize("pair", [(0, 1), (0, 2), (1, 2), (1, 0), (2, 0), (2, 1)])
def test_mixed_derivative_symbol(pair):
    n = 24
    h, X, Y, Z = grid(n)
    C = (X, Y, Z)
    a, b = 2, 3
    i, j = pair
    p = O.make_params(n)
    fi = np.zeros((4, n, n, n))
    fi[2] = np.sin(a * C[i] + b * C[j])
    f = O.with_ghosts(fi)
    dij = O.derij(f, p, 2, i, j)
    expect = -kxy_sym(a, b, h, h) * np.sin(a * C[i] + b * C[j]) * np.ones_like(dij)
    np.testing.assert_allclose(dij, expect, rtol=0, atol=1e-12)
    # the bidiagonal operator is not the tensor product D_i D_j: their symbols differ
    tensor = -k1_sym(a, h) * k1_sym(b, h)
    assert abs(kxy_sym(a, b, h, h) + tensor) > 1e-6


def _poly_field(n, h, fn):
    """f with ghosts holding the polynomial itself (no periodic fill)."""
    g = O.NGHOST
    c = h * (np.arange(n + 2 * g) - g)
    X, Y, Z = c[None, None, :], c[None, :, None], c[:, None, None]
    f = np.zeros((4, n + 2 * g, n + 2 * g, n + 2 * g))
    f[0] = fn(X, Y, Z)
    i = slice(g, g + n)
    return f, X[:, :, i], Y[:, i, :], Z[i, :, :]


def test_polynomial_exactness():
    n, h = 6, 0.1
    L = n * h
    p = O.make_params(n, L=L)
    # first derivative exact to degree 6
    f, X, Y, Z = _poly_field(n, h, lambda X, Y, Z: X ** 6 - 2 * X ** 5 + Y ** 4 * 0 + 3 * Z * X)
    np.testing.assert_allclose(O.der(f, p, 0, 0), 6 * X ** 5 - 10 * X ** 4 + 3 * Z + 0 * Y, atol=1e-10)
    # second derivative exact to degree 7
    f, X, Y, Z = _poly_field(n, h, lambda X, Y, Z: Y ** 7 + X * Y ** 2)
    np.testing.assert_allclose(O.der2(f, p, 0, 1), 42 * Y ** 5 + 2 * X + 0 * Z, atol=1e-9)
    # bidiagonal mixed derivative exact to total degree 7
    f, X, Y, Z = _poly_field(n, h, lambda X, Y, Z: X ** 3 * Z ** 3 + X * Z ** 5 + Y ** 7)
    np.testing.assert_allclose(O.derij(f, p, 0, 0, 2), 9 * X ** 2 * Z ** 2 + 5 * Z ** 4 + 0 * Y, atol=1e-9)
    # ... and not exact at degree 8 (x^3 z^5 has a nonzero O(h^6) error)
    f, X, Y, Z = _poly_field(n, h, lambda X, Y, Z: X ** 3 * Z ** 5)
    err = np.abs(O.derij(f, p, 0, 0, 2) - 15 * X ** 2 * Z ** 4).max()
    assert err > 1e-9


def test_periodic_ghost_fill_matches_wraparound():
    rng = np.random.default_rng(1)
    fi = rng.standard_normal((2, 5, 6, 7))
    f = O.with_ghosts(fi)
    g = O.NGHOST
    idx = lambda n: (np.arange(n + 2 * g) - g) % n
    np.testing.assert_array_equal(f, fi[:, idx(5)][:, :, idx(6)][:, :, :, idx(7)])
