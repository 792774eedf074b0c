"""Pins for the oracle's cubic B-spline (MKernelBSpline.cpp:98-187).

Checks against the mathematics, not a retyping of the formula:
  * normalisation: the lattice sum of W*dV is 1 in 1D/2D/3D (up to the
    reference's truncated pi, global.h:12, a -1.3e-10 relative effect);
  * the gradient is the spatial derivative of W (central differences);
  * C1 continuity at z = 1 and z = 2, compact support, zero at r = 0;
  * the gradient points from j to i (repulsive sign convention).
A wrong normalisation constant, a sign flip or a dropped term fails one of these.
"""
import math

import numpy as np
import pytest

from oracle import oracle as orc

PI_REF = 3.141592654


def lattice_sum(ndim, h, dx):
    n = int(math.ceil(2 * h / dx)) + 1
    axes = [np.arange(-n, n + 1) * dx] * ndim
    grids = np.meshgrid(*axes, indexing="ij")
    pts = np.stack([g.ravel() for g in grids], 1)
    xi = np.zeros(3)
    s = 0.0
    for p in pts:
        xj = np.zeros(3)
        xj[:ndim] = p
        s += orc.kernel_w(xi, xj, ndim, h)
    return s * dx ** ndim


@pytest.mark.parametrize("ndim,h", [(1, 0.7), (2, 1.3), (3, 1.6)])
def test_partition_of_unity(ndim, h):
    s = lattice_sum(ndim, h, 0.1)
    # lattice sums of the cubic spline converge fast; the truncated pi shifts
    # the result by pi/PI_REF - 1 = -1.3e-10
    assert abs(s - 1.0) < 2e-4, s


def test_truncated_pi_normalisation_2d():
    # W(0) = 10/(7*PI*h^2) with PI = 3.141592654 exactly as the reference
    h = 0.37
    w0 = orc.kernel_w(np.zeros(3), np.zeros(3), 2, h)
    assert w0 == 10.0 / (7.0 * PI_REF * (h * h))
    assert w0 != 10.0 / (7.0 * math.pi * h * h)


@pytest.mark.parametrize("ndim", [1, 2, 3])
def test_gradient_is_derivative(ndim):
    rng = np.random.default_rng(1)
    h = 0.9
    for _ in range(200):
        xi = np.zeros(3)
        xj = np.zeros(3)
        xi[:ndim] = rng.uniform(-2 * h, 2 * h, ndim)
        xj[:ndim] = rng.uniform(-2 * h, 2 * h, ndim)
        r = np.linalg.norm(xi - xj)
        if r < 1e-3 or abs(r / h - 1) < 1e-3 or r / h > 1.999:
            continue
        g = orc.kernel_gradw(xi, xj, ndim, h)
        eps = 1e-6
        for d in range(ndim):
            e = np.zeros(3)
            e[d] = eps
            fd = (orc.kernel_w(xi + e, xj, ndim, h) - orc.kernel_w(xi - e, xj, ndim, h)) / (2 * eps)
            assert abs(g[d] - fd) <= 1e-6 * max(1.0, abs(fd)), (ndim, r / h, d, g[d], fd)
        assert np.all(g[ndim:] == 0.0)


def test_continuity_and_support():
    h = 1.0
    xi = np.zeros(3)
    for z in (1.0, 2.0):
        lo = np.array([z - 1e-9, 0, 0])
        hi = np.array([z + 1e-9, 0, 0])
        assert abs(orc.kernel_w(xi, lo, 2, h) - orc.kernel_w(xi, hi, 2, h)) < 1e-8
        assert np.abs(orc.kernel_gradw(xi, lo, 2, h) - orc.kernel_gradw(xi, hi, 2, h)).max() < 1e-8
    far = np.array([2.0000001, 0, 0])
    assert orc.kernel_w(xi, far, 2, h) == 0.0
    assert np.all(orc.kernel_gradw(xi, far, 2, h) == 0.0)
    assert np.all(orc.kernel_gradw(xi, xi, 3, h) == 0.0)


def test_gradient_direction():
    # dW/dr < 0 inside the support: grad_i W points from i towards j
    xi = np.array([0.0, 0.0, 0.0])
    xj = np.array([0.5, 0.2, 0.0])
    g = orc.kernel_gradw(xi, xj, 2, 1.0)
    assert np.dot(g[:2], (xj - xi)[:2]) > 0


@pytest.mark.parametrize("ndim", [1, 2, 3])
def test_dwdr_is_radial_derivative(ndim):
    """dWdr of GradW (MKernelBSpline.cpp:174-185) is dW/dr: central
    differences of W along the separation, and |grad W| = |dW/dr|."""
    rng = np.random.default_rng(2)
    h = 0.8
    xi = np.zeros(3)
    for _ in range(100):
        r = rng.uniform(0.0, 2.2 * h)
        if r < 1e-3 or abs(r / h - 1) < 1e-3 or abs(r / h - 2) < 1e-3:
            continue
        u = np.zeros(3)
        u[:ndim] = rng.normal(size=ndim)
        u /= np.linalg.norm(u)
        xj = r * u
        dr = orc.kernel_dwdr(xi, xj, ndim, h)
        eps = 1e-6
        fd = (orc.kernel_w(xi, (r + eps) * u, ndim, h) - orc.kernel_w(xi, (r - eps) * u, ndim, h)) / (2 * eps)
        assert abs(dr - fd) <= 1e-6 * max(1.0, abs(fd)), (ndim, r / h, dr, fd)
        g = orc.kernel_gradw(xi, xj, ndim, h)
        assert abs(np.linalg.norm(g) - abs(dr)) <= 1e-12 * max(1.0, abs(dr))
    assert orc.kernel_dwdr(xi, xi, ndim, h) == 0.0
