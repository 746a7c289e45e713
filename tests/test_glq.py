"""Gauss-Legendre quadrature: the oracle (Golub-Welsch, glq.rs:179-222) and the library
(Newton on the Legendre recurrence) pinned by the reference's 20-point table and by
the exactness of Gauss quadrature."""

import json
import math

import numpy as np
import pytest

import fem_2d_amd as F
import oracle
from conftest import golden

TABLE = json.load(open(golden("glq_20.json")))


@pytest.mark.parametrize("impl", ["oracle", "library"])
def test_glq_table_20(impl):
    """glq.rs:325-343: nodes, weights and scaled nodes at 1e-9."""
    if impl == "oracle":
        x, w = oracle.gauss_quadrature_points(20)
    else:
        x, w = F.gauss_quadrature_points(20)
    acc = TABLE["accuracy"]
    assert np.max(np.abs(x - TABLE["x_20"])) < acc
    assert np.max(np.abs(w - TABLE["w_20"])) < acc
    # scale_gauss_quad_points(points, 0.25, 0.5) (glq.rs:238-249)
    s, c = (0.5 - 0.25) / 2.0, (0.5 + 0.25) / 2.0
    assert abs(s - 0.125) < 1e-14
    assert np.max(np.abs(x * s + c - TABLE["x_20_scaled_0p25_0p5"])) < acc


def test_glq_with_endpoints():
    """glq.rs:345-374: endpoints prepended/appended with weight 1."""
    x, w = oracle.gauss_quadrature_points(20, include_endpoints=True)
    assert len(x) == 22 and abs(x[0] + 1) < 1e-14 and abs(x[-1] - 1) < 1e-14
    assert abs(w[0] - 1) < 1e-14 and abs(w[-1] - 1) < 1e-14
    assert np.max(np.abs(x[1:21] - TABLE["x_20"])) < 1e-9


@pytest.mark.parametrize("n", [1, 2, 4, 5, 8, 10, 12, 16, 20, 32, 64])
def test_oracle_and_library_nodes_agree(n):
    xo, wo = oracle.gauss_quadrature_points(n)
    xl, wl = F.gauss_quadrature_points(n)
    assert np.max(np.abs(xo - xl)) < 1e-14
    assert np.max(np.abs(wo - wl)) < 1e-14
    assert np.all(np.diff(xl) > 0)
    assert abs(np.sum(xl)) < 1e-13  # glq.rs:167 (sum of symmetric nodes)
    assert abs(np.sum(wl) - 2.0) < 1e-13


@pytest.mark.parametrize("n", [4, 8, 10, 12])
def test_exact_for_degree_2n_minus_1(n):
    """Gauss-Legendre with n points integrates x^k exactly for k <= 2n-1."""
    for impl in (oracle.gauss_quadrature_points, F.gauss_quadrature_points):
        x, w = impl(n)
        for k in range(2 * n):
            exact = 0.0 if k % 2 else 2.0 / (k + 1)
            assert abs(np.dot(w, x ** k) - exact) < 1e-13
        # degree 2n is not integrated exactly (guards against a wrong node count)
        assert abs(np.dot(w, x ** (2 * n)) - 2.0 / (2 * n + 1)) > 1e-12


def test_tensor_integral_u2v2():
    """glq.rs:4-17: the 2D tensor sum of u^2 v^2 over (-1,1)^2 is 4/9."""
    x, w = oracle.gauss_quadrature_points(10)
    sol = 0.0
    for m in range(10):
        inner = 0.0
        for n in range(10):
            inner += x[m] ** 2 * x[n] ** 2 * w[n]
        sol += inner * w[m]
    assert abs(sol - 4.0 / 9.0) < 1e-12


@pytest.mark.parametrize("order,expected", [(1, 4), (2, 8), (3, 16), (4, 16), (5, 32), (6, 32), (8, 32), (9, 64)])
def test_default_ngq(order, expected):
    """basis.rs:171-177: 4 * order rounded up to a power of two."""
    assert oracle.default_ngq(order) == expected
    assert 2 ** math.ceil(math.log2(4 * order)) == expected
