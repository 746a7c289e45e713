"""Shape-function spaces of the oracle pinned by closed forms.

HierPoly (hierarchical_basis_fns.rs:4-13): N_n = x^n, T_0 = 1-x, T_1 = 1+x,
T_n = x^n - 1 (even), x^n - x (odd).
KOLShapeFn (Legendre space of hierarchical_basis_fns.rs:201-647): N_n = L_n and, with
the Q-table identity sum_{k<n, k=n mod 2} (2k+1) L_k = L'_{n+1} - (2n+1) L_n = L'_{n-1},
T_n = c_n (L_n - L'_{n-1} / (n(n-1)/2)), which vanishes at +-1. The reference's order-10
row (denominator 40) breaks that; the oracle and the library use 45."""

import numpy as np
import numpy.polynomial.legendre as L
import pytest

import oracle

X = np.linspace(-1.0, 1.0, 23)
NORM = [0.968246, 2.561738, 0.838525, 4.248161, 0.816397, 5.882766, 0.808509]


def test_hierpoly_closed_form():
    n, nd, t, td = oracle.shape_tables(0, 12, X)
    for k in range(13):
        assert np.allclose(n[k], X ** k, rtol=0, atol=1e-15)
        assert np.allclose(nd[k], k * X ** (k - 1) if k else 0 * X, rtol=0, atol=1e-14)
        if k == 0:
            et, etd = 1 - X, -np.ones_like(X)
        elif k == 1:
            et, etd = 1 + X, np.ones_like(X)
        elif k % 2 == 0:
            et, etd = X ** k - 1, k * X ** (k - 1)
        else:
            et, etd = X ** k - X, k * X ** (k - 1) - 1
        assert np.allclose(t[k], et, rtol=0, atol=1e-14)
        assert np.allclose(td[k], etd, rtol=0, atol=1e-13)


def _leg(k, x, der=0):
    c = np.zeros(k + 1)
    c[k] = 1
    return L.legval(x, L.legder(c, der) if der else c)


@pytest.mark.parametrize("k", list(range(0, 21)))
def test_kol_closed_form(k):
    n, nd, t, td = oracle.shape_tables(1, 20, X)
    assert np.allclose(n[k], _leg(k, X), atol=1e-13)
    assert np.allclose(nd[k], _leg(k, X, 1), atol=1e-11)
    if k >= 2:
        c = NORM[k - 2] if k - 2 < len(NORM) else 1.0
        den = k * (k - 1) / 2
        et = c * (_leg(k, X) - _leg(k - 1, X, 1) / den)
        etd = c * (_leg(k, X, 1) - _leg(k - 1, X, 2) / den)
        assert np.allclose(t[k], et, atol=1e-12)
        assert np.allclose(td[k], etd, atol=1e-10)
        # vanishes at both endpoints (the order-10 row of the reference table does not)
        e = oracle.shape_tables(1, 20, np.array([-1.0, 1.0]))[2][k]
        assert np.max(np.abs(e)) < 1e-12


def test_reference_order10_row_is_inconsistent():
    """Q_NUMERATORS[8] / Q_DENOMINATORS[8] (hierarchical_basis_fns.rs:220, 225) sum to -5/40 at x=1."""
    nums = [-1, 0, -5, 0, -9, 0, -13, 0, -17, 0, 40]
    assert sum(nums) / 40 != 0  # T_10(1) != 0 with the published table
    assert sum(nums[:-1]) + 45 == 0  # the top coefficient that makes it vanish
