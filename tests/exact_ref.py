"""Exact-arithmetic evaluation of the reference's assembly for HierPoly, used to pin the oracle.

An independent reading of galerkin.rs:73-184, basis.rs:225-252/365-423 and
integrals.rs:26-92/292-354: for axis-aligned Elems every integrand is a polynomial
of degree <= 2*max_order in each direction, which n-point Gauss-Legendre
quadrature integrates exactly when n > max_order. The reference's sums therefore
equal exact integrals of these polynomials, computed here with fractions.Fraction.
Only small domains are feasible (pure Python)."""

from __future__ import annotations

from fractions import Fraction
from functools import lru_cache

import numpy as np

# polynomials: tuple of Fraction coefficients, index = power


def padd(a, b):
    n = max(len(a), len(b))
    return tuple((a[k] if k < len(a) else 0) + (b[k] if k < len(b) else 0) for k in range(n))


def pmul(a, b):
    out = [Fraction(0)] * (len(a) + len(b) - 1)
    for i, x in enumerate(a):
        if x:
            for j, y in enumerate(b):
                out[i + j] += x * y
    return tuple(out)


def pscale(a, s):
    return tuple(x * s for x in a)


def pcompose_affine(a, s, c):
    """a(s*x + c)"""
    out = (Fraction(0),)
    lin = (Fraction(c), Fraction(s))
    powk = (Fraction(1),)
    for coef in a:
        out = padd(out, pscale(powk, coef))
        powk = pmul(powk, lin)
    return out


def pint(a):
    """integral over [-1, 1]"""
    return sum((coef * Fraction(2, k + 1) for k, coef in enumerate(a) if k % 2 == 0), Fraction(0))


@lru_cache(maxsize=None)
def hier_poly(n):
    """(N, N', T, T') of HierPoly order n (hierarchical_basis_fns.rs:4-13)."""
    one = Fraction(1)
    N = tuple([Fraction(0)] * n + [one])
    Nd = tuple([Fraction(0)] * (n - 1) + [Fraction(n)]) if n > 0 else (Fraction(0),)
    if n == 0:
        T, Td = (one, -one), (-one,)
    elif n == 1:
        T, Td = (one, one), (one,)
    elif n % 2 == 0:
        T = padd(N, (-one,))
        Td = Nd
    else:
        T = padd(N, (Fraction(0), -one))
        Td = padd(Nd, (-one,))
    return N, Nd, T, Td


def sub_range(loc, r):
    (u0, u1), (v0, v1) = r
    mu, mv = (u0 + u1) / 2, (v0 + v1) / 2
    u, v = (u0, u1), (v0, v1)
    if loc in (0, 2, 4):
        u = (u0, mu)
    if loc in (1, 3, 5):
        u = (mu, u1)
    if loc in (0, 1, 6):
        v = (v0, mv)
    if loc in (2, 3, 7):
        v = (mv, v1)
    return (u, v)


class ExactAssembly:
    """Exact A (CurlCurl) and B (L2Inner) of a flattened domain (numpy arrays of Domain.arrays())."""

    def __init__(self, arrays):
        self.a = arrays
        self.parent = [int(p) for p in arrays["parent"]]
        self.children = [[] for _ in self.parent]
        for e, p in enumerate(self.parent):
            if p >= 0:
                self.children[p].append(e)

    def path(self, e):
        p = []
        while e >= 0:
            p.append(e)
            e = self.parent[e]
        return p[::-1]

    def range_from(self, e, anc):
        r = ((Fraction(-1), Fraction(1)), (Fraction(-1), Fraction(1)))
        p = self.path(e)
        k = p.index(anc) if anc is not None else 0
        for x in p[k + 1:]:
            r = sub_range(int(self.a["loc"][x]), r)
        return r

    def geometry(self, e):
        """(a, b) half extents of Elem e in real space (element.rs:33-50)."""
        (u0, u1), (v0, v1) = self.range_from(e, None)
        x0, y0, x3, y3 = (Fraction(float(v)) for v in self.a["element_box"][e])
        mr = lambda val, lo, hi: (val + 1) * (hi - lo) / 2 + lo  # noqa: E731
        a = (mr(u1, x0, x3) - mr(u0, x0, x3)) / 2
        b = (mr(v1, y0, y3) - mr(v0, y0, y3)) / 2
        return a, b

    def descendants(self, e):
        out = []
        for c in self.children[e]:
            out.append(c)
            out.extend(self.descendants(c))
        return out

    def specs(self, e):
        o0, o1 = int(self.a["bs_offsets"][e]), int(self.a["bs_offsets"][e + 1])
        return [(int(self.a["bs_dir"][t]), int(self.a["bs_i"][t]), int(self.a["bs_j"][t]), int(self.a["bs_dof"][t]))
                for t in range(o0, o1)]

    def pair(self, p, q, P, Q):
        """(CurlCurl, L2Inner) of basis spec p on Elem P sampled over Q and q on Q."""
        pd, pi, pj, _ = p
        qd, qi, qj, _ = q
        aP, bP = self.geometry(P)
        aQ, bQ = self.geometry(Q)
        if P == Q:
            su = sv = Fraction(1)
            cu = cv = Fraction(0)
        else:
            (u0, u1), (v0, v1) = self.range_from(Q, P)
            su, cu = (u1 - u0) / 2, (u1 + u0) / 2
            sv, cv = (v1 - v0) / 2, (v1 + v0) / 2
        eps = Fraction(float(self.a["materials"][P][0]))
        mu = Fraction(float(self.a["materials"][P][1]))
        detP, detQ = aP * bP, aQ * bQ
        # p's 1D factors at the mapped points, q's at the local points
        NP, NdP, TP, TdP = (pcompose_affine(f, su, cu) for f in hier_poly(pi))
        NPy, NdPy, TPy, TdPy = (pcompose_affine(f, sv, cv) for f in hier_poly(pj))
        NQ, NdQ, TQ, TdQ = hier_poly(qi)
        NQy, NdQy, TQy, TdQy = hier_poly(qj)
        # curl of a U spec: -(1/a) N_i(x) T'_j(y) * ps[0]; of a V spec: (1/b) T'_i(x) N_j(y) * ps[1]
        # p's derivative is scaled by q's para_scale (1, 1); q's by p's (su, sv)
        cc = Fraction(0)
        l2 = Fraction(0)
        if pd == 0:
            cp_x, cp_y, cp_s = NP, TdPy, -Fraction(1) / aP
        elif pd == 1:
            cp_x, cp_y, cp_s = TdP, NPy, Fraction(1) / bP
        if qd == 0:
            cq_x, cq_y, cq_s = NQ, TdQy, -Fraction(1) / aQ * su
        elif qd == 1:
            cq_x, cq_y, cq_s = TdQ, NQy, Fraction(1) / bQ * sv
        if pd in (0, 1) and qd in (0, 1):
            ratio = Fraction(1)
            if pd == 0 and qd == 0:
                ratio = (aP / bP) if detP >= detQ else (aQ / bQ)
            if pd == 1 and qd == 1:
                ratio = (bP / aP) if detP >= detQ else (bQ / aQ)
            cc = (1 / mu) * cp_s * cq_s * ratio * pint(pmul(cp_x, cq_x)) * pint(pmul(cp_y, cq_y))
            dmax = max(detP, detQ)
            if pd == 0 and qd == 0:
                l2 = eps * su * sv * (1 / aP) * (1 / aQ) * dmax * pint(pmul(NP, NQ)) * pint(pmul(TPy, TQy))
            elif pd == 1 and qd == 1:
                l2 = eps * su * sv * (1 / bP) * (1 / bQ) * dmax * pint(pmul(TP, TQ)) * pint(pmul(NPy, NQy))
        return cc, l2

    def assemble(self):
        """{(row, col): (A, B)} over the upper triangle, loop order of galerkin.rs:73-184."""
        out = {}

        def add(k, v):
            r, c = min(k), max(k)
            a0, b0 = out.get((r, c), (Fraction(0), Fraction(0)))
            out[(r, c)] = (a0 + v[0], b0 + v[1])

        for e in range(len(self.parent)):
            loc = self.specs(e)
            for i, p in enumerate(loc):
                for q in loc[i:]:
                    add((p[3], q[3]), self.pair(p, q, e, e))
            for d in self.descendants(e):
                for p in loc:
                    for q in self.specs(d):
                        add((p[3], q[3]), self.pair(p, q, e, d))
        return out


def compare_to_exact(res, arrays, rtol=1e-13):
    """Max relative deviation of an oracle/GPU result from the exact values (same floor as _compare)."""
    ex = ExactAssembly(arrays).assemble()
    rows = res.rows()
    keys = list(zip(rows.tolist(), res.col.tolist()))
    assert set(keys) == set(ex.keys())
    A = np.array([float(ex[k][0]) for k in keys])
    B = np.array([float(ex[k][1]) for k in keys])
    from _compare import entry_errors

    ea = entry_errors(rows, res.col, res.a_val, A)
    eb = entry_errors(rows, res.col, res.b_val, B)
    return float(ea.max()), float(eb.max())
