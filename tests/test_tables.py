"""Basis and quadrature tables: transcription, exactness, orthonormality and
an exact regeneration of the basis (CPU only)."""
import json
import math
import os
from decimal import Decimal, getcontext
from fractions import Fraction

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden", "ref_tables.json")
N_Q = {0: 1, 1: 3, 2: 6, 3: 12, 4: 16, 5: 25}
# measured orthonormality defects of the reference's tables, max |M - I|
ORTHO_BOUND = {0: 4.4e-16, 1: 1e-14, 2: 6.2e-14, 3: 7.8e-12, 4: 1.3e-9, 5: 1.9e-7}


@pytest.fixture(scope="module")
def ref():
    with open(GOLD) as f:
        return json.load(f)


def monomials(i):
    d = 0
    while (d + 1) * (d + 2) // 2 <= i:
        d += 1
    b = i - d * (d + 1) // 2
    return d - b, b


def phi_np(coeff, i, r, s):
    return sum(coeff[i][k] * r ** monomials(k)[0] * s ** monomials(k)[1] for k in range(i + 1))


def test_product_tables_are_the_reference_values(dg, ref):
    """r1, r2, w/2, s_r, w1 bit-identical to the reference's literals; basis
    table values equal the reference polynomials at the reference points."""
    coeff = ref["basis_coeff"]
    for p in range(6):
        t = dg.tables(p)
        q2 = np.array(ref["quad2d"][str(1 if p == 0 else 2 * p)]).reshape(-1, 3)
        assert len(q2) == N_Q[p]
        assert np.array_equal(t["r1"], q2[:, 0]) and np.array_equal(t["r2"], q2[:, 1])
        assert np.array_equal(t["w"], q2[:, 2] / 2.0)
        q1 = np.array(ref["quad1d"][str(p + 1)]).reshape(-1, 2)
        assert np.array_equal(t["s_r"], q1[:, 0]) and np.array_equal(t["w1"], q1[:, 1])
        np_ = t["n_p"]
        B = t["basis"].reshape(np_, -1)
        for i in range(np_):
            want = np.array([phi_np(coeff, i, r, s) for r, s in zip(q2[:, 0], q2[:, 1])])
            # cancellation in the monomial sum: bound by the sum of |terms|
            cond = sum(abs(c) for c in coeff[i])
            assert np.abs(B[i] - want).max() <= 4e-16 * cond, (p, i)


def test_oracle_tables_match_product(dg, orc):
    from helpers import mesh

    m = mesh(2)
    for p in range(6):
        t = dg.tables(p)
        o = orc.Oracle.from_mesh(m, p, dg.euler_vortex())
        to = o.tables()
        for k in ("r1", "r2", "w", "s_r", "w1"):
            assert np.array_equal(t[k], to[k]), (p, k)
        for k in ("basis", "gx", "gy", "side", "vertex"):
            assert np.array_equal(t[k], to[k]), (p, k)


def test_2d_rules_integrate_monomials(ref):
    """Rule of degree d integrates r^a s^b (a + b <= d) exactly:
    int_T r^a s^b = a! b! / (a + b + 2)!; the 15-digit literals bound the error."""
    for d in range(1, 11):
        q = np.array(ref["quad2d"][str(d)]).reshape(-1, 3)
        worst = 0.0
        for a in range(d + 1):
            for b in range(d + 1 - a):
                exact = math.factorial(a) * math.factorial(b) / math.factorial(a + b + 2)
                got = np.sum(q[:, 2] / 2 * q[:, 0] ** a * q[:, 1] ** b)
                worst = max(worst, abs(got - exact) / exact)
        assert worst < 5e-13, (d, worst)
        # one degree higher is not integrated exactly (the rule is of degree d)
        errs = []
        for a in range(d + 2):
            b = d + 1 - a
            exact = math.factorial(a) * math.factorial(b) / math.factorial(a + b + 2)
            errs.append(abs(np.sum(q[:, 2] / 2 * q[:, 0] ** a * q[:, 1] ** b) - exact) / exact)
        if d not in (1, 3, 5, 7, 9):  # odd-degree symmetric rules are exact one degree higher
            assert max(errs) > 1e-10, d


def test_1d_gauss_rules(ref):
    for n in range(1, 9):
        q = np.array(ref["quad1d"][str(n)]).reshape(-1, 2)
        tol = 1e-15 if n <= 5 else 2e-8  # 8-digit literals above 5 points
        for k in range(2 * n):
            exact = 2.0 / (k + 1) if k % 2 == 0 else 0.0
            assert abs(np.sum(q[:, 1] * q[:, 0] ** k) - exact) <= tol * 4, (n, k)
        # the points are the roots of P_n
        roots = np.sort(np.polynomial.legendre.leggauss(n)[0])
        assert np.abs(np.sort(q[:, 0]) - roots).max() <= tol * 10


def test_orthonormality_bounds(dg):
    for p in range(6):
        t = dg.tables(p)
        np_ = t["n_p"]
        B = t["basis"].reshape(np_, -1)
        M = (B * t["w"]) @ B.T
        err = np.abs(M - np.eye(np_)).max()
        assert err <= ORTHO_BOUND[p], (p, err)


def _gram_schmidt_exact(n):
    """Orthonormal basis of graded monomials on the reference triangle with
    inner product int_T f g: Cholesky of the exact Gram matrix in 60-digit
    decimal arithmetic.  Returns the coefficient matrix (row i = phi_i)."""
    getcontext().prec = 60
    mon = [monomials(i) for i in range(n)]

    def integ(a, b):
        return Fraction(math.factorial(a) * math.factorial(b), math.factorial(a + b + 2))

    G = [[integ(mon[i][0] + mon[j][0], mon[i][1] + mon[j][1]) for j in range(n)] for i in range(n)]
    Gd = [[Decimal(g.numerator) / Decimal(g.denominator) for g in row] for row in G]
    L = [[Decimal(0)] * n for _ in range(n)]
    for i in range(n):
        for j in range(i + 1):
            s = Gd[i][j] - sum(L[i][k] * L[j][k] for k in range(j))
            L[i][j] = s.sqrt() if i == j else s / L[j][j]
    # phi = L^{-1} m  (row i of L^{-1})
    Linv = [[Decimal(0)] * n for _ in range(n)]
    for i in range(n):
        Linv[i][i] = 1 / L[i][i]
        for j in range(i):
            Linv[i][j] = -sum(L[i][k] * Linv[k][j] for k in range(j, i)) / L[i][i]
    return [[float(x) for x in row] for row in Linv]


def test_basis_is_gram_schmidt_of_graded_monomials(ref):
    n = 21
    exact = np.array(_gram_schmidt_exact(n))
    coeff = np.array(ref["basis_coeff"])
    for i in range(15):
        scale = np.abs(exact[i]).max()
        rel = np.abs(coeff[i, : i + 1] - exact[i, : i + 1]).max() / scale
        assert rel <= 1e-9, (i, rel)
    # the high modes (p = 5) carry larger table errors but the same polynomials
    for i in range(15, 21):
        scale = np.abs(exact[i]).max()
        assert np.abs(coeff[i, : i + 1] - exact[i, : i + 1]).max() / scale <= 1e-5, i


def test_generated_table_file_matches_fixture(ref):
    """dg_ref_tables.inc holds exactly the fixture's numbers."""
    import re

    path = os.path.join(os.path.dirname(__file__), "..", "dgcuda_amd", "csrc", "dg_ref_tables.inc")
    txt = open(path).read()
    body = re.search(r"DG_BASIS_COEFF\[21\]\[21\] = \{(.*?)\n\};", txt, re.S).group(1)
    rows = [[float(x) for x in r.split(",")] for r in re.findall(r"\{([^{}]*)\}", body)]
    assert rows == ref["basis_coeff"]
