"""Pins the oracle's building blocks against independent references (CPU only).

* Gauss-Legendre: numpy.polynomial.legendre.leggauss, and YETI's own 30-digit table through the
  compiled reference (main/GaussLegendre.f90:27-160), orders 1..9, volume and face orderings.
* B-spline values / derivatives: scipy.interpolate.BSpline (an independent library
  implementation), partition of unity, and YETI's dersbasisfuns (main/dersbasisfuns.f90:30).
"""
import numpy as np
import pytest
from scipy.interpolate import BSpline

from oracle import oracle as O


@pytest.mark.parametrize("n", range(1, 10))
def test_gauss_matches_numpy(n):
    x, w = O.oracle_gauss(n)
    xr, wr = np.polynomial.legendre.leggauss(n)
    assert np.allclose(x, xr, rtol=0, atol=2e-16 * 4)
    assert np.allclose(w, wr, rtol=0, atol=2e-16 * 8)
    assert abs(w.sum() - 2.0) < 1e-15


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built")
@pytest.mark.parametrize("n", range(1, 10))
def test_gauss_matches_reference_table(n):
    x, w = O.oracle_gauss(n)
    g = O.ref_gauss(n, dim=1)
    assert np.allclose(g[1], x, rtol=0, atol=1e-15)
    assert np.allclose(g[0], w, rtol=0, atol=1e-15)
    # 3D ordering: xi slowest, zeta fastest (GaussLegendre.f90:175-186)
    g3 = O.ref_gauss(n, dim=3)
    k = 0
    for i in range(n):
        for j in range(n):
            for l in range(n):
                assert np.allclose(g3[1:, k], [x[i], x[j], x[l]], atol=1e-15)
                assert abs(g3[0, k] - w[i] * w[j] * w[l]) < 1e-15
                k += 1


def _knots(p, nel, seed):
    rng = np.random.default_rng(seed)
    inner = np.sort(rng.uniform(0, 1, nel - 1))
    return np.concatenate([np.zeros(p + 1), inner, np.ones(p + 1)])


@pytest.mark.parametrize("p", [1, 2, 3, 4, 5])
def test_bspline_matches_scipy(p):
    U = _knots(p, 7, p)
    ncp = U.size - p - 1
    rng = np.random.default_rng(p + 10)
    for span0 in range(p, ncp):
        if U[span0] == U[span0 + 1]:
            continue
        for u in rng.uniform(U[span0], U[span0 + 1], 5):
            N, dN = O.oracle_dersbasisfuns(span0 + 1, p, u, U)
            for r in range(p + 1):
                c = np.zeros(ncp)
                c[span0 - p + r] = 1.0
                s = BSpline(U, c, p)
                assert abs(N[r] - s(u)) < 1e-14
                assert abs(dN[r] - s.derivative()(u)) < 1e-11 * max(1.0, abs(dN[r]))
            assert abs(N.sum() - 1.0) < 1e-14
            assert abs(dN.sum()) < 1e-10


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built")
@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_bspline_matches_reference(p):
    U = _knots(p, 6, 100 + p)
    ncp = U.size - p - 1
    rng = np.random.default_rng(p)
    for span0 in range(p, ncp):
        if U[span0] == U[span0 + 1]:
            continue
        u = rng.uniform(U[span0], U[span0 + 1])
        N, dN = O.oracle_dersbasisfuns(span0 + 1, p, u, U)
        Nr, dNr = O.ref_dersbasisfuns(span0 + 1, p, u, U)
        assert np.allclose(N, Nr, rtol=0, atol=1e-15)
        assert np.allclose(dN, dNr, rtol=1e-13, atol=1e-12)
