"""Series coefficients of the Chebyshev propagator store (host routine of the library, no GPU):
against scipy.special.ive, the normalisation e^y = I_0(y) + 2 sum I_k(y), the truncation
tolerance, and exp itself on the scalar spectrum [-rho, 0]."""
import ctypes

import numpy as np
import pytest
import scipy.special

from neuwon_amd import _lib


def _coefs(rho, tol=1e-10):
    lib = _lib.load()
    n = lib.nw_chebyshev_exp_coefficients(rho, tol, None, 0)
    assert n > 0
    c = np.empty(n, dtype=np.float64)
    assert lib.nw_chebyshev_exp_coefficients(rho, tol, _lib.ptr(c), n) == n
    return c


@pytest.mark.parametrize("rho", [0.0, 1e-3, 0.5, 12.0, 600.0, 2438.0, 10437.0, 40000.0])
def test_coefficients_against_scipy(rho):
    c = _coefs(rho)
    K = len(c) - 1
    want = 2.0 * scipy.special.ive(np.arange(K + 1), rho / 2.0)
    want[0] /= 2.0
    tail = 2.0 * scipy.special.ive(np.arange(K + 1, K + 4000), rho / 2.0).sum()
    want[0] += tail                    # the dropped terms are folded into c_0
    assert np.abs(c - want).max() <= 1e-14 * max(1.0, np.abs(want).max()) + 1e-300
    # truncation: the first dropped coefficient is below the tolerance, the last kept one above
    nxt = 2.0 * scipy.special.ive(K + 1, rho / 2.0)
    assert nxt <= 1e-10 and (K == 0 or c[-1] > 1e-10)
    # sum c_k T_k(1) = exp(0) = 1 exactly; sum c_k T_k(-1) = exp(-rho) within twice the tail
    assert c.sum() == pytest.approx(1.0, abs=1e-15)
    alt = (c * (-1.0) ** np.arange(K + 1)).sum()
    assert alt == pytest.approx(np.exp(-rho), abs=2 * tail + 1e-15)
    # exp(lambda) at points of [-rho, 0] (t = 2 lambda / rho + 1), Clenshaw against numpy
    if rho > 0:
        lam = -rho * np.linspace(0.0, 1.0, 41)
        t = 2.0 * lam / rho + 1.0
        series = np.polynomial.chebyshev.chebval(t, c)
        assert np.abs(series - np.exp(lam)).max() < 2 * tail + 1e-14
        assert tail < 3e-9


def test_bad_arguments_are_refused():
    lib = _lib.load()
    assert lib.nw_chebyshev_exp_coefficients(-1.0, 1e-10, None, 0) < 0
    assert lib.nw_chebyshev_exp_coefficients(1.0, 0.0, None, 0) < 0
    assert b"bad argument" in lib.nw_last_error()
