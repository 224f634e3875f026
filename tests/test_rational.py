"""The rational store's approximation of exp (host constants of the library, no GPU): its error
against exp on the whole negative axis, r(0) = 1, the poles' half plane, the generator that made
the constants, and the level-order tree elimination the kernels run, restated in numpy, against
scipy.linalg.expm of the tree operator."""
import numpy as np
import pytest
import scipy.linalg

from neuwon_amd import _lib, geometry, rational, synth
from oracle import neuwon_oracle as O


def _library_coefficients():
    out = np.empty(21, dtype=np.float64)
    assert _lib.load().nw_rational_coefficients(_lib.ptr(out), 21) == 21
    a0 = out[0]
    theta = out[1::4] + 1j * out[2::4]
    alpha = out[3::4] + 1j * out[4::4]
    return a0, alpha, theta


def test_error_on_the_negative_axis():
    a0, alpha, theta = _library_coefficients()
    x = -np.concatenate([np.linspace(0.0, 60.0, 300001), np.logspace(-10, 9, 300001)])
    err = np.abs(rational.evaluate(x, a0, alpha, theta) - np.exp(x))
    assert err.max() < 3e-10
    assert rational.evaluate(0.0, a0, alpha, theta) == pytest.approx(1.0, abs=1e-15)
    assert np.all(theta.imag > 1.0)
    # negative controls: a dropped pole (even the one of the smallest residue), or a residue
    # with the sign of its imaginary part flipped, misses the bound by orders of magnitude
    for j in range(5):
        keep = np.arange(5) != j
        assert np.abs(rational.evaluate(x, a0, alpha[keep], theta[keep]) - np.exp(x)).max() > 1e-4
    bad = alpha.copy()
    bad[2] = bad[2].conjugate()
    assert np.abs(rational.evaluate(x, a0, bad, theta) - np.exp(x)).max() > 1e-4


def test_generator_reproduces_the_library_constants():
    a0, alpha, theta = _library_coefficients()
    b0, beta, phi = rational.coefficients()
    assert np.allclose(phi, theta, rtol=1e-7, atol=0)
    assert np.allclose(beta, alpha, rtol=1e-6, atol=1e-9)
    assert abs(b0 - a0) < 1e-9


def _level_order_product(parent, B, x, a0, alpha, theta):
    """exp(B) x by the kernels' algorithm: rows in level order, per pole q_i = (x_i -
    sum_children B[i][c] q_c) / d_i leaves to root, z_i = q_i - B[i][parent] / d_i z_parent root to
    leaves; out = a0 x + 2 Re sum alpha_j z_j.  Pivots from the FP32-rounded edge coefficients,
    zero row sums."""
    n = len(parent)
    kids = [[] for _ in range(n)]
    for c in range(1, n):
        kids[parent[c]].append(c)
    order, levels = [0], [0, 1]
    while levels[-1] > levels[-2]:
        for k in range(levels[-2], levels[-1]):
            order += kids[order[k]]
        levels.append(len(order))
    levels = levels[:-1]
    pos = np.empty(n, int)
    pos[order] = np.arange(n)
    up = np.array([np.float32(B[parent[i], i]) if i else 0.0 for i in order], np.float64)
    dn = np.array([np.float32(B[i, parent[i]]) if i else 0.0 for i in order], np.float64)
    ch = [[pos[c] for c in kids[i]] for i in order]
    par = [pos[parent[i]] if i else -1 for i in order]
    diag = np.array([-dn[k] - sum(up[c] for c in ch[k]) for k in range(n)])
    xs = x[order]
    out = a0 * xs
    for a, th in zip(alpha, theta):
        inv = np.zeros(n, complex)
        for k in range(n - 1, -1, -1):
            inv[k] = 1.0 / (diag[k] - th - sum(up[c] * dn[c] * inv[c] for c in ch[k]))
        q = np.zeros(n, complex)
        for L in range(len(levels) - 2, -1, -1):
            for k in range(levels[L], levels[L + 1]):
                q[k] = (xs[k] - sum(up[c] * q[c] for c in ch[k])) * inv[k]
        z = np.zeros(n, complex)
        for L in range(len(levels) - 1):
            for k in range(levels[L], levels[L + 1]):
                z[k] = q[k] - (dn[k] * inv[k] * z[par[k]] if par[k] >= 0 else 0.0)
        out = out + 2.0 * (a * z).real
    res = np.empty(n)
    res[order] = out
    return res


@pytest.mark.parametrize("n,seed", [(1, 0), (2, 1), (40, 2), (300, 3)])
def test_level_order_solve_against_expm(n, seed):
    a0, alpha, theta = _library_coefficients()
    rng = np.random.default_rng(seed)
    cell = synth.random_cell(n, rng)
    parent = np.asarray(cell.parent, np.int64)
    geo = geometry.geometry(parent, cell.length, cell.diam)
    R, C = geometry.passive(geo["length"], geo["cross_sectional_area"], geo["area_at_creation"])
    B = O.coefficient_matrix(parent, R, C, 0.05).toarray()
    x = rng.uniform(-80.0, 40.0, n)
    got = _level_order_product(parent, B, x, a0, alpha, theta)
    # the FP32-rounded edges perturb B by ~6e-8 relative; the operator with those edges:
    Bq = B.copy()
    for i in range(1, n):
        Bq[parent[i], i] = np.float32(B[parent[i], i])
        Bq[i, parent[i]] = np.float32(B[i, parent[i]])
    np.fill_diagonal(Bq, 0.0)
    np.fill_diagonal(Bq, -Bq.sum(axis=1))
    want = scipy.linalg.expm(Bq) @ x
    assert np.abs(got - want).max() < 1e-9 * np.abs(x).max()
    # constants are kept: exp(B) 1 = 1
    ones = _level_order_product(parent, B, np.ones(n), a0, alpha, theta)
    assert np.abs(ones - 1.0).max() < 1e-13
    # a transposed operator is a different propagator (catches swapped edge directions)
    if n > 2:
        assert np.abs(scipy.linalg.expm(Bq.T) @ x - want).max() > 1e-6
