"""Rational approximation of exp on the negative real axis, the series of the rational store.

    exp(x) ~ r(x) = a0 + 2 Re sum_j alpha_j / (x - theta_j),   x <= 0,

with five poles theta_j in the upper half plane (type (10, 10)).  The poles come from the AAA
algorithm (Nakatsukasa, Sete and Trefethen, SIAM J. Sci. Comput. 40 (2018)) run on
F(t) = exp(s (t - 1) / (t + 1)), t in [-1, 1) (s = 9), the Moebius transplant of (-inf, 0] used by
Trefethen, Weideman and Schmelzer (BIT 46 (2006)) for the Caratheodory-Fejer approximants; the
residues are the real least-squares fit of exp on 20,000 transplanted Chebyshev points, and a0 is
then moved so that r(0) = 1 exactly (constant vectors are kept, as exp(B) 1 = 1 for a generator
with zero row sums).  Maximum error on (-inf, 0]: 3e-10 (pinned by the tests).

The library carries the printed constants (nw_rational_coefficients); this module regenerates them.
"""
import numpy as np

POLES = 5
TRANSPLANT = 9.0


def _aaa(Z, F, m):
    """AAA with m support points; returns the poles of the barycentric approximant."""
    J = np.ones(len(Z), bool)
    z, f = [], []
    R = np.full(len(Z), F.mean())
    for _ in range(m):
        j = int(np.argmax(np.abs(F - R) * J))
        z.append(Z[j])
        f.append(F[j])
        J[j] = False
        zz, ff = np.array(z), np.array(f)
        with np.errstate(divide="ignore", invalid="ignore"):
            C = 1.0 / (Z[:, None] - zz[None, :])
        C[~J, :] = 0.0
        A = (F[:, None] - ff[None, :]) * C
        w = np.linalg.svd(A[J], full_matrices=False)[2][-1].conj()
        R = F.copy()
        R[J] = (C[J] @ (w * ff)) / (C[J] @ w)
    zz = np.array(z)
    E = np.zeros((m + 1, m + 1), complex)
    E[0, 1:] = w
    E[1:, 0] = 1.0
    E[1:, 1:] = np.diag(zz)
    B = np.eye(m + 1)
    B[0, 0] = 0.0
    import scipy.linalg
    pol = scipy.linalg.eigvals(E, B)
    return pol[np.isfinite(pol)]


def coefficients(poles=POLES, s=TRANSPLANT):
    """(a0, alpha[poles], theta[poles]) with Im theta > 0."""
    t = np.cos(np.linspace(0.0, np.pi, 4000))
    t = t[t > -1.0 + 1e-12]
    F = np.exp(s * (t - 1.0) / (t + 1.0))
    pol = _aaa(t.astype(complex), F.astype(complex), 2 * poles + 1)
    x_pol = s * (pol - 1.0) / (pol + 1.0)
    theta = np.sort_complex(x_pol[x_pol.imag > 0])
    assert len(theta) == poles
    tt = np.cos(np.linspace(0.0, np.pi, 20000))
    tt = tt[tt > -1.0 + 1e-14]
    x = s * (tt - 1.0) / (tt + 1.0)
    cols = [np.ones_like(x)]
    for p in theta:
        g = 1.0 / (x - p)
        cols += [2.0 * g.real, -2.0 * g.imag]
    c = np.linalg.lstsq(np.stack(cols, 1), np.exp(x), rcond=None)[0]
    alpha = c[1::2] + 1j * c[2::2]
    a0 = c[0]
    a0 += 1.0 - evaluate(0.0, a0, alpha, theta)
    return a0, alpha, theta


def evaluate(x, a0, alpha, theta):
    """r(x) for real x (array or scalar), in the order the kernels sum it."""
    x = np.asarray(x, dtype=np.float64)
    r = a0 * np.ones_like(x)
    for a, p in zip(alpha, theta):
        r = r + 2.0 * (a / (x - p)).real
    return r


if __name__ == "__main__":
    a0, alpha, theta = coefficients()
    print(f"a0 = {float(a0)!r}")
    for a, p in zip(alpha, theta):
        print(f"theta {float(p.real)!r}, {float(p.imag)!r}   alpha {float(a.real)!r}, {float(a.imag)!r}")
