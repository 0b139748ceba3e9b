"""Pins of the adjoint oracle (oracle/adjoint.py) on the CPU: the blocked
reverse mode against Murray's closed form (a different algorithm), both
against central differences of numpy's Cholesky, the solve adjoint and the
GP gradient against central differences."""
import numpy as np
import pytest

from oracle import adjoint as adj


def _spd(n, seed):
    rng = np.random.default_rng(seed)
    G = rng.standard_normal((n, n))
    return G @ G.T / max(n, 1) + np.eye(n)


@pytest.mark.parametrize("n,nb", [(1, 64), (7, 1), (7, 3), (64, 64), (65, 64), (150, 16), (150, 64)])
def test_blocked_matches_closed_form(n, nb):
    rng = np.random.default_rng(n + nb)
    L = np.linalg.cholesky(_spd(n, n))
    Lb = np.tril(rng.standard_normal((n, n)))
    a = adj.chol_adjoint_blocked(L, Lb, nb)
    c = adj.chol_adjoint_closed(L, Lb)
    assert np.abs(np.triu(a, 1)).max(initial=0.0) == 0.0
    assert np.abs(a - c).max() <= 1e-12 * max(1.0, np.abs(c).max())


def test_cholesky_adjoint_central_differences():
    """sum Lbar * dL along a symmetric direction E equals sum_{i>=j} Abar E."""
    n, h = 12, 1e-6
    rng = np.random.default_rng(5)
    A = _spd(n, 3)
    Lb = np.tril(rng.standard_normal((n, n)))
    E = rng.standard_normal((n, n))
    E = E + E.T
    fd = (np.sum(Lb * np.linalg.cholesky(A + h * E)) - np.sum(Lb * np.linalg.cholesky(A - h * E))) / (2 * h)
    Ab = adj.chol_adjoint_blocked(np.linalg.cholesky(A), Lb, nb=4)
    assert abs(np.sum(np.tril(Ab) * np.tril(E)) - fd) <= 1e-7 * max(1.0, abs(fd))
    # the symmetric gradient gives the same directional derivative over the full E
    S = adj.lower_to_symmetric(Ab)
    assert abs(np.sum(S * E) - fd) <= 1e-7 * max(1.0, abs(fd))


@pytest.mark.parametrize("trans", ["N", "T"])
def test_trsm_adjoint_central_differences(trans):
    n, k, h = 10, 2, 1e-6
    rng = np.random.default_rng(6)
    L = np.linalg.cholesky(_spd(n, 4))
    B = rng.standard_normal((n, k))
    Xb = rng.standard_normal((n, k))
    op = (lambda M: M.T) if trans == "T" else (lambda M: M)
    X = np.linalg.solve(op(L), B)
    Bb, Lb = adj.trsm_adjoint(trans, L, X, Xb)
    dB = rng.standard_normal((n, k))
    dL = np.tril(rng.standard_normal((n, n)))
    f = lambda Lm, Bm: np.sum(Xb * np.linalg.solve(op(Lm), Bm))
    fd_b = (f(L, B + h * dB) - f(L, B - h * dB)) / (2 * h)
    fd_l = (f(L + h * dL, B) - f(L - h * dL, B)) / (2 * h)
    assert abs(np.sum(Bb * dB) - fd_b) <= 1e-7 * max(1.0, abs(fd_b))
    assert abs(np.sum(Lb * dL) - fd_l) <= 1e-7 * max(1.0, abs(fd_l))


def test_gp_gradient_central_differences():
    n, ell, noise, h = 40, 0.3, 1e-2, 1e-6
    rng = np.random.default_rng(3)
    P = rng.uniform(0, 1, (n, 3))
    y = np.linalg.cholesky(adj.gp_kernel(P, ell, noise)) @ rng.standard_normal(n)
    dK, dell = adj.gp_nll_grad(P, y, ell, noise)
    fd = (adj.gp_nll(P, y, ell + h, noise) - adj.gp_nll(P, y, ell - h, noise)) / (2 * h)
    assert abs(dell - fd) <= 1e-6 * max(1.0, abs(fd))
    # one off-diagonal entry, perturbed symmetrically, and one diagonal entry
    K = adj.gp_kernel(P, ell, noise)

    def nll_K(Km):
        L = np.linalg.cholesky(Km)
        a = np.linalg.solve(L, y)
        return 0.5 * a @ a + np.log(np.diag(L)).sum()
    for (i, j) in [(5, 2), (7, 7)]:
        E = np.zeros((n, n))
        E[i, j] = E[j, i] = 1.0
        fdk = (nll_K(K + h * E) - nll_K(K - h * E)) / (2 * h)
        assert abs(np.sum(dK * E) - fdk) <= 1e-6 * max(1.0, abs(fdk))


def test_gp_all_gradients_central_differences():
    """gp_nll_grad_all (closed forms through dK) against central differences of
    gp_nll in every parameter: the length scale, the amplitude, the noise, one
    coordinate of one point, and one entry of y."""
    n, ell, noise, sf2, h = 36, 0.35, 2e-2, 1.7, 1e-6
    rng = np.random.default_rng(11)
    P = rng.uniform(0, 1, (n, 3))
    y = np.linalg.cholesky(adj.gp_kernel(P, ell, noise, sf2)) @ rng.standard_normal(n)
    g = adj.gp_nll_grad_all(P, y, ell, noise, sf2)
    f = lambda **kw: adj.gp_nll(kw.get("P", P), kw.get("y", y), kw.get("ell", ell), kw.get("noise", noise),
                                kw.get("sf2", sf2))
    fd = lambda name, v: (f(**{name: v + h}) - f(**{name: v - h})) / (2 * h)
    for name, v in (("ell", ell), ("noise", noise), ("sf2", sf2)):
        assert abs(g["d" + name] - fd(name, v)) <= 1e-6 * max(1.0, abs(g["d" + name])), name
    for (i, q) in [(4, 0), (17, 2)]:
        Pp, Pm = P.copy(), P.copy()
        Pp[i, q] += h
        Pm[i, q] -= h
        fdp = (f(P=Pp) - f(P=Pm)) / (2 * h)
        assert abs(g["dP"][i, q] - fdp) <= 1e-6 * max(1.0, abs(fdp))
    yp, ym = y.copy(), y.copy()
    yp[9] += h
    ym[9] -= h
    assert abs(g["dy"][9] - (f(y=yp) - f(y=ym)) / (2 * h)) <= 1e-6 * max(1.0, abs(g["dy"][9]))
    # with sf2 = 1 the older function agrees
    assert adj.gp_nll_grad(P, y, ell, noise)[1] == pytest.approx(adj.gp_nll_grad_all(P, y, ell, noise)["dell"],
                                                                 rel=1e-12)
