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
