"""Adjoint oracle (numpy) -- TEST INFRASTRUCTURE ONLY.

Only ``tests/`` and ``bench.py``'s CPU legs may import this module, as the
checker.  The reference has no derivatives (its factor is task_dpotrf,
tasks_potrf.c:20-24, its solves bblas_dltrsm, bblas_trsm.h:116-160); these
restate the reverse-mode algorithms the HIP path implements, plus independent
closed forms to pin them:

* ``chol_adjoint_blocked`` -- blocked reverse mode of the block Cholesky
  (I. Murray, "Differentiation of the Cholesky decomposition", 2016), the
  same block recurrences as the device code.  Pinned by
  ``chol_adjoint_closed`` (Murray's symmetric closed form, a different
  algorithm) and by central differences of ``numpy.linalg.cholesky``.
* ``trsm_adjoint`` -- closed form of the triangular-solve adjoint, pinned by
  central differences of ``numpy.linalg.solve``.
* ``gp_nll`` / ``gp_nll_grad`` -- Gaussian-process negative log marginal
  likelihood and its gradient through dK = (K^-1 - a a^T) / 2, pinned by
  central differences in the length scale and in single entries of K.

Gradients with respect to a symmetric A are given in the *lower
representation* (dLoss = sum_{i>=j} Abar_ij dA_ij, A read from its lower
triangle) unless a function says otherwise.
"""
from __future__ import annotations

import numpy as np
from scipy.linalg import solve_triangular


def chol_level2_adjoint(D: np.ndarray, Dbar: np.ndarray) -> np.ndarray:
    """Unblocked reverse mode of D = chol(P) on one block: returns Pbar (lower)."""
    X = np.tril(np.array(Dbar, dtype=np.float64))
    m = D.shape[0]
    for j in range(m - 1, -1, -1):
        d = D[j, j]
        X[j + 1:, j] /= d
        X[j, j] = (X[j, j] - X[j + 1:, j] @ D[j + 1:, j]) / (2.0 * d)
        X[j + 1:, :j] -= np.outer(X[j + 1:, j], D[j, :j])
        X[j, :j] -= X[j + 1:, j] @ D[j + 1:, :j] + 2.0 * X[j, j] * D[j, :j]
    return X


def chol_adjoint_blocked(L: np.ndarray, Lbar: np.ndarray, nb: int = 64) -> np.ndarray:
    """Abar (lower representation) from L and Lbar, blocks of nb columns from
    the last to the first:  Qbar = Cbar D^-1;  Dbar -= tril(Qbar^T C);
    Pbar = level-2 adjoint;  Bbar -= Qbar R;  Rbar -= Qbar^T B + (Pbar + Pbar^T) R."""
    n = L.shape[0]
    Ab = np.tril(np.array(Lbar, dtype=np.float64))
    for j0 in range(((n - 1) // nb) * nb if n else -1, -1, -nb):
        j1 = min(n, j0 + nb)
        D, R, B, C = L[j0:j1, j0:j1], L[j0:j1, :j0], L[j1:, :j0], L[j1:, j0:j1]
        Q = solve_triangular(D, Ab[j1:, j0:j1].T, lower=True, trans=1).T  # Cbar D^-1
        Ab[j1:, j0:j1] = Q
        Ab[j0:j1, j0:j1] -= np.tril(Q.T @ C)
        P = chol_level2_adjoint(D, Ab[j0:j1, j0:j1])
        Ab[j0:j1, j0:j1] = P
        Ab[j1:, :j0] -= Q @ R
        Ab[j0:j1, :j0] -= Q.T @ B + (P + P.T) @ R
    return Ab


def chol_adjoint_closed(L: np.ndarray, Lbar: np.ndarray) -> np.ndarray:
    """Murray's closed form: G = L^-T Phi(L^T Lbar) L^-1 (Phi: lower triangle,
    halved diagonal); symmetric gradient S = (G + G^T)/2; lower representation
    2 S off the diagonal, S on it."""
    n = L.shape[0]
    M = L.T @ np.tril(Lbar)
    Phi = np.tril(M) - 0.5 * np.diag(np.diag(M))
    T = solve_triangular(L, Phi, lower=True, trans=1)          # L^-T Phi
    G = solve_triangular(L, T.T, lower=True, trans=1).T        # (L^-T Phi) L^-1
    S = 0.5 * (G + G.T)
    return np.tril(2.0 * S) - np.diag(np.diag(S)) if n else S


def lower_to_symmetric(Abar: np.ndarray) -> np.ndarray:
    """The symmetric gradient from the lower representation."""
    low = np.tril(Abar, -1)
    return 0.5 * (low + low.T) + np.diag(np.diag(Abar))


def trsm_adjoint(trans: str, L: np.ndarray, X: np.ndarray, Xbar: np.ndarray):
    """X = op(L)^-1 B:  Bbar = op(L)^-T Xbar;  Lbar = -tril(Bbar X^T) (op = N)
    or -tril(X Bbar^T) (op = T)."""
    t = trans.upper() == "T"
    Bbar = solve_triangular(L, Xbar, lower=True, trans=0 if t else 1)
    Lbar = -np.tril(X @ Bbar.T) if t else -np.tril(Bbar @ X.T)
    return Bbar, Lbar


def gp_kernel(P: np.ndarray, ell: float, noise: float, sf2: float = 1.0) -> np.ndarray:
    """K_ij = sf2 exp(-|p_i - p_j|^2 / (2 ell^2)) + noise delta_ij."""
    d2 = ((P[:, None, :] - P[None, :, :]) ** 2).sum(-1)
    return sf2 * np.exp(-d2 / (2.0 * ell * ell)) + noise * np.eye(P.shape[0])


def gp_nll(P: np.ndarray, y: np.ndarray, ell: float, noise: float, sf2: float = 1.0) -> float:
    """y^T K^-1 y / 2 + log det K / 2 + n log(2 pi) / 2."""
    K = gp_kernel(P, ell, noise, sf2)
    L = np.linalg.cholesky(K)
    a = solve_triangular(L, y, lower=True)
    return float(0.5 * a @ a + np.log(np.diag(L)).sum() + 0.5 * len(y) * np.log(2.0 * np.pi))


def gp_nll_grad(P: np.ndarray, y: np.ndarray, ell: float, noise: float):
    """(dK symmetric, d ell): dK = (K^-1 - a a^T) / 2 with a = K^-1 y;
    d ell = sum_ij dK_ij dK_ij/d ell, dK_ij/d ell = K0_ij |p_i - p_j|^2 / ell^3."""
    d2 = ((P[:, None, :] - P[None, :, :]) ** 2).sum(-1)
    K0 = np.exp(-d2 / (2.0 * ell * ell))
    K = K0 + noise * np.eye(P.shape[0])
    Ki = np.linalg.inv(K)
    a = Ki @ y
    dK = 0.5 * (Ki - np.outer(a, a))
    dK = 0.5 * (dK + dK.T)
    return dK, float((dK * K0 * d2).sum() / ell ** 3)


def gp_nll_grad_all(P: np.ndarray, y: np.ndarray, ell: float, noise: float, sf2: float = 1.0) -> dict:
    """Every gradient of gp_nll through the symmetric dK = (K^-1 - a a^T) / 2:
    d ell = sum_ij dK_ij k_ij |p_i - p_j|^2 / ell^3, d sf2 = sum_ij dK_ij k_ij / sf2,
    d noise = trace(dK), dP_i = -2 / ell^2 sum_j dK_ij k_ij (p_i - p_j)
    (k = sf2 exp(...), the noise-free part of K)."""
    diff = P[:, None, :] - P[None, :, :]
    d2 = (diff ** 2).sum(-1)
    k = sf2 * np.exp(-d2 / (2.0 * ell * ell))
    K = k + noise * np.eye(P.shape[0])
    Ki = np.linalg.inv(K)
    a = Ki @ y
    dK = 0.5 * (Ki - np.outer(a, a))
    dK = 0.5 * (dK + dK.T)
    W = dK * k
    return {"dell": float((W * d2).sum() / ell ** 3), "dsf2": float(W.sum() / sf2), "dnoise": float(np.trace(dK)),
            "dP": -2.0 / (ell * ell) * (W[:, :, None] * diff).sum(1), "dy": a}
