"""torch.autograd.Function wrappers over the library's FP64 Cholesky, triangular
solves and their adjoints (a ctypes harness over the C ABI), and a Gaussian-
process negative log marginal likelihood built from them.

The functions take and return CPU float64 tensors; every factor, solve and
adjoint runs on the GPU inside liblinompss (host arrays are staged through
device buffers by the library).  The Cholesky backward recomputes the factor
from A on the device instead of keeping it from the forward pass
(linompss_dpotrf_adjoint); the Gaussian-process likelihoods keep their factor
on the device between the passes.

Import this module (or ``linompss_amd``) before anything that initialises
torch's own GPU runtime: torch bundles a HIP runtime with the same soname,
and the process keeps whichever copy is loaded first.
"""
from __future__ import annotations

import ctypes
import itertools
import math

import numpy as np

from . import DeviceMatrix, LinompssError, _check, copy_dev, lib, ompss_dchol_rl

import torch  # noqa: E402  (after the library, see the module docstring)

_p = ctypes.c_void_p
_i = ctypes.c_int
lib.linompss_dpotrf_adjoint.restype = _i
lib.linompss_dpotrf_adjoint.argtypes = [_i, _i, _p, _i, _p, _i, ctypes.c_double]
lib.linompss_dtrsm_adjoint.restype = _i
lib.linompss_dtrsm_adjoint.argtypes = [ctypes.c_char, _i, _i, _p, _i, _p, _i, _p, _i, _p, _i]
_i64 = ctypes.c_int64
_d = ctypes.c_double
lib.linompss_dpotrf_adjoint_dev.restype = _i
lib.linompss_dpotrf_adjoint_dev.argtypes = [_i, _p, _i64, _p, _i64, _d, _p]
lib.linompss_dtrsm_adjoint_dev.restype = _i
lib.linompss_dtrsm_adjoint_dev.argtypes = [_i, _i, _i, _p, _i64, _p, _i64, _p, _i64, _p, _i64, _i, _p]
lib.linompss_trsm_dev.restype = _i
lib.linompss_trsm_dev.argtypes = [_i, _i, _i, _i, _p, _i64, _p, _i64, _p]
lib.linompss_dlogdet_dev.restype = _i
lib.linompss_dlogdet_dev.argtypes = [_i, _p, _i64, ctypes.POINTER(_d), _p]


lib.linompss_gp_kernel_dev.restype = _i
lib.linompss_gp_kernel_dev.argtypes = [_i, _i, _p, _i64, _d, _d, _d, _p, _i64, _p]
lib.linompss_gp_grad_dev.restype = _i
lib.linompss_gp_grad_dev.argtypes = [_i, _i, _p, _i64, _d, _d, _p, _i64, ctypes.POINTER(_d), _p, _i64, _p]


def _f64(t: torch.Tensor) -> np.ndarray:
    """Column-major float64 copy of a 2-D CPU tensor."""
    return np.asfortranarray(t.detach().to(dtype=torch.float64, device="cpu").numpy())


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_p)


def potrf_adjoint(A: np.ndarray, Lbar: np.ndarray, b: int = 256, logdet_weight: float = 0.0) -> np.ndarray:
    """Abar (lower representation) of A = L L^T given Lbar (plus the adjoint of
    logdet_weight * log det A); the factor is recomputed from A on the device."""
    n = A.shape[0]
    A = np.asfortranarray(A, dtype=np.float64)
    out = np.asfortranarray(np.tril(Lbar), dtype=np.float64)
    _check("linompss_dpotrf_adjoint",
           lib.linompss_dpotrf_adjoint(n, b, _ptr(A), n, _ptr(out), n, float(logdet_weight)), False)
    return out


def trsm_adjoint(trans: str, L: np.ndarray, X: np.ndarray, Xbar: np.ndarray, want_lbar: bool = True):
    """(Bbar, Lbar) for X = op(L)^-1 B (op = 'N' or 'T')."""
    n, k = X.shape
    L = np.asfortranarray(L, dtype=np.float64)
    X = np.asfortranarray(X, dtype=np.float64)
    Bb = np.asfortranarray(Xbar, dtype=np.float64).copy(order="F")
    Lb = np.zeros((n, n), order="F") if want_lbar else None
    rc = lib.linompss_dtrsm_adjoint(trans.encode(), n, k, _ptr(L), n, _ptr(X), n, _ptr(Bb), n,
                                    _ptr(Lb) if Lb is not None else None, n)
    _check("linompss_dtrsm_adjoint", rc, False)
    return Bb, Lb


class Cholesky(torch.autograd.Function):
    """L = chol(A) (lower; A symmetric, its lower triangle read).  Backward
    returns the symmetric gradient."""

    @staticmethod
    def forward(ctx, A: torch.Tensor, b: int = 256):
        L = _f64(A).copy(order="F")
        info = ompss_dchol_rl(L, b)
        if info != 0:
            raise LinompssError("ompss_dchol_rl", info)
        ctx.save_for_backward(A)  # the factor itself is recomputed in backward
        ctx.b = b
        return torch.from_numpy(np.tril(L)).to(A.dtype)

    @staticmethod
    def backward(ctx, Lbar: torch.Tensor):
        (A,) = ctx.saved_tensors
        Ab = potrf_adjoint(_f64(A), _f64(Lbar), ctx.b)
        low = np.tril(Ab, -1)
        S = 0.5 * (low + low.T) + np.diag(np.diag(Ab))
        return torch.from_numpy(S).to(A.dtype), None


class TriSolve(torch.autograd.Function):
    """X = op(L)^-1 B for lower L (op 'N' or 'T')."""

    @staticmethod
    def forward(ctx, L: torch.Tensor, B: torch.Tensor, trans: str = "N"):
        Ln = _f64(L)
        X = _f64(B if B.dim() == 2 else B.reshape(-1, 1)).copy(order="F")
        n, k = X.shape
        rc = lib.linompss_dltrsm(b"L", trans.encode(), b"N", 256, k, n, k, 1.0, _ptr(Ln), n, _ptr(X), n)
        _check("linompss_dltrsm", rc, False)
        ctx.trans = trans
        ctx.vec = B.dim() == 1
        Xt = torch.from_numpy(X).to(B.dtype)
        ctx.save_for_backward(L, Xt)
        return Xt.reshape(-1) if ctx.vec else Xt

    @staticmethod
    def backward(ctx, Xbar: torch.Tensor):
        L, X = ctx.saved_tensors
        Xb = Xbar.reshape(-1, 1) if ctx.vec else Xbar
        Bb, Lb = trsm_adjoint(ctx.trans, _f64(L), _f64(X), _f64(Xb), ctx.needs_input_grad[0])
        gB = torch.from_numpy(Bb).to(X.dtype)
        gL = torch.from_numpy(Lb).to(X.dtype) if Lb is not None else None
        return gL, (gB.reshape(-1) if ctx.vec else gB), None


class GaussianNLL(torch.autograd.Function):
    """nll(K, y) = y^T K^-1 y / 2 + log det K / 2 + n log(2 pi) / 2 with the
    factor L kept on the device between the passes: forward = FP64 factor, one
    sweep and the log det; backward = the solve adjoint and the Cholesky
    adjoint (with the log det term) on the forward's L, so only K goes up and
    dK (symmetric) comes down.  ``timing`` (a dict, optional) receives the
    device seconds of each pass."""

    @staticmethod
    def forward(ctx, K: torch.Tensor, y: torch.Tensor, b: int = 256, timing: dict | None = None):
        import time
        n = K.shape[0]
        dL = DeviceMatrix.from_host(_f64(K))
        da = DeviceMatrix.from_host(y.detach().to(torch.float64).numpy().reshape(-1, 1))
        lib.linompss_dev_sync()
        t0 = time.perf_counter()
        _check("linompss_chol_dev", lib.linompss_chol_dev(0, 0, n, b, dL.ptr, n, None, None))
        _check("linompss_trsm_dev", lib.linompss_trsm_dev(0, 0, n, 1, dL.ptr, n, da.ptr, n, None), False)
        ld = _d(0.0)
        _check("linompss_dlogdet_dev", lib.linompss_dlogdet_dev(n, dL.ptr, n, ctypes.byref(ld), None), False)
        lib.linompss_dev_sync()
        if timing is not None:
            timing["forward_s"] = time.perf_counter() - t0
        a = da.to_host()[:, 0]
        ctx.dL, ctx.da, ctx.b, ctx.timing = dL, da, b, timing
        ctx.dtype = K.dtype
        nll = 0.5 * float(a @ a) + 0.5 * ld.value + 0.5 * n * math.log(2.0 * math.pi)
        return torch.tensor(nll, dtype=K.dtype)

    @staticmethod
    def backward(ctx, g: torch.Tensor):
        import time
        dL, da = ctx.dL, ctx.da
        n = dL.rows
        dLb = DeviceMatrix(n, n)
        dbb = DeviceMatrix(n, 1)
        lib.linompss_dev_sync()
        t0 = time.perf_counter()
        copy_dev(dbb, da)  # abar = a (from a^T a / 2)
        # bbar = L^-T a (= K^-1 y) and Lbar = -tril(bbar a^T), written without reading Lbar
        _check("linompss_dtrsm_adjoint_dev", lib.linompss_dtrsm_adjoint_dev(
            0, n, 1, dL.ptr, n, da.ptr, n, dbb.ptr, n, dLb.ptr, n, 0, None), False)
        # + the log det term, then the reverse sweep: Lbar -> Kbar (lower representation)
        _check("linompss_dpotrf_adjoint_dev", lib.linompss_dpotrf_adjoint_dev(
            n, dL.ptr, n, dLb.ptr, n, 0.5, None), False)
        lib.linompss_dev_sync()
        if ctx.timing is not None:
            ctx.timing["backward_s"] = time.perf_counter() - t0
        Ab = dLb.to_host()
        ybar = dbb.to_host()[:, 0]
        # L and a are only read here and stay on the device until ctx is
        # released, so a second backward (retain_graph=True) gives the same
        # gradient
        for m in (dLb, dbb):
            m.free()
        low = np.tril(Ab, -1)
        S = 0.5 * (low + low.T) + np.diag(np.diag(Ab))
        gv = float(g)
        return torch.from_numpy(S * gv).to(ctx.dtype), torch.from_numpy(ybar * gv).to(ctx.dtype), None, None


def gp_kernel(points: torch.Tensor, lengthscale: torch.Tensor, noise: float) -> torch.Tensor:
    """K_ij = exp(-|p_i - p_j|^2 / (2 l^2)) + noise delta_ij (differentiable in l)."""
    sq = (points * points).sum(-1)
    d2 = (sq[:, None] + sq[None, :] - 2.0 * points @ points.T).clamp_min(0.0)
    return torch.exp(-d2 / (2.0 * lengthscale * lengthscale)) + noise * torch.eye(points.shape[0], dtype=points.dtype)


def gp_nll(points: torch.Tensor, y: torch.Tensor, lengthscale: torch.Tensor, noise: float = 1e-2,
           b: int = 256, K: torch.Tensor | None = None) -> torch.Tensor:
    """Negative log marginal likelihood y^T K^-1 y / 2 + log det K / 2 + n log(2 pi) / 2:
    FP64 Cholesky, log det from its diagonal, one triangular solve."""
    if K is None:
        K = gp_kernel(points, lengthscale, noise)
    L = Cholesky.apply(K, b)
    a = TriSolve.apply(L, y, "N")
    return 0.5 * (a * a).sum() + torch.log(torch.diagonal(L)).sum() + 0.5 * y.shape[0] * math.log(2.0 * math.pi)


def gp_nll_fused(points: torch.Tensor, y: torch.Tensor, lengthscale: torch.Tensor, noise: float = 1e-2,
                 b: int = 256, K: torch.Tensor | None = None, timing: dict | None = None) -> torch.Tensor:
    """The same quantity through GaussianNLL (K resident on the device between passes)."""
    if K is None:
        K = gp_kernel(points, lengthscale, noise)
    return GaussianNLL.apply(K, y, b, timing)


# ---- Gaussian-process likelihood with K built on the device --------------------

class _GPWork:
    """Device buffers of one problem shape, reused across calls: K (the factor
    in place), Kbar, the points, y, a = L^-1 y, bbar and dX.  ``owner`` names
    the forward pass whose points, y, factor and a the buffers hold (0: none);
    the backward passes only read those four."""

    def __init__(self, n: int, d: int):
        self.n, self.d = n, d
        self.owner = 0
        self.K = DeviceMatrix(n, n)
        self.Kb = DeviceMatrix(n, n)
        self.P = DeviceMatrix(n, d)
        self.y = DeviceMatrix(n, 1)
        self.a = DeviceMatrix(n, 1)
        self.bb = DeviceMatrix(n, 1)
        self.dX = DeviceMatrix(n, d)

    def free(self):
        for m in (self.K, self.Kb, self.P, self.y, self.a, self.bb, self.dX):
            m.free()


_GP_WS: dict = {}
_GP_PASS = itertools.count(1)  # forward passes of GPMarginalLikelihood


def _gp_work(n: int, d: int) -> _GPWork:
    ws = _GP_WS.get((n, d))
    if ws is None:
        for k in list(_GP_WS):  # one shape at a time: K and Kbar are n^2 each
            _GP_WS.pop(k).free()
        ws = _GP_WS[(n, d)] = _GPWork(n, d)
    return ws


def gp_release_workspace() -> None:
    """Frees the device buffers GPMarginalLikelihood keeps between calls."""
    for k in list(_GP_WS):
        _GP_WS.pop(k).free()


def gp_kernel_dev(points: np.ndarray, ell: float, sf2: float = 1.0, noise: float = 0.0) -> np.ndarray:
    """K built on the device (linompss_gp_kernel_dev), downloaded."""
    P = np.asfortranarray(points, dtype=np.float64)
    n, d = P.shape
    dP = DeviceMatrix.from_host(P)
    dK = DeviceMatrix(n, n)
    try:
        _check("linompss_gp_kernel_dev",
               lib.linompss_gp_kernel_dev(n, d, dP.ptr, n, float(ell), float(sf2), float(noise), dK.ptr, n, None),
               False)
        return dK.to_host()
    finally:
        dP.free()
        dK.free()


def _scalar(v) -> float:
    return float(v.detach()) if isinstance(v, torch.Tensor) else float(v)


def _gp_forward_dev(ws: _GPWork, ell: float, sf2: float, noise: float, b: int) -> float:
    """K from the points, L in place, a = L^-1 y; returns log det K."""
    n, d = ws.n, ws.d
    _check("linompss_gp_kernel_dev",
           lib.linompss_gp_kernel_dev(n, d, ws.P.ptr, n, ell, sf2, noise, ws.K.ptr, n, None), False)
    _check("linompss_chol_dev", lib.linompss_chol_dev(0, 0, n, b, ws.K.ptr, n, None, None))
    copy_dev(ws.a, ws.y)
    _check("linompss_trsm_dev", lib.linompss_trsm_dev(0, 0, n, 1, ws.K.ptr, n, ws.a.ptr, n, None), False)
    ld = _d(0.0)
    _check("linompss_dlogdet_dev", lib.linompss_dlogdet_dev(n, ws.K.ptr, n, ctypes.byref(ld), None), False)
    return ld.value


class GPMarginalLikelihood(torch.autograd.Function):
    """nll(X, y; ell, sf2, noise) = y^T K^-1 y / 2 + log det K / 2 + n log(2 pi) / 2
    with K_ij = sf2 exp(-|x_i - x_j|^2 / (2 ell^2)) + noise delta_ij built on
    the device from the points: only X (n x d), y and scalars cross PCIe.
    Forward: K, FP64 factor in place, one sweep, log det.  Backward: the
    forward's factor and a = L^-1 y from the shape's workspace (recomputed only
    when another forward of the same shape has used the workspace since), the
    solve adjoint, the blocked Cholesky adjoint with the log det term, then the
    contraction of Kbar with dK/dell, dK/dsf2, dK/dnoise and dK/dX on the
    device.  Gradients flow to X, y and to ell / sf2 / noise when they are
    tensors.  ``timing`` (a dict, optional) receives each pass's device seconds."""

    @staticmethod
    def forward(ctx, points, y, ell, sf2, noise, b: int = 256, timing: dict | None = None):
        import time
        P = _f64(points)
        n, d = P.shape
        ws = _gp_work(n, d)
        ws.P.upload(P)
        ws.y.upload(y.detach().to(torch.float64).numpy().reshape(-1, 1))
        vals = (_scalar(ell), _scalar(sf2), _scalar(noise))
        lib.linompss_dev_sync()
        t0 = time.perf_counter()
        logdet = _gp_forward_dev(ws, *vals, b)
        ws.owner = ctx.owner = next(_GP_PASS)
        lib.linompss_dev_sync()
        if timing is not None:
            timing["forward_s"] = time.perf_counter() - t0
        a = ws.a.to_host()[:, 0]
        ctx.save_for_backward(points, y)
        ctx.vals, ctx.b, ctx.timing = vals, b, timing
        ctx.is_tensor = [isinstance(v, torch.Tensor) for v in (ell, sf2, noise)]
        nll = 0.5 * float(a @ a) + 0.5 * logdet + 0.5 * n * math.log(2.0 * math.pi)
        return torch.tensor(nll, dtype=torch.float64)

    @staticmethod
    def backward(ctx, g: torch.Tensor):
        import time
        points, y = ctx.saved_tensors
        ell, sf2, noise = ctx.vals
        P = _f64(points)
        n, d = P.shape
        ws = _gp_work(n, d)
        recompute = ws.owner != ctx.owner
        if recompute:
            ws.P.upload(P)
            ws.y.upload(y.detach().to(torch.float64).numpy().reshape(-1, 1))
        lib.linompss_dev_sync()
        t0 = time.perf_counter()
        if recompute:  # another forward of this shape ran in between
            _gp_forward_dev(ws, ell, sf2, noise, ctx.b)
            ws.owner = ctx.owner
        copy_dev(ws.bb, ws.a)  # abar = a (from a^T a / 2)
        # bbar = L^-T a (= K^-1 y) and Lbar = -tril(bbar a^T), written without reading Lbar
        _check("linompss_dtrsm_adjoint_dev", lib.linompss_dtrsm_adjoint_dev(
            0, n, 1, ws.K.ptr, n, ws.a.ptr, n, ws.bb.ptr, n, ws.Kb.ptr, n, 0, None), False)
        # + the log det term, then the reverse sweep: Lbar -> Kbar (lower representation)
        _check("linompss_dpotrf_adjoint_dev", lib.linompss_dpotrf_adjoint_dev(
            n, ws.K.ptr, n, ws.Kb.ptr, n, 0.5, None), False)
        grads = (_d * 3)()
        want_x = ctx.needs_input_grad[0]
        _check("linompss_gp_grad_dev", lib.linompss_gp_grad_dev(
            n, d, ws.P.ptr, n, ell, sf2, ws.Kb.ptr, n, grads, ws.dX.ptr if want_x else None, n, None), False)
        lib.linompss_dev_sync()
        if ctx.timing is not None:
            ctx.timing["backward_s"] = time.perf_counter() - t0
        gv = float(g)
        gx = torch.from_numpy(ws.dX.to_host() * gv).to(points.dtype) if want_x else None
        gy = torch.from_numpy(ws.bb.to_host()[:, 0] * gv).to(y.dtype) if ctx.needs_input_grad[1] else None
        gs = [torch.tensor(grads[q] * gv, dtype=torch.float64) if ctx.is_tensor[q] else None for q in range(3)]
        return gx, gy, gs[0], gs[1], gs[2], None, None


def gp_nll_device(points: torch.Tensor, y: torch.Tensor, lengthscale, sf2=1.0, noise=1e-2, b: int = 256,
                  timing: dict | None = None) -> torch.Tensor:
    """The GP negative log marginal likelihood with K built on the device
    (GPMarginalLikelihood)."""
    return GPMarginalLikelihood.apply(points, y, lengthscale, sf2, noise, b, timing)
