"""GPU parity of the adjoint kernels and of the autograd harness over them.

The reference has no derivatives; the bars come from the oracle's
restatement of the blocked reverse mode (oracle/adjoint.py, pinned on the CPU
by Murray's closed form and central differences):
* Cholesky adjoint: relative max error <= 1e-12 against the oracle;
* triangular-solve adjoint: <= 1e-12;
* GP negative log marginal likelihood gradient (lengthscale and K) through
  torch.autograd: <= 1e-9 against the FP64 closed form dK = (K^-1 - a a^T)/2,
  and against a central difference (h = 1e-6) of the library's own forward
  within 1e-6.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _spd(n, seed, shift=1.0):
    rng = np.random.default_rng(seed)
    G = rng.standard_normal((n, n))
    return np.asfortranarray(G @ G.T / max(n, 1) + shift * np.eye(n))


@pytest.mark.parametrize("n", [1, 5, 64, 65, 100, 300, 1000])
def test_potrf_adjoint(lx, dev, n):
    from oracle import adjoint as adj
    import linompss_amd.autograd as ag
    print(f"[potrf adjoint] n={n}", flush=True)
    A = _spd(n, n)
    L = np.linalg.cholesky(A)
    Lb = np.tril(np.random.default_rng(n + 1).standard_normal((n, n)))
    got = ag.potrf_adjoint(A, Lb, b=64)
    ref = adj.chol_adjoint_blocked(L, Lb, 64)
    scale = np.abs(ref).max()
    assert np.abs(np.tril(got) - ref).max() <= 1e-12 * scale
    assert np.abs(np.tril(got) - adj.chol_adjoint_closed(L, Lb)).max() <= 1e-11 * scale


def test_potrf_adjoint_dev_lda(lx, dev):
    """The device entry point with leading dimensions larger than n."""
    from oracle import adjoint as adj
    print("[potrf adjoint dev] n=200 ld=211", flush=True)
    n, ld = 200, 211
    A = _spd(n, 7)
    L = np.asfortranarray(np.linalg.cholesky(A))
    Lb = np.tril(np.random.default_rng(8).standard_normal((n, n)))
    hL = np.zeros((ld, n), order="F")
    hL[:n] = L
    hB = np.full((ld, n), 7.0, order="F")
    hB[:n] = Lb
    dL = lx.DeviceMatrix(ld, n)
    dB = lx.DeviceMatrix(ld, n)
    dL.upload(hL)
    dB.upload(hB)
    fn = lx.lib.linompss_dpotrf_adjoint_dev
    fn.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64, ctypes.c_double,
                   ctypes.c_void_p]
    assert fn(n, dL.ptr, ld, dB.ptr, ld, 0.5, None) == 0
    out = dB.to_host()
    # with logdet_weight 1/2: Lbar_ii += 1 / L_ii before the sweep
    ref = adj.chol_adjoint_blocked(L, Lb + np.diag(1.0 / np.diag(L)), 64)
    assert np.abs(np.tril(out[:n]) - ref).max() <= 1e-12 * np.abs(ref).max()
    assert (out[n:] == 7.0).all()  # rows past n untouched
    dL.free()
    dB.free()


@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("n,k", [(300, 3), (1000, 1), (64, 16)])
def test_trsm_adjoint(lx, dev, trans, n, k):
    from oracle import adjoint as adj
    import linompss_amd.autograd as ag
    print(f"[trsm adjoint] {trans} n={n} k={k}", flush=True)
    rng = np.random.default_rng(n * k)
    L = np.linalg.cholesky(_spd(n, 9))
    B = rng.standard_normal((n, k))
    X = np.linalg.solve(L.T if trans == "T" else L, B)
    Xb = rng.standard_normal((n, k))
    Bb, Lb = ag.trsm_adjoint(trans, L, X, Xb)
    Bref, Lref = adj.trsm_adjoint(trans, L, X, Xb)
    assert np.abs(Bb - Bref).max() <= 1e-12 * np.abs(Bref).max()
    assert np.abs(Lb - Lref).max() <= 1e-12 * np.abs(Lref).max()


@pytest.mark.parametrize("n", [300, 2048])
def test_gp_autograd(lx, dev, n):
    """dL/dl and dL/dK of the GP negative log marginal likelihood (8192-point
    workload shape at reduced n: points uniform in [0,1]^3, l = 0.3,
    noise 1e-2, y ~ N(0, K))."""
    import torch
    from oracle import adjoint as adj
    import linompss_amd.autograd as ag
    print(f"[gp autograd] n={n}", flush=True)
    rng = np.random.default_rng(3)
    P = rng.uniform(0, 1, (n, 3))
    ell0, noise = 0.3, 1e-2
    y = np.linalg.cholesky(adj.gp_kernel(P, ell0, noise)) @ rng.standard_normal(n)
    pts = torch.from_numpy(P)
    yt = torch.from_numpy(y)
    ell = torch.tensor(ell0, dtype=torch.float64, requires_grad=True)
    K = ag.gp_kernel(pts, ell, noise)
    K.retain_grad()
    loss = ag.gp_nll(pts, yt, ell, noise, K=K)
    loss.backward()
    dK_ref, dell_ref = adj.gp_nll_grad(P, y, ell0, noise)
    assert abs(loss.item() - adj.gp_nll(P, y, ell0, noise)) <= 1e-10 * abs(loss.item())
    assert abs(ell.grad.item() - dell_ref) <= 1e-9 * abs(dell_ref)
    assert np.abs(K.grad.numpy() - dK_ref).max() <= 1e-9 * np.abs(dK_ref).max()
    # central difference of the library's own forward
    h = 1e-6
    with torch.no_grad():
        fp = ag.gp_nll(pts, yt, torch.tensor(ell0 + h, dtype=torch.float64), noise).item()
        fm = ag.gp_nll(pts, yt, torch.tensor(ell0 - h, dtype=torch.float64), noise).item()
    fd = (fp - fm) / (2 * h)
    assert abs(ell.grad.item() - fd) <= 1e-6 * abs(fd)


@pytest.mark.parametrize("n", [300, 2048])
def test_gp_fused(lx, dev, n):
    """GaussianNLL (the forward's factor kept on the device for backward):
    value, dL/dl, dL/dK and dL/dy against the FP64 closed forms."""
    import torch
    from oracle import adjoint as adj
    import linompss_amd.autograd as ag
    print(f"[gp fused] n={n}", flush=True)
    rng = np.random.default_rng(4)
    P = rng.uniform(0, 1, (n, 3))
    ell0, noise = 0.3, 1e-2
    y = np.linalg.cholesky(adj.gp_kernel(P, ell0, noise)) @ rng.standard_normal(n)
    pts = torch.from_numpy(P)
    yt = torch.from_numpy(y).requires_grad_(True)
    ell = torch.tensor(ell0, dtype=torch.float64, requires_grad=True)
    K = ag.gp_kernel(pts, ell, noise)
    K.retain_grad()
    tm = {}
    loss = ag.gp_nll_fused(pts, yt, ell, noise, K=K, timing=tm)
    loss.backward()
    assert tm["forward_s"] > 0 and tm["backward_s"] > 0
    dK_ref, dell_ref = adj.gp_nll_grad(P, y, ell0, noise)
    assert abs(loss.item() - adj.gp_nll(P, y, ell0, noise)) <= 1e-10 * abs(loss.item())
    assert abs(ell.grad.item() - dell_ref) <= 1e-9 * abs(dell_ref)
    assert np.abs(K.grad.numpy() - dK_ref).max() <= 1e-9 * np.abs(dK_ref).max()
    ybar = np.linalg.solve(adj.gp_kernel(P, ell0, noise), y)
    assert np.abs(yt.grad.numpy() - ybar).max() <= 1e-9 * np.abs(ybar).max()


def test_gp_fused_second_backward(lx, dev):
    """backward twice on one graph (retain_graph=True): L and a stay on the
    device until the graph is released and backward only reads them, so the
    second pass gives the same gradient instead of reading freed buffers."""
    import torch
    from oracle import adjoint as adj
    import linompss_amd.autograd as ag
    n = 500
    print(f"[gp second backward] n={n}", flush=True)
    rng = np.random.default_rng(5)
    P = rng.uniform(0, 1, (n, 3))
    y = np.linalg.cholesky(adj.gp_kernel(P, 0.3, 1e-2)) @ rng.standard_normal(n)
    ell = torch.tensor(0.3, dtype=torch.float64, requires_grad=True)
    loss = ag.gp_nll_fused(torch.from_numpy(P), torch.from_numpy(y), ell, 1e-2)
    loss.backward(retain_graph=True)
    g1 = ell.grad.item()
    ell.grad = None
    loss.backward()
    assert ell.grad.item() == g1
