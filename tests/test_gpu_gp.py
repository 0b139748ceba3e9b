"""The Gaussian-process likelihood with K built on the device
(linompss_gp_kernel_dev, linompss_gp_grad_dev, GPMarginalLikelihood): the
kernel matrix against the oracle's, and every gradient (length scale,
amplitude, noise, the points, y) against the oracle's FP64 closed forms,
which tests/test_adjoint_oracle.py pins by central differences.  The
workload's shape (points uniform in [0,1]^3, l = 0.3, noise 1e-2, y ~ N(0, K))
at reduced n, and at the full 8192 points in test_gpu_configs.py."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _problem(n, seed, ell, noise, sf2=1.0, d=3):
    from oracle import adjoint as adj
    rng = np.random.default_rng(seed)
    P = rng.uniform(0, 1, (n, d))
    y = np.linalg.cholesky(adj.gp_kernel(P, ell, noise, sf2)) @ rng.standard_normal(n)
    return P, y


@pytest.mark.parametrize("n,d", [(1, 3), (63, 3), (200, 1), (700, 5), (1000, 16)])
def test_gp_kernel_matches_oracle(lx, dev, n, d):
    from oracle import adjoint as adj
    import linompss_amd.autograd as ag
    print(f"[gp kernel] n={n} d={d}", flush=True)
    P = np.random.default_rng(n).uniform(-1, 1, (n, d))
    K = ag.gp_kernel_dev(P, 0.45, 1.3, 0.02)
    Kref = adj.gp_kernel(P, 0.45, 0.02, 1.3)
    assert np.abs(K - Kref).max() <= 4e-16 * np.abs(Kref).max()
    assert np.array_equal(K, K.T)


def test_gp_kernel_argument_checks(lx):
    L = lx.lib
    p = ctypes.c_void_p(16)
    assert L.linompss_gp_kernel_dev(-1, 3, p, 1, 0.3, 1.0, 0.0, p, 1, None) == -1
    assert L.linompss_gp_kernel_dev(4, 17, p, 4, 0.3, 1.0, 0.0, p, 4, None) == -2
    assert L.linompss_gp_kernel_dev(4, 3, p, 3, 0.3, 1.0, 0.0, p, 4, None) == -4
    assert L.linompss_gp_kernel_dev(4, 3, p, 4, 0.0, 1.0, 0.0, p, 4, None) == -5
    assert L.linompss_gp_kernel_dev(4, 3, p, 4, 0.3, 1.0, 0.0, p, 3, None) == -9
    g = (ctypes.c_double * 3)()
    assert L.linompss_gp_grad_dev(4, 3, p, 4, 0.3, 0.0, p, 4, g, None, 4, None) == -6
    assert L.linompss_gp_grad_dev(4, 3, p, 4, 0.3, 1.0, p, 3, g, None, 4, None) == -8
    assert L.linompss_gp_grad_dev(4, 3, p, 4, 0.3, 1.0, p, 4, None, None, 4, None) == -9
    assert L.linompss_gp_grad_dev(4, 3, p, 4, 0.3, 1.0, p, 4, g, p, 3, None) == -11


@pytest.mark.parametrize("n,sf2,b", [(300, 1.0, 256), (1000, 1.7, 128), (2048, 0.6, 512)])
def test_gp_device_gradients(lx, dev, n, sf2, b):
    import torch
    from oracle import adjoint as adj
    import linompss_amd.autograd as ag
    print(f"[gp device] n={n} sf2={sf2} b={b}", flush=True)
    ell0, noise = 0.3, 1e-2
    P, y = _problem(n, 7, ell0, noise, sf2)
    pts = torch.from_numpy(P).requires_grad_(True)
    yt = torch.from_numpy(y).requires_grad_(True)
    ell = torch.tensor(ell0, dtype=torch.float64, requires_grad=True)
    s = torch.tensor(sf2, dtype=torch.float64, requires_grad=True)
    nz = torch.tensor(noise, dtype=torch.float64, requires_grad=True)
    tm = {}
    loss = ag.gp_nll_device(pts, yt, ell, s, nz, b=b, timing=tm)
    loss.backward()
    assert tm["forward_s"] > 0 and tm["backward_s"] > 0
    ref = adj.gp_nll_grad_all(P, y, ell0, noise, sf2)
    assert abs(loss.item() - adj.gp_nll(P, y, ell0, noise, sf2)) <= 1e-10 * abs(loss.item())
    for name, t in (("dell", ell), ("dsf2", s), ("dnoise", nz)):
        assert abs(t.grad.item() - ref[name]) <= 1e-9 * max(1.0, abs(ref[name])), name
    assert np.abs(pts.grad.numpy() - ref["dP"]).max() <= 1e-9 * np.abs(ref["dP"]).max()
    assert np.abs(yt.grad.numpy() - ref["dy"]).max() <= 1e-9 * np.abs(ref["dy"]).max()


def test_gp_device_scalars_as_floats_and_reuse(lx, dev):
    """Plain floats for the parameters (no gradient asked), two problem sizes
    in a row (the workspace is replaced), and a repeat bit for bit."""
    import torch
    import linompss_amd.autograd as ag
    grads = []
    for n in (500, 260, 260):
        P, y = _problem(n, 9, 0.3, 1e-2)
        ell = torch.tensor(0.3, dtype=torch.float64, requires_grad=True)
        loss = ag.gp_nll_device(torch.from_numpy(P), torch.from_numpy(y), ell, 1.0, 1e-2)
        loss.backward()
        grads.append(ell.grad.item())
    assert grads[1] == grads[2] and grads[0] != grads[1]
    ag.gp_release_workspace()


def test_gp_device_interleaved_passes(lx, dev):
    """Backward reuses the forward's factor from the shape's workspace only
    while no other forward of that shape has run: two graphs of one shape
    with their forwards first (the first backward recomputes its factor), and
    a retained second backward, give the gradients of a forward followed at
    once by its backward, bit for bit."""
    import torch
    import linompss_amd.autograd as ag
    n = 700
    print(f"[gp interleaved] n={n}", flush=True)
    P, y = _problem(n, 11, 0.3, 1e-2)

    def run(ell0):
        ell = torch.tensor(ell0, dtype=torch.float64, requires_grad=True)
        return ell, ag.gp_nll_device(torch.from_numpy(P), torch.from_numpy(y), ell, 1.0, 1e-2)

    ref = {}
    for ell0 in (0.3, 0.35):
        ell, loss = run(ell0)
        loss.backward()
        ref[ell0] = (loss.item(), ell.grad.item())
    (ea, la), (eb, lb) = run(0.3), run(0.35)
    la.backward(retain_graph=True)  # the workspace holds the 0.35 forward: recomputed
    lb.backward()                   # and now the 0.3 one: recomputed again
    assert (la.item(), ea.grad.item()) == ref[0.3]
    assert (lb.item(), eb.grad.item()) == ref[0.35]
    ea.grad = None
    la.backward()                   # the workspace still holds the 0.3 pass: reused
    assert ea.grad.item() == ref[0.3][1]
    ag.gp_release_workspace()
