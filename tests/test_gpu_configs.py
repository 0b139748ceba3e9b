"""The workload configurations at their full sizes, through the C-ABI.

* FP32-factor + FP64 refinement at n = 32768, b = 512, cond2 in {1e2, 1e4,
  1e6}, 1 and 16 right-hand sides uniform in [-1, 1] (the headline's
  configuration): backward error <= 1e-14 as reported by the library AND as
  recomputed on the host in FP64 from panels of A, no FP64 fallback, and the
  solution within what the two residuals allow of the plain FP64 factor +
  solve's: ||x - x_fp64||_2 <= (||r||_2 + ||r_fp64||_2) / lambda_min(A).
* n = 65536, b = 1024, cond2 ~ 1e4 (seed 4), 8 right-hand sides, split
  low-precision trailing updates: the same bars.
* The Gaussian-process likelihood gradient at 8192 points (seed 3) against
  the FP64 closed form of the oracle and a central difference (h = 1e-6).

Refinement contract: src/itref/itref.c:999-1101; factor: chol_rlmain.c:22-53;
sweeps: bblas_trsm.h:116-160.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def host_berr(lx, dA, Xs, Bs, panel=2048):
    """max_c ||B_c - A X_c||_inf / (||A||_inf ||X_c||_inf) in FP64 on the host,
    A streamed from the device in column panels (both triangles stored by the
    generator), for several solution sets at once."""
    n = dA.rows
    Rs = [B.copy() for B in Bs]
    rows = np.zeros(n)
    for c0 in range(0, n, panel):
        c1 = min(n, c0 + panel)
        P = dA.cols_to_host(c0, c1)
        rows += np.abs(P).sum(axis=1)
        for R, X in zip(Rs, Xs):
            R -= P @ X[c0:c1]
    an = rows.max()
    out = []
    for R, X in zip(Rs, Xs):
        out.append(max(np.abs(R[:, c]).max() / (an * np.abs(X[:, c]).max()) for c in range(X.shape[1])))
    return out, [np.linalg.norm(R, axis=0) for R in Rs]


def _solve_and_check(lx, n, b, seed, kappa, nrhs_list, update_prec=2):
    print(f"[config] n={n} b={b} seed={seed} kappa={kappa:g} nrhs={nrhs_list}", flush=True)
    dA, alpha, lmax = lx.gen_gauss_spd(n, seed, kappa)
    nmax = max(nrhs_list)
    Bmax = lx.gen_uniform(n, nmax, seed + 100).to_host()
    Xs, Bs = [], []
    for nr in nrhs_list:
        B = np.asfortranarray(Bmax[:, :nr])
        dB = lx.DeviceMatrix.from_host(B)
        dX = lx.DeviceMatrix(n, nr)
        st = lx.dposv_ir_dev(dA, dB, dX, b, lx.IrOpts(30, 1e-14, 0.5, update_prec))
        print(f"  nrhs={nr}: sweeps={st.sweeps} fallback={st.fallback} berr={st.berr:.2e} "
              f"factor={st.t_factor_s * 1e3:.1f} ms", flush=True)
        assert st.info == 0
        assert st.fallback == 0
        assert st.berr <= 1e-14
        assert st.sweeps <= 5
        Xs.append(dX.to_host())
        Bs.append(B)
        dX.free()
        dB.free()
    # the plain FP64 factor + solve of the same systems (the widest RHS set)
    dL = lx.DeviceMatrix(n, n)
    lx.copy_dev(dL, dA)
    dXd = lx.DeviceMatrix.from_host(Bmax)
    assert lx.posv_dev(dL, dXd, b) == 0
    dL.free()
    Xd = dXd.to_host()
    berrs, rnorms = host_berr(lx, dA, Xs + [Xd], Bs + [Bmax])
    print(f"  host berr: {['%.2e' % e for e in berrs]}", flush=True)
    assert all(e <= 1e-14 for e in berrs)
    for X, r in zip(Xs, rnorms):
        nr = X.shape[1]
        # A (x - xd) = r_d - r and lambda_min(A) >= alpha (A = G G^T / n + alpha I):
        # ||x - xd||_2 <= (||r||_2 + ||r_d||_2) / alpha, column by column
        d = np.linalg.norm(X - Xd[:, :nr], axis=0)
        bound = (r + rnorms[-1][:nr]) / alpha
        print(f"  ||x - x_fp64||_2 = {d.max():.2e}, bound {bound.min():.2e}", flush=True)
        assert np.all(d <= 1.01 * bound)


@pytest.mark.timeout(900)
@pytest.mark.parametrize("kappa", [1e2, 1e4, 1e6])
def test_refinement_n32768(lx, dev, kappa):
    _solve_and_check(lx, 32768, 512, 1, kappa, [1, 16])


@pytest.mark.timeout(1200)
def test_refinement_n65536_b1024(lx, dev):
    _solve_and_check(lx, 65536, 1024, 4, 1e4, [8])


@pytest.mark.timeout(900)
def test_gp_gradient_8192(lx, dev):
    """dL/dl of the GP negative log marginal likelihood at the workload's full
    size: 8192 points uniform in [0,1]^3 (seed 3), l = 0.3, noise 1e-2,
    y = chol(K) z, through the torch.autograd harness (K resident on the
    device, factor recomputed in backward)."""
    import torch
    from oracle import adjoint as adj
    import linompss_amd.autograd as ag
    n = 8192
    print(f"[gp] n={n}", flush=True)
    rng = np.random.default_rng(3)
    P = rng.uniform(0, 1, (n, 3))
    ell0, noise = 0.3, 1e-2
    y = np.linalg.cholesky(adj.gp_kernel(P, ell0, noise)) @ rng.standard_normal(n)
    pts = torch.from_numpy(P)
    yt = torch.from_numpy(y)
    ell = torch.tensor(ell0, dtype=torch.float64, requires_grad=True)
    loss = ag.gp_nll_fused(pts, yt, ell, noise)
    loss.backward()
    g = ell.grad.item()
    dK_ref, dell_ref = adj.gp_nll_grad(P, y, ell0, noise)
    print(f"  nll={loss.item():.12e} dl={g:.12e} ref={dell_ref:.12e}", flush=True)
    assert abs(loss.item() - adj.gp_nll(P, y, ell0, noise)) <= 1e-10 * abs(loss.item())
    assert abs(g - dell_ref) <= 1e-9 * abs(dell_ref)
    h = 1e-6
    with torch.no_grad():
        fp = ag.gp_nll_fused(pts, yt, torch.tensor(ell0 + h, dtype=torch.float64), noise).item()
        fm = ag.gp_nll_fused(pts, yt, torch.tensor(ell0 - h, dtype=torch.float64), noise).item()
    fd = (fp - fm) / (2 * h)
    print(f"  central difference {fd:.12e}", flush=True)
    assert abs(g - fd) <= 1e-5 * abs(fd)


@pytest.mark.timeout(900)
def test_gp_device_gradients_8192(lx, dev):
    """The same workload through GPMarginalLikelihood (K built on the device
    from the points): the value and every gradient against the oracle's FP64
    closed forms (pinned by central differences in test_adjoint_oracle.py)."""
    import torch
    from oracle import adjoint as adj
    import linompss_amd.autograd as ag
    n = 8192
    print(f"[gp device] n={n}", flush=True)
    rng = np.random.default_rng(3)
    P = rng.uniform(0, 1, (n, 3))
    ell0, noise = 0.3, 1e-2
    y = np.linalg.cholesky(adj.gp_kernel(P, ell0, noise)) @ rng.standard_normal(n)
    pts = torch.from_numpy(P).requires_grad_(True)
    yt = torch.from_numpy(y).requires_grad_(True)
    ell = torch.tensor(ell0, dtype=torch.float64, requires_grad=True)
    nz = torch.tensor(noise, dtype=torch.float64, requires_grad=True)
    loss = ag.gp_nll_device(pts, yt, ell, 1.0, nz)
    loss.backward()
    ref = adj.gp_nll_grad_all(P, y, ell0, noise)
    print(f"  dl={ell.grad.item():.12e} ref={ref['dell']:.12e}", flush=True)
    assert abs(loss.item() - adj.gp_nll(P, y, ell0, noise)) <= 1e-10 * abs(loss.item())
    assert abs(ell.grad.item() - ref["dell"]) <= 1e-9 * abs(ref["dell"])
    assert abs(nz.grad.item() - ref["dnoise"]) <= 1e-9 * abs(ref["dnoise"])
    assert np.abs(pts.grad.numpy() - ref["dP"]).max() <= 1e-8 * np.abs(ref["dP"]).max()
    assert np.abs(yt.grad.numpy() - ref["dy"]).max() <= 1e-9 * np.abs(ref["dy"]).max()
    ag.gp_release_workspace()
