"""GPU parity of the workload generators and of the device backward error.

The headline's input matrix and its accuracy figure come from device code:
``linompss_gen_*`` (Philox4x32-10 streams, the Gram matrix, the Lanczos
estimate of lambda_max that sets alpha) and ``linompss_dberr_dev``.  Each is checked here
against the oracle (``oracle.gen_uniform`` / ``gen_gauss`` /
``genmat_covariant`` / ``backward_error``, themselves pinned on the CPU by
Random123's known answers and closed forms) and against numpy's eigenvalues
for the conditioning the generator promises.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("m,n,seed", [(1, 1, 0), (257, 3, 1), (1000, 17, 123456)])
def test_gen_uniform_bit_exact(lx, orc, dev, m, n, seed):
    """U[lo, hi) from Philox4x32-10: integer stream and the affine map are
    exact, so the device equals the oracle bit for bit (FP32: the FP64 value
    rounded once)."""
    print(f"[gen uniform] {m}x{n} seed={seed}", flush=True)
    ref = orc.gen_uniform(m, n, seed, -1.0, 1.0)
    assert np.array_equal(lx.gen_uniform(m, n, seed).to_host(), ref)
    assert np.array_equal(lx.gen_uniform(m, n, seed, dtype=np.float32).to_host(), ref.astype(np.float32))
    ref2 = orc.gen_uniform(m, n, seed, 0.25, 3.5)
    assert np.array_equal(lx.gen_uniform(m, n, seed, 0.25, 3.5).to_host(), ref2)


@pytest.mark.parametrize("m,n,seed", [(255, 3, 1), (1000, 16, 7)])
def test_gen_gauss_vs_oracle(lx, orc, dev, m, n, seed):
    """N(0, 1) by Box-Muller over the same Philox blocks: the device's log,
    sqrt, cos and sin may differ from the host libm in the last bits, so the
    values agree to a few units in the last place."""
    print(f"[gen gauss] {m}x{n} seed={seed}", flush=True)
    ref = orc.gen_gauss(m, n, seed)
    got = lx.gen_gauss(m, n, seed).to_host()
    assert np.abs(got - ref).max() <= 1e-14
    assert np.mean(got == ref) > 0.5  # most values bit-identical


@pytest.mark.parametrize("n,kappa", [(600, 1e2), (1024, 1e4)])
def test_gen_gauss_spd_matches_oracle(lx, orc, dev, n, kappa):
    """A = G G^T / n + alpha I with G the oracle's stream (seed 1), alpha =
    lambda_max / (kappa - 1) with lambda_max (Lanczos) within 1e-10 of
    numpy's largest eigenvalue of G G^T / n."""
    print(f"[gen spd] n={n} kappa={kappa:g}", flush=True)
    dA, alpha, lmax = lx.gen_gauss_spd(n, 1, kappa)
    A = dA.to_host()
    G = orc.gen_gauss(n, n, 1)
    W = G @ G.T / n
    lam = np.linalg.eigvalsh(W)
    assert abs(lmax - lam[-1]) <= 1e-10 * lam[-1]
    assert alpha == lmax / (kappa - 1.0)
    ref = W + alpha * np.eye(n)
    assert np.abs(A - ref).max() <= 1e-13 * np.abs(ref).max()
    assert np.array_equal(A, A.T)


@pytest.mark.parametrize("kappa", [1e2, 1e4, 1e6])
def test_gen_gauss_spd_condition(lx, dev, kappa):
    """cond2(A) = (lambda_max + alpha) / (lambda_min + alpha) = kappa /
    (1 + lambda_min / alpha): within 1 % of kappa where lambda_min / alpha is
    negligible (1e2, 1e4) and within the predicted deficit at 1e6, where
    lambda_min of a square Wishart matrix (~1/n^2) is a few percent of alpha."""
    n = 2048
    print(f"[gen cond] n={n} kappa={kappa:g}", flush=True)
    dA, alpha, lmax = lx.gen_gauss_spd(n, 1, kappa)
    lam = np.linalg.eigvalsh(dA.to_host())
    cond = lam[-1] / lam[0]
    # the spectrum is that of G G^T / n shifted by alpha: the largest
    # eigenvalue is lambda_max + alpha with lambda_max the Lanczos estimate
    assert abs(lam[-1] - (lmax + alpha)) <= 1e-10 * lam[-1]
    assert 0.0 < lam[0] - alpha < alpha  # lambda_min of the Wishart part: positive, below alpha
    assert abs(cond - kappa * lam[-1] / (lmax + alpha) / (1 + (lam[0] - alpha) / alpha)) <= 1e-6 * cond
    assert cond <= kappa * (1 + 1e-8)
    if kappa <= 1e4:
        assert cond >= 0.99 * kappa
    else:
        assert cond >= 0.8 * kappa


@pytest.mark.parametrize("n,p,k", [(1, 2.0, 2.0), (300, 2.0, 2.0), (1000, 3.0, 1.5)])
def test_gen_covariant_vs_oracle(lx, orc, dev, n, p, k):
    """GENMAT_COVARIANT (genmat.c:263-302) on the device against the oracle:
    pow on either side may differ in the last place."""
    print(f"[gen covar] n={n} p={p} k={k}", flush=True)
    got = lx.gen_covariant(n, p, k, ld=n + 3).to_host()
    ref = orc.genmat_covariant(n, p, k)
    assert np.all(np.abs(got - ref) <= 2 * np.spacing(np.abs(ref)))
    assert np.array_equal(got, got.T)
    with pytest.raises(lx.LinompssError):
        lx.gen_covariant(4, 0.0, 2.0)


@pytest.mark.parametrize("n,nr,pad", [(1, 1, 0), (700, 1, 5), (1000, 4, 0), (2049, 16, 3), (1500, 21, 2)])
def test_dberr_dev_vs_oracle(lx, orc, dev, n, nr, pad):
    """linompss_dberr_dev (the headline's accuracy figure) against the
    oracle's max_c ||b_c - A x_c||_inf / (||A||_inf ||x_c||_inf), for one and
    many right-hand sides (21: past the old 16-column limit), lda > n, and a
    column whose residual is exactly zero (b = 0, x = 0: berr 0, not 0/0)."""
    print(f"[dberr] n={n} nr={nr} pad={pad}", flush=True)
    rng = np.random.default_rng(n + nr)
    A = orc.genmat_spd(n, 2, diag_mode=1)
    X = np.asfortranarray(rng.uniform(-1, 1, (n, nr)))
    B = np.asfortranarray(A @ X + 1e-6 * rng.standard_normal((n, nr)))
    if nr > 1:
        X[:, 1] = 0.0
        B[:, 1] = 0.0
    Ap = np.zeros((n + pad, n), order="F")
    Ap[:n] = A
    Ap[np.triu_indices(n, 1)] = np.nan  # only the lower triangle is read
    dA = lx.DeviceMatrix(n, n, ld=n + pad)
    lx.lib.linompss_memcpy(dA.ptr, Ap.ctypes.data, Ap.nbytes, 1)
    got = lx.dberr_dev(dA, lx.DeviceMatrix.from_host(X), lx.DeviceMatrix.from_host(B))
    ref = orc.backward_error(A, X, B)
    assert np.isfinite(got)
    # the residuals of two summation orders differ by at most n eps |A| |x|,
    # so the backward errors by at most n eps (||(|A| |x|)||_inf <= ||A||_inf ||x||_inf)
    assert abs(got - ref) <= 1e-10 * ref + 2 * n * np.finfo(float).eps
    assert got > 1e3 * n * np.finfo(float).eps  # the residual is far above rounding level
    if nr > 1:
        Z = np.zeros((n, 1), order="F")
        assert lx.dberr_dev(dA, lx.DeviceMatrix.from_host(Z), lx.DeviceMatrix.from_host(Z)) == 0.0
