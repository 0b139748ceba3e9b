"""The factor autotuner on the device: it times candidate tile sizes and unit
shapes, records the fastest for its shape, and the factors that then run with
the recorded configuration still match LAPACK (the configuration changes how
tile tasks are cut into work units, never the arithmetic's result beyond
rounding).  Reference task graph: chol_rlmain.c:25-50, chol_llmain.c:26-53."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _clean(lx):
    lx.clear_factor_configs()
    yield
    lx.clear_factor_configs()


@pytest.mark.parametrize("dtype,variant,update_prec", [(np.float64, "rl", 0), (np.float64, "ll", 0),
                                                       (np.float32, "rl", 2)])
def test_tune_records_and_factors_stay_exact(lx, orc, dev, dtype, variant, update_prec):
    n = 2048
    print(f"[tune] n={n} dtype={np.dtype(dtype)} {variant} up={update_prec}", flush=True)
    best, tried = lx.tune_factor(n, 0, dtype, variant, update_prec, reps=1)
    tiles = {128, 256, 384, 512, 768, 1024}  # the small tiles are searched up to n = 16384
    assert best.b in tiles and best.ms > 0
    assert len(tried) >= 10 and min(t["ms"] for t in tried) == pytest.approx(best.ms)
    assert {t["b"] for t in tried} == tiles
    rec, got = lx.get_factor_config(n, best.b, dtype, variant, update_prec)
    assert rec and got.as_dict() == best.as_dict()
    for b in tiles - {best.b}:  # losing tile sizes leave nothing behind
        assert not lx.get_factor_config(n, b, dtype, variant, update_prec)[0]
    # a factor of that shape now runs with the recorded configuration
    A0 = orc.genmat_spd(n, seed=5, diag_mode=1)
    Lref = np.linalg.cholesky(A0)
    if dtype == np.float64:
        A = A0.copy(order="F")
        fn = lx.ompss_dchol_rl if variant == "rl" else lx.ompss_dchol_ll
        assert fn(A, best.b) == 0
        assert np.linalg.norm(np.tril(A) - Lref) / np.linalg.norm(Lref) <= 1e-12
        assert np.array_equal(np.triu(A, 1), np.triu(A0, 1))
    else:
        # the split-FP16 updates need |L| <= 16: the refinement driver's equilibrated form
        d = 2.0 ** -np.ceil(np.log2(np.sqrt(np.diag(A0))))
        As = (A0 * d[:, None] * d[None, :]).astype(np.float32, order="F")
        dA = lx.DeviceMatrix.from_host(As)
        assert lx.chol_dev(dA, best.b, variant, lx.ExecOpts(1, 128, 128, 0, 0.0, 0, 2)) == 0
        L = np.tril(dA.to_host().astype(np.float64)) / d[:, None]
        assert np.linalg.norm(L - Lref) / np.linalg.norm(Lref) <= 1e-5
        dA.free()


def test_tune_fixed_tile_and_prior_kept(lx, dev):
    """b > 0 keeps the tile; a configuration recorded for another tile size is untouched."""
    other = lx.FactorConfig(1024, 128, 128, 256, 64, 0, 1, 1.0)
    lx.set_factor_config(1536, other)
    best, tried = lx.tune_factor(1536, 384, np.float64, "rl", 0, reps=1)
    assert best.b == 384 and all(t["b"] == 384 for t in tried)
    assert lx.get_factor_config(1536, 1024)[1].as_dict() == other.as_dict()
