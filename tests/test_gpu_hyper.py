"""GPU parity of the hypermatrix factor (ompss_{d,s}chol_{hrl,hll,nhrl,nhll},
src/cholesky/dense/chol_hrlmain.c:16-37, chol_hllmain.c:17-45): tiles as
separate b x b buffers, tile (i, j) of the factor in place.

Bars: each tile against the oracle's CHOL_HRL / CHOL_HLL on the same tiles and
against LAPACK on the assembled matrix (FP64 1e-12 relative, FP32 1e-6); the
tiled run against the column-major factor of the same matrix bit for bit (the
same plan, unit bodies and arithmetic order, only the addressing differs); the
strict upper part of each diagonal tile untouched; info = the first
non-positive global column.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _fn(lx, dt, variant):
    p = "d" if dt == np.float64 else "s"
    return getattr(lx, f"ompss_{p}chol_{variant}")


@pytest.mark.parametrize("variant", ["hrl", "hll", "nhrl", "nhll"])
@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("mt,b", [(1, 64), (3, 100), (4, 256), (5, 48)])
def test_hypermatrix_vs_oracle(lx, orc, dev, dt, variant, mt, b):
    print(f"[hyper {variant}] {np.dtype(dt).name} mt={mt} b={b}", flush=True)
    n = mt * b
    A0 = orc.genmat_spd(n, 7, diag_mode=1)
    tiles = orc.to_tiles(A0.astype(dt), b)
    ref = {k: v.copy(order="F") for k, v in tiles.items()}
    assert orc.chol_tiles(ref, mt, b, "hrl" if variant.endswith("rl") else "hll") == 0
    assert _fn(lx, dt, variant)(tiles, mt, b, t=16) == 0
    L = np.linalg.cholesky(A0)
    tol = 1e-12 if dt == np.float64 else 1e-6
    got = np.tril(orc.from_tiles(tiles, mt, b))
    assert np.linalg.norm(got - L) <= tol * np.linalg.norm(L)
    assert np.linalg.norm(got - np.tril(orc.from_tiles(ref, mt, b))) <= tol * np.linalg.norm(L)
    for k in range(mt):
        assert np.array_equal(np.triu(tiles[(k, k)], 1), np.triu(A0[k * b:(k + 1) * b, k * b:(k + 1) * b].astype(dt), 1))


@pytest.mark.parametrize("variant", ["rl", "ll"])
def test_hypermatrix_equals_colmajor(lx, orc, dev, variant):
    """n = 2048 in 512-wide tiles: the tiled factor and ompss_dchol_{rl,ll} on the
    assembled matrix agree bit for bit when the column-major plan is the tiled
    one (no paired updates: a hypermatrix plan never lets a unit span two
    panels), and within rounding with the column-major default."""
    print(f"[hyper vs colmajor {variant}]", flush=True)
    mt, b = 4, 512
    n = mt * b
    A0 = orc.genmat_spd(n, 8, diag_mode=1)
    fn = lx.ompss_dchol_rl if variant == "rl" else lx.ompss_dchol_ll
    A = A0.copy(order="F")
    assert fn(A, b) == 0  # the default plan
    lx.clear_factor_configs()
    _, cfg = lx.get_factor_config(n, b, np.float64, variant)
    cfg.update_group = 1
    lx.set_factor_config(n, cfg, np.float64, variant)
    try:
        As = A0.copy(order="F")
        assert fn(As, b) == 0
    finally:
        lx.clear_factor_configs()
    tiles = orc.to_tiles(A0, b)
    assert (lx.ompss_dchol_hrl if variant == "rl" else lx.ompss_dchol_hll)(tiles, mt, b) == 0
    L = np.tril(orc.from_tiles(tiles, mt, b))
    assert np.array_equal(L, np.tril(As))
    assert np.linalg.norm(L - np.tril(A)) <= 1e-14 * np.linalg.norm(L)


def test_hypermatrix_info_and_arguments(lx, orc, dev):
    print("[hyper info]", flush=True)
    A = orc.genmat_spd(300, 2, diag_mode=1)
    A[150, 150] = -1.0
    for fn in (lx.ompss_dchol_hrl, lx.ompss_dchol_hll):
        assert fn(orc.to_tiles(A, 60), 5, 60) == 151
    assert lx.lib.ompss_dchol_hrl(0, 64, 0, None) == 0
    assert lx.lib.ompss_dchol_hrl(-1, 64, 0, None) == -1
    assert lx.lib.ompss_dchol_hrl(2, 0, 0, None) == -2
    assert lx.lib.ompss_dchol_hrl(2, 64, 0, None) == -4
    with pytest.raises(ValueError):
        lx.ompss_dchol_hrl({(0, 0): np.eye(4, order="F")}, 2, 4)  # lower tiles missing


def test_chol_tiles_dev(lx, orc, dev):
    """The device entry point on tiles allocated separately in device memory."""
    print("[hyper dev]", flush=True)
    mt, b = 3, 128
    A0 = orc.genmat_spd(mt * b, 9, diag_mode=1)
    host = orc.to_tiles(A0, b)
    dtiles = {k: lx.DeviceMatrix.from_host(v) for k, v in host.items()}
    try:
        assert lx.chol_tiles_dev(dtiles, mt, b, "rl") == 0
        got = {k: d.to_host() for k, d in dtiles.items()}
    finally:
        for d in dtiles.values():
            d.free()
    L = np.linalg.cholesky(A0)
    assert np.linalg.norm(np.tril(orc.from_tiles(got, mt, b)) - L) <= 1e-12 * np.linalg.norm(L)
