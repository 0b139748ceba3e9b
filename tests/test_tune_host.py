"""Recorded factor configurations (the autotuner's output) on the host: the
defaults a shape uses, set / get / clear, argument checks, and that a recorded
configuration changes the plan the executor would run (unit counts from the
planner, no device)."""
import ctypes

import numpy as np
import pytest


def test_defaults_and_roundtrip(lx):
    lx.clear_factor_configs()
    rec, d64 = lx.get_factor_config(32768, 512)
    assert not rec and d64.b == 512 and d64.ms == 0.0
    assert (d64.potrf_sub, d64.gemm_wide_m, d64.gemm_wide_n, d64.trsm_rows_wide, d64.wide_tail,
            d64.update_group) == (256, 256, 512, 64, -1, 4)
    # FP64 tiles below 512: 128-wide diagonal sub-tiles
    assert lx.get_factor_config(8192, 256)[1].potrf_sub == 128
    assert lx.get_factor_config(8192, 384)[1].potrf_sub == 128
    rec, d32 = lx.get_factor_config(32768, 512, np.float32, update_prec=2)
    assert not rec and (d32.potrf_sub, d32.gemm_wide_m, d32.trsm_rows_wide, d32.wide_tail, d32.update_group) == \
        (128, 512, 128, 0, 4)
    c = lx.FactorConfig(1024, 512, 128, 1024, 128, 8, 1, 201.5)
    lx.set_factor_config(32768, c)
    rec, got = lx.get_factor_config(32768, 1024)
    assert rec and got.as_dict() == c.as_dict()
    # other shapes are untouched: another b, n, variant, dtype or update arithmetic
    assert not lx.get_factor_config(32768, 512)[0]
    assert not lx.get_factor_config(16384, 1024)[0]
    assert not lx.get_factor_config(32768, 1024, variant="ll")[0]
    assert not lx.get_factor_config(32768, 1024, np.float32)[0]
    assert lx.get_factor_config(32768, 1024)[1].update_group == 1  # recorded: c
    assert lx.get_factor_config(32768, 2048, np.float32)[1].update_group == 1  # default: 2048 / b
    assert lx.get_factor_config(32768, 256, np.float32)[1].update_group == 8
    c.wide_tail = 0
    lx.set_factor_config(32768, c)  # replaces the entry
    assert lx.get_factor_config(32768, 1024)[1].wide_tail == 0
    lx.clear_factor_configs()
    assert not lx.get_factor_config(32768, 1024)[0]


def test_argument_checks(lx):
    L = lx.lib
    ok = lx.FactorConfig(512, 256, 256, 512, 64, -1, 2, 0.0)
    assert L.linompss_set_factor_config(2, 0, 100, 0, ctypes.byref(ok)) == -1
    assert L.linompss_set_factor_config(0, 2, 100, 0, ctypes.byref(ok)) == -2
    assert L.linompss_set_factor_config(0, 0, 0, 0, ctypes.byref(ok)) == -3
    assert L.linompss_set_factor_config(0, 0, 100, 2, ctypes.byref(ok)) == -4   # FP64 has no split updates
    assert L.linompss_set_factor_config(0, 0, 100, 0, None) == -5
    for bad in (lx.FactorConfig(0, 256, 256, 512, 64, -1, 1, 0), lx.FactorConfig(512, 96, 256, 512, 64, -1, 1, 0),
                lx.FactorConfig(512, 256, 192, 512, 64, -1, 1, 0), lx.FactorConfig(512, 256, 256, 0, 64, -1, 1, 0),
                lx.FactorConfig(512, 256, 256, 512, 32, -1, 1, 0), lx.FactorConfig(512, 256, 256, 512, 64, -2, 1, 0),
                lx.FactorConfig(512, 256, 256, 512, 64, -1, 0, 0), lx.FactorConfig(512, 256, 256, 512, 64, -1, 9, 0)):
        assert L.linompss_set_factor_config(0, 0, 100, 0, ctypes.byref(bad)) == -5
    out = lx.FactorConfig()
    assert L.linompss_get_factor_config(0, 0, 100, 0, 0, ctypes.byref(out)) == -4
    assert L.linompss_get_factor_config(0, 0, 100, 64, 0, None) == -6
    best = lx.FactorConfig()
    assert L.linompss_tune_factor(3, 0, 100, 0, 0, 1, ctypes.byref(best), None, 0, None) == -1
    assert L.linompss_tune_factor(0, 0, 100, -1, 0, 1, ctypes.byref(best), None, 0, None) == -4
    assert L.linompss_tune_factor(0, 0, 100, 0, 1, 1, ctypes.byref(best), None, 0, None) == -5
    assert L.linompss_tune_factor(0, 0, 100, 0, 0, 1, None, None, 0, None) == -7
    assert L.linompss_tune_factor(0, 0, 100, 0, 0, 1, ctypes.byref(best), None, 4, None) == -9


def test_recorded_config_reaches_the_planner(lx):
    """plan_info builds the plan the executor would run for a shape; with a
    recorded configuration of 128-row update units the plan has more units
    than with the defaults' taller ones."""
    lx.clear_factor_configs()
    rec, d = lx.get_factor_config(8192, 512)
    assert not rec
    u0 = lx.plan_info(0, 8192, 512)[0]
    lx.set_factor_config(8192, lx.FactorConfig(512, d.potrf_sub, 128, 128, 64, 0, 1, 0.0))
    try:
        u1 = lx.plan_info(0, 8192, 512)[0]
    finally:
        lx.clear_factor_configs()
    assert u1 > u0


def test_grouped_updates_fewer_units_same_tasks(lx):
    """Grouped trailing updates cut the plan's units, never its tasks: the task
    graph (chol_rlmain.c:25-50) is the same for every group size."""
    lx.clear_factor_configs()
    rec, d = lx.get_factor_config(8192, 512)
    units = []
    try:
        for g in (1, 2, 3, 4):
            d.update_group = g
            lx.set_factor_config(8192, d)
            u, t, _ = lx.plan_info(0, 8192, 512)
            assert t == 816  # mt = 16: 16 POTRF + 120 TRSM + 120 SYRK + 560 GEMM
            units.append(u)
    finally:
        lx.clear_factor_configs()
    assert units[0] > units[1] > units[3]
