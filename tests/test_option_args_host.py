"""Argument checks of the option entry points, on the host: the library
refuses a matching mode out of range, the standardized modes without their
null table, an unknown standardized distance, a missing per-reference
threshold array and an unknown Wiener gain before any device work (dummy
device pointers are never dereferenced)."""

import ctypes

import pytest

from bmnd_amd import _lib

I32X3, I64X3 = ctypes.c_int32 * 3, ctypes.c_int64 * 3
DUMMY = ctypes.c_void_p(16)
SHAPE, BLOCK, RADIUS = I64X3(24, 24, 24), I32X3(4, 4, 4), I32X3(7, 7, 7)


def _error(rc):
    assert rc != 0
    return _lib.lib().bmnd_hip_last_error().decode()


@pytest.mark.parametrize("mode", [-1, 10])
def test_mode_out_of_range(mode):
    L = _lib.lib()
    rc = L.bmnd_hip_blockmatch(DUMMY, SHAPE, BLOCK, RADIUS, DUMMY, 10, 1.0, 16, 2, 1, 1, 1.0,
                               mode, 0, DUMMY, DUMMY, DUMMY, None)
    assert "mode must be in 0..9" in _error(rc)


@pytest.mark.parametrize("mode", [7, 8, 9])
def test_standardized_modes_need_the_table(mode):
    L = _lib.lib()
    rc = L.bmnd_hip_blockmatch(DUMMY, SHAPE, BLOCK, RADIUS, DUMMY, 10, 1.0, 16, 2, 1, 1, 1.0,
                               mode, 0, DUMMY, DUMMY, DUMMY, None)
    assert "null-moment table" in _error(rc)
    rc = L.bmnd_hip_blockmatch_standardized(DUMMY, SHAPE, BLOCK, RADIUS, DUMMY, 10, 1.0, 16, 2,
                                            1, 1, 1.0, mode - 6, DUMMY, 0, DUMMY, DUMMY, DUMMY,
                                            None)
    assert "null-moment table" in _error(rc)


@pytest.mark.parametrize("distance", [0, 4])
def test_standardized_distance_codes(distance):
    rc = _lib.lib().bmnd_hip_blockmatch_standardized(
        DUMMY, SHAPE, BLOCK, RADIUS, DUMMY, 10, 1.0, 16, 2, 1, 1, 1.0, distance, DUMMY, 512,
        DUMMY, DUMMY, DUMMY, None)
    assert "distance must be 1" in _error(rc)


def test_thresholds_entry_needs_the_array():
    rc = _lib.lib().bmnd_hip_blockmatch_thresholds(
        DUMMY, SHAPE, BLOCK, RADIUS, DUMMY, None, 10, 16, 2, 1, 1, 1.0, 4, 0, DUMMY, DUMMY,
        DUMMY, None)
    assert "ref_thresholds must not be null" in _error(rc)


@pytest.mark.parametrize("code,msg", [
    (-1, "unknown bits"), (2048, "unknown bits"), (3 << 3, "weight model must be"),
    (1024, "matrix-core member transforms are a Wiener-stage option"),
    (2 << 3, "risk weight model is a Wiener-stage option"),
    (1 | 4, "soft thresholding is a hard-threshold-stage option"),
    (512, "the exact group risk needs the variance"),
    (512 | 32 | 64 | (1 << 3), "the exact group risk needs the variance"),
    (512 | 32 | 64, "k > 0"),
    (512 | 32 | 64 | 256, "the Poisson model")])
def test_filter_code_argument(code, msg):
    rc = _lib.lib().bmnd_hip_filter_groups(
        code, DUMMY, DUMMY, DUMMY, SHAPE, BLOCK, RADIUS, DUMMY, DUMMY, 10, 16, DUMMY, DUMMY,
        3.0, 0, 1.0, DUMMY, DUMMY, None)
    assert msg in _error(rc)


def test_exact_risk_entry_checks_its_workspace():
    L = _lib.lib()
    need = L.bmnd_hip_exact_risk_workspace(10, 16, 4, BLOCK)
    assert need >= 10 * 4 * 64 * 4 + 10 * 8 and need % 256 == 0
    assert L.bmnd_hip_exact_risk_workspace(10, 16, 0, BLOCK) == -1
    rc = L.bmnd_hip_filter_groups_exact(
        32 | 64, DUMMY, DUMMY, DUMMY, SHAPE, BLOCK, RADIUS, DUMMY, DUMMY, 10, 16, DUMMY, DUMMY,
        3.0, 4, 1.0, DUMMY, DUMMY, None, DUMMY, need - 1, None)
    assert "smaller than bmnd_hip_exact_risk_workspace" in _error(rc)
    rc = L.bmnd_hip_filter_groups_exact(
        32 | 64, DUMMY, DUMMY, DUMMY, SHAPE, BLOCK, RADIUS, DUMMY, DUMMY, 10, 16, DUMMY, DUMMY,
        3.0, 0, 1.0, DUMMY, DUMMY, None, DUMMY, need, None)
    assert "k >= 1" in _error(rc)


def test_gaussian_filter_needs_the_coefficient_variances():
    rc = _lib.lib().bmnd_hip_filter_groups(
        256, DUMMY, DUMMY, None, SHAPE, BLOCK, RADIUS, DUMMY, DUMMY, 10, 16, DUMMY, DUMMY,
        3.0, 0, 1.0, DUMMY, DUMMY, None)
    assert "coefficient variances" in _error(rc)


@pytest.mark.parametrize("gamma,noise,thr,msg", [
    (-1.0, DUMMY, 1.0, "gamma must be finite"), (float("nan"), DUMMY, 1.0, "gamma must be finite"),
    (3.0, None, 1.0, "needs the noise SSD"), (0.0, None, -1.0, "finite and non-negative"),
    (0.0, None, float("inf"), "finite and non-negative")])
def test_ssd_matching_arguments(gamma, noise, thr, msg):
    rc = _lib.lib().bmnd_hip_blockmatch_ssd(DUMMY, SHAPE, BLOCK, RADIUS, DUMMY, 10, thr, 16, 2, 1,
                                            noise, gamma, DUMMY, DUMMY, DUMMY, None)
    assert msg in _error(rc)


@pytest.mark.parametrize("gamma,scale,thr,msg", [
    (-1.0, 1.0, 1.0, "gamma must be finite"), (float("inf"), 1.0, 1.0, "gamma must be finite"),
    (3.0, 0.0, 1.0, "poisson_count_scale must be finite"),
    (3.0, float("nan"), 1.0, "poisson_count_scale must be finite"),
    (3.0, 1.0, -1.0, "finite and non-negative")])
def test_poisson_ssd_matching_arguments(gamma, scale, thr, msg):
    rc = _lib.lib().bmnd_hip_blockmatch_ssd_poisson(DUMMY, SHAPE, BLOCK, RADIUS, DUMMY, 10, thr,
                                                    16, 2, 1, gamma, scale, DUMMY, DUMMY, DUMMY,
                                                    None)
    assert msg in _error(rc)


@pytest.mark.parametrize("domain,msg", [((0, 4, 4), "positive"), ((4096, 4096, 4096), "too large")])
def test_gaussian_variance_arguments(domain, msg):
    rc = _lib.lib().bmnd_hip_gaussian_variance(DUMMY, I32X3(*domain), SHAPE, BLOCK, RADIUS, DUMMY,
                                               DUMMY, 10, 16, DUMMY, 4, DUMMY, None)
    assert msg in _error(rc)


def test_value_range_arguments():
    L = _lib.lib()
    assert "must not be null" in _error(L.bmnd_hip_value_range(DUMMY, 10, None, None))
    assert "empty volume" in _error(L.bmnd_hip_value_range(DUMMY, 0, DUMMY, None))


@pytest.mark.parametrize("table_max,scale", [(101, 1.0), (-1, 1.0), (50, 2.0)])
def test_count_table_mode_refuses_what_its_table_cannot_hold(table_max, scale):
    # mode 2 reads terms from a (table_max+1)^2 table of integer counts at
    # unit count scale; a table past 100 counts or another scale is refused
    L = _lib.lib()
    rc = L.bmnd_hip_blockmatch(DUMMY, SHAPE, BLOCK, RADIUS, DUMMY, 10, 1.0, 16, 2, 1, 1, scale,
                               2, table_max, DUMMY, DUMMY, DUMMY, None)
    assert "table mode needs count_scale == 1 and 0 <= table_max <= 100" in _error(rc)


def test_stage_driver_count_table_mode_needs_unit_scale():
    L = _lib.lib()
    ws = L.bmnd_hip_stage_workspace(SHAPE, BLOCK, 10, 16, 4)
    rc = L.bmnd_hip_stage(0, DUMMY, DUMMY, DUMMY, DUMMY, SHAPE, BLOCK, RADIUS, DUMMY, 10, 1.0,
                          16, 2, 2.0, 2, DUMMY, DUMMY, 3.0, 4, DUMMY, ws, DUMMY, DUMMY, DUMMY,
                          DUMMY, None)
    assert "mode 2 needs count_scale == 1" in _error(rc)
