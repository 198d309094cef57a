"""The GENERATED schedule built on the device (bmnd_hip_schedule_generated)
against the host restatement (bmnd_amd.shifts, itself pinned against the
oracle's C schedule and the reference's known answers in
tests/test_oracle_schedule.py): the same references in the same order,
element for element, for 1- to 4-D grids, default and non-default steps and
densities, clipped and one-patch grids and both 256^3 BM4D stages (610,597
and 1,220,714 references); and the per-axis shift tables against the
reference's rule (shifts.py:229-247) on the host."""

import ctypes

import numpy as np
import pytest

from bmnd_amd import _lib
from bmnd_amd.shifts import device_schedule, generated_axis_shifts, generated_schedule_arrays


@pytest.mark.parametrize("density", [0.0, 0.5, 1.0, 2.0, 3.0, 7.5])
def test_axis_shift_tables_follow_the_reference_rule(density):
    out = (ctypes.c_int32 * 64)()
    L = _lib.lib()
    for block in range(1, 13):
        for step in range(0, 10):
            for allow in (0, 1):
                n = L.bmnd_hip_schedule_axis_shifts(block, step, allow, density, out)
                want = generated_axis_shifts(block, step, allow_multiple=bool(allow),
                                             density=density)
                assert list(out[:n]) == [int(x) for x in want], (block, step, allow)


def test_capacity_bounds_the_candidates():
    L = _lib.lib()
    c = (ctypes.c_int64 * 3)(252, 252, 252)
    b = (ctypes.c_int32 * 3)(5, 5, 5)
    s = (ctypes.c_int32 * 3)(3, 3, 3)
    assert L.bmnd_hip_schedule_capacity(3, c, b, s, 2.0, 2) == 85 ** 3 * 2
    s0 = (ctypes.c_int32 * 3)(0, 3, 3)
    assert L.bmnd_hip_schedule_capacity(3, c, b, s0, 2.0, 2) == -1


CASES = [
    ((57, 57), (8, 8), (3, 3), 2.0, 2),
    ((505, 505), (8, 8), (3, 3), 2.0, 2),
    ((1, 1), (8, 8), (3, 3), 2.0, 2),
    ((33, 91), (8, 8), (4, 2), 1.0, 3),
    ((21, 24, 19), (4, 4, 4), (3, 3, 3), 2.0, 2),
    ((4, 33, 30), (5, 5, 5), (3, 3, 3), 2.0, 2),
    ((37, 41, 29), (5, 5, 5), (2, 3, 4), 3.0, 1),
    ((40,), (6,), (3,), 2.0, 2),
    ((9, 8, 7, 6), (4, 4, 4, 3), (2, 3, 2, 2), 2.0, 2),
    ((253, 253, 253), (4, 4, 4), (3, 3, 3), 2.0, 2),
    ((252, 252, 252), (5, 5, 5), (3, 3, 3), 2.0, 2),
]


@pytest.mark.gpu
@pytest.mark.parametrize("counts,block,step,sd,dens", CASES,
                         ids=["x".join(map(str, c[0])) for c in CASES])
def test_device_schedule_equals_host(counts, block, step, sd, dens):
    print(f"device schedule {counts} block {block} step {step}", flush=True)
    lin, slot = device_schedule(counts, block, step, sd, dens, "cuda", with_slots=True)
    host = generated_schedule_arrays(counts, block, step, sd, dens)
    want = np.ravel_multi_index(tuple(host.reference_abs.T), counts)
    got = lin.cpu().numpy().astype(np.int64)
    assert np.array_equal(got, want)
    # each reference's candidate slot gives its shifted (slot-base) position
    s = slot.cpu().numpy().astype(np.int64) // dens
    last = np.array(counts) - 1
    slots = (last + np.array(step) - 1) // np.array(step) + 1
    base = np.minimum(np.stack(np.unravel_index(s, tuple(slots)), axis=1) * np.array(step), last)
    assert np.array_equal(base, host.reference_shifted)
    if counts == (253, 253, 253):
        assert got.size == 610597
    if counts == (252, 252, 252):
        assert got.size == 1220714
