"""Reference schedule: the reference's known answers (tests/test_shifts_unit.py)
and agreement of the two independent implementations (oracle C / bmnd_amd)."""

import numpy as np
import pytest

import oracle
from bmnd_amd.shifts import generated_schedule_arrays, generated_shifts, get_reference_schedule


def test_generated_shifts_known_answer():
    # tests/test_shifts_unit.py:86-88 (shift_density 1.0)
    assert [s.tolist() for s in generated_shifts((8, 8), (3, 3), 1.0)] == [[0, 4, 7], [0]]


def test_generated_schedule_covers_far_boundary_without_duplicates():
    # tests/test_shifts_unit.py:91-97
    items = get_reference_schedule((9, 9), (8, 8), (3, 3), mode="generated")
    abs_ = [it.reference_abs for it in items]
    assert abs_[0] == (0, 0) and abs_[-1] == (8, 8)
    assert len(abs_) == len(set(abs_))


def test_generated_schedule_extends_to_4d():
    # tests/test_shifts_unit.py:100-112
    items = get_reference_schedule((9, 9, 9, 9), (5, 5, 5, 5), (3, 3, 3, 3), mode="generated")
    assert items[0].reference_abs == (0, 0, 0, 0)
    assert items[-1].reference_abs == (8, 8, 8, 8)


@pytest.mark.parametrize("counts,block", [((57, 57), (8, 8)), ((9, 9), (8, 8)), ((21, 30, 17), (5, 5, 5)),
                                          ((60, 61, 62), (4, 4, 4)), ((1, 40), (8, 8)), ((5, 6, 7, 8), (4, 4, 4, 4))])
@pytest.mark.parametrize("density", [(2.0, 2), (1.0, 1), (3.0, 3)])
def test_two_implementations_agree(counts, block, density):
    step = (3,) * len(block)
    a = generated_schedule_arrays(counts, block, step, *density).reference_abs
    b = oracle.schedule(counts, block, step, *density)
    assert np.array_equal(a, b)


def test_work_sizes_of_the_benchmark_shapes():
    # reference counts quoted for the 64^2, 512^2 and 256^3 inputs
    assert len(generated_schedule_arrays((57, 57), (8, 8), (3, 3), 2.0, 2)) == 728
    assert len(generated_schedule_arrays((505, 505), (8, 8), (3, 3), 2.0, 2)) == 56514
    assert len(generated_schedule_arrays((253,) * 3, (4,) * 3, (3,) * 3, 2.0, 2)) == 610597
    assert len(generated_schedule_arrays((252,) * 3, (5,) * 3, (3,) * 3, 2.0, 2)) == 1220714


def test_shifted_positions_are_slot_bases():
    s = generated_schedule_arrays((20, 20), (8, 8), (3, 3), 2.0, 2)
    assert np.all(s.reference_shifted - s.shift == s.reference_abs)
    assert np.all(s.reference_shifted % 3 == 0) or np.any(s.reference_shifted == 19)
