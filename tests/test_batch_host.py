"""Host-side checks of the batch records (bmnd_hip_input): the library
rejects records that are not packed in order or whose grids do not match
their shapes before any device work, and the aggregation tile sizes that
tile_off counts are the compiled ones."""

import ctypes

import numpy as np
import pytest

from bmnd_amd import BM3DProfile, NoiseModel, _lib, bmnd_batch
from bmnd_amd.engine import INPUT_DTYPE


def _records(shapes, block):
    recs = np.zeros(len(shapes), dtype=INPUT_DTYPE)
    vox = grid = 0
    for i, s in enumerate(shapes):
        g = tuple(s[d] - block[d] + 1 for d in range(3))
        recs[i] = (vox, grid, 0, s, g)
        vox += int(np.prod(s))
        grid += int(np.prod(g))
    return recs


def _blockmatch(recs, block=(4, 4, 4), radius=(7, 7, 7)):
    L = _lib.lib()
    i32x3 = ctypes.c_int32 * 3
    dummy = ctypes.c_void_p(16)  # never dereferenced: the records are rejected first
    return L.bmnd_hip_blockmatch_batch(dummy, ctypes.c_void_p(recs.ctypes.data), dummy, len(recs),
                                       i32x3(*block), i32x3(*radius), dummy, dummy, None, 10, 1.0, 16,
                                       2, 1, 1, 1.0, 4, 0, None, 0, dummy, dummy, dummy, None)


def test_aggregate_tiles():
    L = _lib.lib()
    t = (ctypes.c_int32 * 3)()
    assert L.bmnd_hip_aggregate_tile((ctypes.c_int32 * 3)(1, 8, 8), t) == 0 and list(t) == [1, 16, 32]
    assert L.bmnd_hip_aggregate_tile((ctypes.c_int32 * 3)(4, 4, 4), t) == 0 and list(t) == [8, 8, 8]
    assert L.bmnd_hip_aggregate_tile((ctypes.c_int32 * 3)(3, 3, 3), t) != 0


@pytest.mark.parametrize("field,value,msg", [
    ("vox_off", 5, b"packed in order"),
    ("grid_off", 7, b"packed in order"),
    ("grid", (3, 3, 3), b"shape - block + 1"),
])
def test_batch_records_are_checked(field, value, msg):
    recs = _records([(10, 12, 9), (8, 8, 8)], (4, 4, 4))
    recs[1][field] = value
    assert _blockmatch(recs) == -1
    assert msg in _lib.lib().bmnd_hip_last_error()


def test_batch_input_smaller_than_block_is_rejected():
    recs = _records([(10, 12, 9), (8, 8, 8)], (4, 4, 4))
    recs[1]["shape"] = (3, 8, 8)
    assert _blockmatch(recs) == -1


def test_batch_rejects_mixed_dimensions():
    with pytest.raises(ValueError, match="same number of dimensions"):
        bmnd_batch([np.ones((16, 16), np.float32), np.ones((8, 8, 8), np.float32)],
                   BM3DProfile(noise_model=NoiseModel.POISSON))


def test_batch_of_all_zero_inputs_needs_no_device():
    z = [np.zeros((9, 12), np.float32), np.zeros((16, 10), np.float32)]
    outs = bmnd_batch(z, BM3DProfile(noise_model=NoiseModel.POISSON))
    assert [o.shape for o in outs] == [(9, 12), (16, 10)] and not any(np.any(o) for o in outs)


def test_stage_workspace_size_and_checks():
    """bmnd_hip_stage_workspace: exact-plane std, filtered patches, weights and
    reference map, each 256-byte aligned; the driver rejects a short
    workspace before any device work."""
    L = _lib.lib()
    i64x3, i32x3 = ctypes.c_int64 * 3, ctypes.c_int32 * 3
    R, M, k = 1000, 32, 4
    got = L.bmnd_hip_stage_workspace(i64x3(30, 40, 50), i32x3(5, 5, 5), R, M, k)
    al = lambda b: (b + 255) // 256 * 256  # noqa: E731
    want = (al(8) + al(R * 4 * 125 * 4) + al(R * M * 125 * 4) + al(R * M * 4)
            + al(26 * 36 * 46 * 4))
    assert got == want
    assert L.bmnd_hip_stage_workspace(i64x3(4, 40, 50), i32x3(5, 5, 5), R, M, k) == -1
    dummy = ctypes.c_void_p(16)
    rc = L.bmnd_hip_stage(0, dummy, dummy, dummy, dummy, i64x3(30, 40, 50), i32x3(5, 5, 5),
                          i32x3(7, 7, 7), dummy, R, 1.0, M, 2, 1.0, -1, dummy, dummy, 3.0, k,
                          dummy, got - 1, dummy, dummy, dummy, dummy, None)
    assert rc == -1 and b"workspace" in L.bmnd_hip_last_error()
