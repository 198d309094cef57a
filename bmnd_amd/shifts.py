"""Reference-patch schedules -- the work queue of each stage.

Restates the GENERATED mode of the reference (src/bmnd/shifts.py:229-266 and
519-576) with numpy array operations instead of a Python loop over slots:
slot tuples are enumerated in C order with the phase index innermost, the
per-axis shift of axis d is chosen by the slot index of axis d+1 plus the
phase, the last slot of every axis is pinned to the far boundary, and
duplicate positions are dropped keeping their first occurrence.  The OFF
mode (shifts.py:486-516) is included; the table-driven modes are not part of
the MI355X path.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .enums import ReferenceScheduleMode, coerce_enum


@dataclass(frozen=True)
class ReferenceScheduleItem:
    shift: tuple[int, ...]
    reference_abs: tuple[int, ...]
    reference_shifted: tuple[int, ...]


def generated_axis_shifts(block: int, step: int, *, allow_multiple: bool,
                          density: float) -> np.ndarray:
    """Offsets of one axis (shifts.py:229-247): round(linspace(0, block-1, n))
    with n = clip(ceil(density*(block-1)/step), 1, block), half-to-even."""
    if not allow_multiple or block < 3 or step < 1 or step > block:
        return np.array([0], dtype=int)
    count = max(1, math.ceil(max(float(density), 0.0) * (block - 1) / step))
    count = min(count, block)
    if count == 1:
        return np.array([0], dtype=int)
    positions = np.linspace(0.0, float(block - 1), num=count)
    return np.array(sorted({int(round(p)) for p in positions} | {0}), dtype=int)


def generated_shifts(block_size, step_size, shift_density: float = 1.0) -> tuple[np.ndarray, ...]:
    ndim = len(block_size)
    return tuple(
        generated_axis_shifts(int(b), int(s), allow_multiple=axis < max(ndim - 1, 0),
                              density=shift_density)
        for axis, (b, s) in enumerate(zip(block_size, step_size)))


@dataclass
class ScheduleArrays:
    """A schedule as arrays: reference origins, their shifted (slot-base)
    positions and the applied shift, one row per reference in order."""

    reference_abs: np.ndarray  # (R, ndim) int64
    reference_shifted: np.ndarray  # (R, ndim) int64
    shift: np.ndarray  # (R, ndim) int64

    def __len__(self) -> int:
        return int(self.reference_abs.shape[0])

    def items(self) -> list[ReferenceScheduleItem]:
        return [ReferenceScheduleItem(shift=tuple(int(x) for x in s),
                                      reference_abs=tuple(int(x) for x in a),
                                      reference_shifted=tuple(int(x) for x in b))
                for s, a, b in zip(self.shift, self.reference_abs, self.reference_shifted)]


def generated_schedule_arrays(counts, block_size, step_size, shift_density: float = 1.0,
                              schedule_density: int = 1) -> ScheduleArrays:
    counts = tuple(int(c) for c in counts)
    ndim = len(counts)
    empty = np.empty((0, ndim), dtype=np.int64)
    if ndim == 0:
        return ScheduleArrays(empty, empty, empty)
    shifts = generated_shifts(block_size, step_size, shift_density)
    last = np.array([max(c - 1, 0) for c in counts], dtype=np.int64)
    step = np.asarray(step_size, dtype=np.int64)
    slots = (last + step - 1) // step + 1
    phases = max(1, int(schedule_density))
    slot = np.indices(tuple(int(s) for s in slots), dtype=np.int64).reshape(ndim, -1).T
    base = np.minimum(slot * step, last)  # (S, ndim)
    abs_all = np.empty((slot.shape[0], phases, ndim), dtype=np.int64)
    for ph in range(phases):
        for d in range(ndim):
            if d < ndim - 1:
                table = np.asarray(shifts[d], dtype=np.int64)
                offset = table[(slot[:, d + 1] + ph) % table.size]
            else:
                offset = 0
            pos = np.clip(base[:, d] - offset, 0, last[d])
            pos = np.where(slot[:, d] == slots[d] - 1, last[d], pos)
            abs_all[:, ph, d] = pos
    abs_flat = abs_all.reshape(-1, ndim)
    shifted_flat = np.repeat(base, phases, axis=0)
    lin = np.ravel_multi_index(tuple(abs_flat.T), counts)
    _, first = np.unique(lin, return_index=True)
    keep = np.sort(first)
    ref_abs = abs_flat[keep]
    ref_shifted = shifted_flat[keep]
    return ScheduleArrays(ref_abs, ref_shifted, ref_shifted - ref_abs)


def off_schedule_arrays(counts, step_size) -> ScheduleArrays:
    starts = []
    for count, step in zip(counts, step_size):
        last = max(int(count) - 1, 0)
        s = list(range(0, last + 1, int(step)))
        if s[-1] != last:
            s.append(last)
        starts.append(np.asarray(s, dtype=np.int64))
    grids = np.meshgrid(*starts, indexing="ij")
    ref = np.stack([g.reshape(-1) for g in grids], axis=1)
    return ScheduleArrays(ref, ref.copy(), np.zeros_like(ref))


def get_reference_schedule_arrays(counts, block_size, step_size,
                                  mode=ReferenceScheduleMode.GENERATED, shift_density: float = 1.0,
                                  schedule_density: int = 1) -> ScheduleArrays:
    mode = coerce_enum(mode, ReferenceScheduleMode, "reference_schedule_mode")
    if mode is ReferenceScheduleMode.GENERATED:
        return generated_schedule_arrays(counts, block_size, step_size, shift_density,
                                         schedule_density)
    if mode is ReferenceScheduleMode.OFF:
        return off_schedule_arrays(counts, step_size)
    raise NotImplementedError(
        f"reference_schedule_mode {mode.value!r} is not supported by the MI355X path "
        "(supported: 'generated', 'off').")


class LazySchedule:
    """A stage schedule whose host arrays are built on first use only (for
    callbacks, per-reference thresholds and work counts); the device
    pipeline takes its reference list from device_schedule."""

    def __init__(self, counts, block_size, step_size, mode=ReferenceScheduleMode.GENERATED,
                 shift_density: float = 1.0, schedule_density: int = 1):
        self.spec = (tuple(int(c) for c in counts), tuple(int(b) for b in block_size),
                     tuple(int(x) for x in step_size), mode, float(shift_density),
                     int(schedule_density))
        self._arrays: ScheduleArrays | None = None

    @property
    def generated(self) -> bool:
        return coerce_enum(self.spec[3], ReferenceScheduleMode,
                           "reference_schedule_mode") is ReferenceScheduleMode.GENERATED

    @property
    def arrays(self) -> ScheduleArrays:
        if self._arrays is None:
            counts, block, step, mode, sd, dens = self.spec
            self._arrays = get_reference_schedule_arrays(counts, block, step, mode, sd, dens)
        return self._arrays

    def __getattr__(self, name):
        if name in ("reference_abs", "reference_shifted", "shift", "items"):
            return getattr(self.arrays, name)
        raise AttributeError(name)

    def __len__(self) -> int:
        return len(self.arrays)

    def device_lin(self, device):
        """The schedule's reference origins as patch-grid linear indices,
        an int32 device tensor: built on the device in GENERATED mode
        (bmnd_hip_schedule_generated), uploaded from the host arrays
        otherwise."""
        import torch
        counts, block, step, _, sd, dens = self.spec
        if self.generated and torch.device(device).type == "cuda":
            return device_schedule(counts, block, step, sd, dens, device)
        lin = np.ravel_multi_index(tuple(self.arrays.reference_abs.T), counts)
        return torch.from_numpy(lin.astype(np.int32)).to(device)


def device_schedule(counts, block_size, step_size, shift_density: float = 1.0,
                    schedule_density: int = 1, device=None, with_slots: bool = False):
    """The GENERATED schedule built on the device: int32 tensor of patch-grid
    linear indices in the reference's order (and, with_slots, each
    reference's candidate index slot * schedule_density + phase)."""
    import ctypes

    import torch

    from . import _lib
    nd = len(counts)
    c = (ctypes.c_int64 * nd)(*[int(x) for x in counts])
    b = (ctypes.c_int32 * nd)(*[int(x) for x in block_size])
    st = (ctypes.c_int32 * nd)(*[int(x) for x in step_size])
    L = _lib.lib()
    cap = L.bmnd_hip_schedule_capacity(nd, c, b, st, float(shift_density), int(schedule_density))
    if cap < 0:
        _lib.check(-1)
    dev = torch.device(device if device is not None else "cuda")
    lin = torch.empty(cap, dtype=torch.int32, device=dev)
    slot = torch.empty(cap, dtype=torch.int32, device=dev) if with_slots else None
    n = ctypes.c_int64(0)
    stream = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    _lib.check(L.bmnd_hip_schedule_generated(
        nd, c, b, st, float(shift_density), int(schedule_density),
        ctypes.c_void_p(lin.data_ptr()), ctypes.c_void_p(0 if slot is None else slot.data_ptr()),
        cap, ctypes.byref(n), stream))
    lin = lin[:n.value]
    return (lin, slot[:n.value]) if with_slots else lin


def get_reference_schedule(counts, block_size, step_size, mode=ReferenceScheduleMode.OFFICIAL,
                           shift_density: float = 1.0,
                           schedule_density: int = 1) -> list[ReferenceScheduleItem]:
    """List-of-items form with the reference's signature (shifts.py:405-483)."""
    return get_reference_schedule_arrays(counts, block_size, step_size, mode, shift_density,
                                         schedule_density).items()
