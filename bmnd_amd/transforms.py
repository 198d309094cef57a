"""Transform descriptors and the basis matrices of the default Poisson path.

`Transform`, `TransformType` and `TransformMode` keep the reference's names
(src/bmnd/transforms.py:73-102).  The matrices are built exactly as the
reference builds them (float32 hierarchical Haar, float32 DCT-II from float64
cosines, the official bior1.5 tables with a float64 inverse rounded to
float32), because their float32 bits are inputs of the bit-exact kernels.
"""

from __future__ import annotations

import enum
from dataclasses import dataclass, field
from typing import Any

import numpy as np
from numpy.typing import NDArray


class TransformType(enum.Enum):
    DCT = 0
    FFT = 1
    DST = 2
    HAAR = 3
    HADAMARD = 4
    BIOR15 = 5
    BIOR22 = 6
    BIOR33 = 7
    DB1 = 8
    DB2 = 9
    STARLET = 10
    ATROUS = 11
    B3SPLINE = 12
    CURVELET = 13


class TransformMode(enum.Enum):
    ND = 0
    GROUP = 1
    SPATIAL = 2


@dataclass
class Transform:
    type: TransformType
    mode: TransformMode
    parameters: dict[str, Any] = field(default_factory=dict)


def haar_matrix(size: int) -> NDArray[np.float32]:
    """Hierarchical orthonormal Haar basis, rows [coarse; details]
    (transforms.py:393-408); built by float32 Kronecker products."""
    if size <= 0 or size & (size - 1):
        raise ValueError(f"Haar group size must be a power of two, got {size}.")
    s = np.float32(1.0 / np.sqrt(2.0))
    m = np.ones((1, 1), dtype=np.float32)
    cur = 1
    while cur < size:
        coarse = np.kron(m, np.array([s, s], dtype=np.float32))
        detail = np.kron(np.eye(cur, dtype=np.float32), np.array([s, -s], dtype=np.float32))
        m = np.vstack((coarse, detail)).astype(np.float32, copy=False)
        cur *= 2
    return m


def dct_matrix(size: int) -> NDArray[np.float32]:
    """Orthonormal DCT-II, entries float32(norm * cos(pi*(n+0.5)*k/N))
    (transforms.py:359-386)."""
    m = np.zeros((size, size), dtype=np.float32)
    for k in range(size):
        norm = np.sqrt(1 / size) if k == 0 else np.sqrt(2 / size)
        for n in range(size):
            m[k, n] = norm * np.cos(np.pi * (n + 0.5) * k / size)
    return m


# The official bior1.5 analysis matrices for lengths 4 and 8
# (transforms.py:711-806).  These literal tables define the transform.
_BIOR15_4 = np.array(
    [
        [0.50000000, 0.50000000, 0.50000000, 0.50000000],
        [0.50000000, 0.50000000, -0.50000000, -0.50000000],
        [0.70710677, -0.70710677, 0.0, 0.0],
        [0.0, 0.0, 0.70710677, -0.70710677],
    ],
    dtype=np.float32,
)
_B8 = (0.343550200747110, -0.225454819240296, -0.461645582253923,
       0.569359398342840, 0.402347308162280, -0.083506045090280, 0.707106781186550)
_BIOR15_8 = np.array(
    [
        [_B8[0]] * 8,
        [_B8[1], _B8[2], _B8[2], _B8[1], -_B8[1], -_B8[2], -_B8[2], -_B8[1]],
        [_B8[3], _B8[4], -_B8[4], -_B8[3], _B8[5], -_B8[5], _B8[5], -_B8[5]],
        [_B8[5], -_B8[5], _B8[5], -_B8[5], _B8[3], _B8[4], -_B8[4], -_B8[3]],
        [_B8[6], -_B8[6], 0, 0, 0, 0, 0, 0],
        [0, 0, _B8[6], -_B8[6], 0, 0, 0, 0],
        [0, 0, 0, 0, _B8[6], -_B8[6], 0, 0],
        [0, 0, 0, 0, 0, 0, _B8[6], -_B8[6]],
    ],
    dtype=np.float32,
)


def bior15_matrices(size: int) -> tuple[NDArray[np.float32], NDArray[np.float32]]:
    """Official bior1.5 analysis matrix and its inverse (float64 inverse of
    the float32 table, rounded to float32; transforms.py:810)."""
    if size == 4:
        fwd = _BIOR15_4.copy()
    elif size == 8:
        fwd = _BIOR15_8.copy()
    else:
        raise NotImplementedError(
            f"bior1.5 of length {size} is not one of the official tables (4, 8) "
            "supported by the MI355X path."
        )
    inv = np.asarray(np.linalg.inv(fwd.astype(np.float64)), dtype=np.float32)
    return fwd, inv


def group_matrices(size: int) -> tuple[NDArray[np.float32], NDArray[np.float32]]:
    """Group-axis transform: Haar for powers of two, DCT-II otherwise
    (transforms.py:411-440); the inverse is the transpose."""
    if size > 0 and size & (size - 1) == 0:
        fwd = haar_matrix(size)
    else:
        fwd = dct_matrix(size)
    return fwd, np.ascontiguousarray(fwd.T)


def spatial_matrices(kind: TransformType, size: int) -> tuple[NDArray[np.float32], NDArray[np.float32]]:
    if kind is TransformType.DCT:
        fwd = dct_matrix(size)
        return fwd, np.ascontiguousarray(fwd.T)
    if kind is TransformType.BIOR15:
        return bior15_matrices(size)
    raise NotImplementedError(
        f"spatial transform {kind.name} is not supported by the MI355X path "
        "(supported: BIOR15, DCT)."
    )
