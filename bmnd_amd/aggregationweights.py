"""Float64 tables of the non-default aggregation weights and of the
group-mass correction, built on the host in the reference's own steps so
that the device kernels and the oracle read the same values.

- `synthesis_metric_diagonals`: the diagonal of the spatial synthesis metric
  S^T diag(w^2) S (S the Kronecker product of the spatial synthesis matrices,
  w the Kaiser window) and of the group metric Ginv^T Ginv
  (aggregationweights.py:79-120), used by the windowed-synthesis domain;
- `constant_synthesis_directions`: per axis the solution of M d = 1 for the
  synthesis matrix M of that axis (poisson.py:120-160), whose outer product
  is the direction of the group-mass correction.
"""

from __future__ import annotations

import numpy as np


def synthesis_metric_diagonals(inverse_spatial, window, group_inverses):
    """(spatial metric diagonal [P], group metric diagonals of every n
    concatenated) in float64."""
    synth = np.array([[1.0]], dtype=np.float64)
    for m in inverse_spatial:
        synth = np.kron(synth, np.asarray(m, dtype=np.float64))
    w = np.asarray(window, dtype=np.float64).reshape(-1)
    if w.size != synth.shape[0]:
        raise ValueError("the window does not match the spatial matrices")
    spatial_metric = synth.T @ ((w ** 2)[:, None] * synth)
    smd = np.diag(spatial_metric).copy()
    gmd = []
    for gi in group_inverses:
        g = np.asarray(gi, dtype=np.float64)
        gmd.append(np.diag(g.T @ g).copy())
    return smd, np.concatenate(gmd) if gmd else np.empty(0)


def constant_synthesis_directions(inverse_mats):
    """Per axis d, the vector with M_d d = 1 (float64 solve)."""
    out = []
    for m in inverse_mats:
        mm = np.asarray(m, dtype=np.float64)
        out.append(np.linalg.solve(mm, np.ones(mm.shape[0], dtype=np.float64)))
    return out
