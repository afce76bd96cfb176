"""Autotuner for the element-by-element product: picks the elements per block
(256 x EPT with EPT in {1, 2, 4, 8}) by timing the product on the device.

The block size trades the fraction of shared (halo) dofs, which falls as
blocks grow, against registers per thread (EPT element matrices are held in
registers) and the number of workgroups.  The best size depends on k, the
precision and the mesh, so it is measured: every candidate operator is built,
given the same values, timed with HIP events over `reps` products, and the
fastest is returned.  Candidates whose LDS footprint does not fit are skipped.
"""
from __future__ import annotations

import numpy as np

from . import DeviceArray, FemGpuError, Operator

CANDIDATES = (256, 512, 1024, 2048)


def tune_ebe(ctx, neq: int, lm: np.ndarray, k: int, values, fp32: bool = False,
             reps: int = 10, candidates=CANDIDATES, edges: bool = False):
    """Return (best_elems_per_block, {epb: ms}, best_operator).

    values: AoS element (or edge) matrices, host array or DeviceArray.
    edges: build edge-by-edge operators (k must be 2)."""
    dt = np.float32 if fp32 else np.float64
    P = ctx.to_device(np.random.default_rng(0).standard_normal(neq + 1).astype(dt))
    Q = DeviceArray(ctx, neq + 1, dt)
    times, ops = {}, {}
    for epb in candidates:
        try:
            op = Operator.ede(ctx, neq, lm, epb) if edges else Operator.ebe(ctx, neq, lm, k, epb)
        except FemGpuError:
            continue
        op.set_values(values)
        times[epb] = op.time_apply(P, Q, reps=reps)
        ops[epb] = op
    if not times:
        raise FemGpuError("no block size fits this mesh")
    best = min(times, key=times.get)
    for epb, op in ops.items():
        if epb != best:
            op.close()
    return best, times, ops[best]
