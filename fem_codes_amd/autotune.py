"""Autotuner for the element-by-element product: picks the elements per block
(256 x EPT with EPT in {1, 2, 4, 8}) and the launch schedule (a block per
workgroup, or a persistent grid of w workgroups per CU that loads the next
block's element matrices while it reduces the current one) by timing the
product on the device.

The block size trades the fraction of shared (halo) dofs, which falls as
blocks grow, against registers per thread (EPT element matrices are held in
registers) and the number of workgroups; the persistent schedule trades
occupancy (two register sets) for loads in flight.  The best pair depends on
k, the precision and the mesh, so it is measured: every candidate operator is
built, given the same values, timed with HIP events over `reps` products, and
the fastest is returned.  Candidates whose LDS footprint does not fit are
skipped.  Every schedule computes identical results.
"""
from __future__ import annotations

import numpy as np

from . import DeviceArray, FemGpuError, Operator

CANDIDATES = (256, 512, 1024, 2048)
PIPELINES = (0, 1, 2, 3, 4, 6)


def tune_ebe(ctx, neq: int, lm: np.ndarray, k: int, values, fp32: bool = False,
             reps: int = 10, candidates=CANDIDATES, edges: bool = False,
             pipelines=PIPELINES):
    """Return ((best_elems_per_block, best_pipeline), {(epb, pipe): ms}, best_operator).

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
        for pipe in pipelines:
            op.set_pipeline(pipe)
            times[(epb, pipe)] = op.time_apply(P, Q, reps=reps)
        ops[epb] = op
    if not times:
        raise FemGpuError("no block size fits this mesh")
    best = min(times, key=times.get)
    for epb, op in ops.items():
        if epb != best[0]:
            op.close()
    op = ops[best[0]]
    op.set_pipeline(best[1])
    return best, times, op
