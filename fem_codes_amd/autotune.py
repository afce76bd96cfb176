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

Two more knobs:
- the element ordering of the block plan (breadth-first growth, or runs of
  consecutive elements in the caller's order, which keeps an ordering the
  caller already chose); it is tried at the best block size, since it
  changes the halo fraction and the gather locality but not the register
  budget.  Both orderings give the same products to rounding (1e-14).
- the restart length of the solve (tune_restart): longer cycles take fewer
  steps on stagnating systems but cost more orthogonalisation per step, so
  the fastest converged solve over a candidate list is kept.  The result
  differs from the reference's own restart length only within the solve
  tolerance.
"""
from __future__ import annotations

import numpy as np

from . import (ORDER_BFS, ORDER_MESH, PRODUCT_AUTO, PRODUCT_BLOCKED, PRODUCT_FP32, PRODUCT_FP64,
               DeviceArray, FemGpuError, Operator)

CANDIDATES = (256, 512, 1024, 2048)
PIPELINES = (0, 1, 2, 3, 4, 6)
ORDERINGS = (ORDER_BFS, ORDER_MESH)
ORDER_NAMES = {ORDER_BFS: "bfs", ORDER_MESH: "mesh"}
RESTARTS = (10, 15, 20, 30, 45, 64)
ROWS = "rows"   # the schedule entry of the sliced-row kernel (edges, k = 3 / 4 elements)


def tune_ebe(ctx, neq: int, lm: np.ndarray, k: int, values, fp32: bool = False,
             reps: int = 10, candidates=CANDIDATES, edges: bool = False,
             pipelines=PIPELINES, orderings=(ORDER_BFS,)):
    """Return ((epb, ordering, pipeline), {(epb, ordering, pipe): ms}, best_operator).

    values: AoS element (or edge) matrices, host array or DeviceArray.
    edges: build edge-by-edge operators (k must be 2).  Where the operator
    has sliced rows (edges; elements with k = 3 or 4), that kernel is timed
    too, as the candidate (0, ORDER_BFS, ROWS).
    orderings: the first is tuned over every block size; the others are
    timed at the block size that won."""
    dt = np.float32 if fp32 else np.float64
    P = ctx.to_device(np.random.default_rng(0).standard_normal(neq + 1).astype(dt))
    Q = DeviceArray(ctx, neq + 1, dt)
    times, ops = {}, {}

    def trial(epb, order):
        try:
            if edges:
                op = Operator.ede(ctx, neq, lm, epb, ordering=order)
            else:
                op = Operator.ebe(ctx, neq, lm, k, epb, ordering=order)
        except FemGpuError:
            return
        op.set_values(values)
        if not ops and op.product_sliced(fp32):
            times[(0, ORDER_BFS, ROWS)] = op.time_apply(P, Q, reps=reps)
        if edges or op.product_sliced(fp32):
            op.set_product_kernel(PRODUCT_BLOCKED)
        for pipe in pipelines:
            op.set_pipeline(pipe)
            times[(epb, order, pipe)] = op.time_apply(P, Q, reps=reps)
        ops[(epb, order)] = op

    for epb in candidates:
        trial(epb, orderings[0])
    if not times:
        raise FemGpuError("no block size fits this mesh")
    blocked = {k: v for k, v in times.items() if k[2] != ROWS}
    epb0 = min(blocked, key=blocked.get)[0]
    for order in orderings[1:]:
        trial(epb0, order)
    best = min(times, key=times.get)
    blocked = {k: v for k, v in times.items() if k[2] != ROWS}
    bbest = min(blocked, key=blocked.get)
    for key, op in ops.items():
        if key != bbest[:2]:
            op.close()
    op = ops[bbest[:2]]
    op.set_pipeline(bbest[2])
    # the tuned precision: sliced rows if they won, else the blocked schedule
    # that won; the other precision: whichever of the two is faster on the
    # returned operator (its timings in op.other_precision_ms)
    mine, other = (PRODUCT_FP32, PRODUCT_FP64) if fp32 else (PRODUCT_FP64, PRODUCT_FP32)
    op.set_product_kernel(PRODUCT_AUTO if best[2] == ROWS else PRODUCT_BLOCKED, mine)
    if ROWS in [k[2] for k in times]:
        odt = np.float64 if fp32 else np.float32
        oname = "fp64" if fp32 else "fp32"
        Po = ctx.to_device(np.random.default_rng(1).standard_normal(neq + 1).astype(odt))
        Qo = DeviceArray(ctx, neq + 1, odt)
        op.set_product_kernel(PRODUCT_AUTO, other)
        t_rows = op.time_apply(Po, Qo, reps=reps)
        op.set_product_kernel(PRODUCT_BLOCKED, other)
        t_blk = op.time_apply(Po, Qo, reps=reps)
        op.other_precision_ms = {oname: {"sliced": t_rows, "blocked": t_blk}}
        op.set_product_kernel(PRODUCT_AUTO if t_rows <= t_blk else PRODUCT_BLOCKED, other)
    return best, times, op


def tune_restart(solve, restarts=RESTARTS):
    """solve(kmax) -> (seconds, converged) runs one solve with restart
    length kmax.  Returns (best_kmax, {kmax: seconds or None}); a solve that
    did not converge is recorded as None and never chosen."""
    times = {}
    for km in restarts:
        sec, ok = solve(int(km))
        times[int(km)] = float(sec) if ok else None
    done = {k: v for k, v in times.items() if v is not None}
    if not done:
        raise FemGpuError("no restart length converged")
    return min(done, key=done.get), times


def key_str(key):
    """(epb, ordering, pipe) -> 'epb/ordering/pipe' for JSON records"""
    return f"{key[0]}/{ORDER_NAMES.get(key[1], key[1])}/{key[2]}"
