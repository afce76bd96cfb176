"""Node rows of the 4-dof-per-node item products (k = 12 triangles,
ebemvNDOF4; k = 8 edges, edemvNDOF4): structured and Delaunay meshes whose
free nodes carry four consecutive equations (numbered in mesh order or at
random), whole nodes prescribed on the inflow side and at random, random item
values.  The node rows must be built and give the reference's product bit for
bit in FP64 (the oracle's row-sum product) and FP32 (a float32 restatement),
after new values (AoS scatter) and after Block_scaling_EBE (the rows
refreshed from the blocked copy); a mesh with a partly prescribed node keeps
the blocked kernel."""
import numpy as np
import pytest
from scipy.spatial import Delaunay

import fem_codes_amd as F
import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = F.Context(0)
    yield c
    c.close()


def _ids(nn, fixed, rng, shuffle):
    """4 consecutive equations per free node (node order, or the free nodes
    in a random order); -1 for prescribed dofs"""
    free = np.nonzero(~fixed)[0]
    order = rng.permutation(len(free)) if shuffle else np.arange(len(free))
    ids = np.full((nn, 4), -1, np.int64)
    ids[free[order]] = 4 * np.arange(len(free))[:, None] + np.arange(4)
    return ids, 4 * len(free)


def _lm(tri, ids, neq):
    lm = ids[tri].reshape(-1)
    return np.where(lm < 0, neq, lm).astype(np.int32)


def _ebemv_f32(neq, lm, A, P, k=12):
    A = np.asarray(A, np.float32).reshape(-1, k, k)
    L = lm.reshape(-1, k)
    p = np.asarray(P, np.float32).copy()
    p[neq] = 0
    pe = p[L]
    s = A[:, :, 0] * pe[:, None, 0]
    for c in range(1, k):
        s = s + A[:, :, c] * pe[:, None, c]
    Q = np.zeros(neq + 1, np.float32)
    np.add.at(Q, L.ravel(), s.ravel())
    Q[neq] = 0
    return Q


def _mesh(kind, seed):
    rng = np.random.default_rng(seed)
    if kind == "structured":
        x, y, t, v = F.structured_mesh2d(40)
        tri = np.asarray(v, np.int64).reshape(-1, 3)
        nn = len(x)
        fixed = (np.abs(x) <= 1e-15) | (rng.random(nn) < 0.1)
        shuffle = False
    else:
        pts = rng.random((2500, 2))
        tri = Delaunay(pts).simplices.astype(np.int64)
        nn = len(pts)
        fixed = rng.random(nn) < 0.1
        shuffle = True
    ids, neq = _ids(nn, fixed, rng, shuffle)
    return tri, nn, ids, neq


def _edges(tri):
    """the unique edges of a triangulation, in order of first appearance"""
    e = np.stack([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]], 1).reshape(-1, 2)
    key = np.minimum(e[:, 0], e[:, 1]) * (tri.max() + 1) + np.maximum(e[:, 0], e[:, 1])
    _, first = np.unique(key, return_index=True)
    return e[np.sort(first)]


@pytest.mark.parametrize("k", [12, 8])
@pytest.mark.parametrize("kind,seed", [("structured", 1), ("delaunay", 2)])
def test_node_rows_bitwise(ctx, kind, seed, k):
    tri, nn, ids, neq = _mesh(kind, seed)
    items = tri if k == 12 else _edges(tri)
    lm = _lm(items, ids, neq)
    nel = len(items)
    rng = np.random.default_rng(seed + 10)
    orc = O.Oracle()
    op = F.Operator.ebe(ctx, neq, lm, k)
    A = rng.standard_normal(k * k * nel)
    op.set_values(ctx.to_device(A))
    assert op.product_sliced(False) and op.product_sliced(True)
    P = rng.standard_normal(neq + 1)
    P[neq] = 0.0

    def check(vals):
        assert np.array_equal(op.matvec(P), orc.ebemv(neq, nel, k, vals, lm, P.copy()))
        P32 = P.astype(np.float32)
        assert np.array_equal(op.matvec(P32, np.float32), _ebemv_f32(neq, lm, vals, P32, k))

    check(A)
    A2 = rng.uniform(0.5, 1.5, k * k * nel)
    op.set_values(ctx.to_device(A2))
    check(A2)
    # a scaling rewrites the blocked values; the node rows follow them
    # (Block_scaling_EBE for the element operator, which the reference scales
    # that way only; the Left/Right Diag scaling for the edge operator)
    if k == 12:
        ids32 = ids.reshape(-1).astype(np.int32)
        pc = F.Preconditioner.blockdiag(op, nn, items.reshape(-1).astype(np.int32), ids32)
        pc.block_scaling(ctx.to_device(rng.standard_normal(neq + 1)))
    else:
        pc = F.Preconditioner(op, F.PREC_DIAG)
        pc.scaling(F.SCALING_LEFTRIGHT, ctx.to_device(np.ones(neq + 1)))
    scaled = op.get_values()
    assert not np.array_equal(scaled, A2)
    check(scaled)
    # the blocked kernel on the same operator: the same product to 1e-14
    op.set_product_kernel(F.PRODUCT_BLOCKED)
    Qb = op.matvec(P)
    Qo = orc.ebemv(neq, nel, k, scaled, lm, P.copy())
    assert np.max(np.abs(Qb - Qo)) <= 1e-14 * np.max(np.abs(Qo))
    pc.close()
    op.close()


def test_node_rows_not_built_for_partial_nodes(ctx):
    """one dof of a node prescribed: no node rows, the blocked product"""
    tri, nn, ids, neq = _mesh("structured", 3)
    free = np.nonzero(ids[:, 0] >= 0)[0]
    ids = ids.copy()
    ids[free[len(free) // 2], 2] = -1           # a hole in one node's group
    flat = ids.reshape(-1)
    keep = flat >= 0
    flat[keep] = np.arange(int(keep.sum()))    # renumber: groups no longer whole
    neq = int(keep.sum())
    ids = flat.reshape(-1, 4)
    lm = _lm(tri, ids, neq)
    nel = len(tri)
    rng = np.random.default_rng(5)
    op = F.Operator.ebe(ctx, neq, lm, 12)
    A = rng.standard_normal(144 * nel)
    op.set_values(ctx.to_device(A))
    assert not op.product_sliced(False) and not op.product_sliced(True)
    P = rng.standard_normal(neq + 1)
    P[neq] = 0.0
    Qo = O.Oracle().ebemv(neq, nel, 12, A, lm, P.copy())
    assert np.max(np.abs(op.matvec(P) - Qo)) <= 1e-14 * np.max(np.abs(Qo))
    op.close()
