"""Element rows on unstructured meshes: Delaunay triangulations and
tetrahedralisations of random points, random Dirichlet nodes and random
element values, so that the sliced rows have odd widths (the unpaired last
entry of the FP64 triangle layout), widths above 8 (whole chunks and the
clamped rest) and mixed widths within a slice.  The default product must stay
the reference's ebemv / ebemv3Dndof1 bit for bit in FP64 (oracle) and FP32
(a float32 restatement), after new values (AoS scatter into the rows), after
a LeftRight scaling (the rows refreshed from the blocked copy) and for the
Diag setup read from the rows."""
import numpy as np
import pytest
from scipy.spatial import Delaunay

import fem_codes_amd as F
import oracle as O

pytestmark = pytest.mark.gpu


def _mesh(dim, npts, seed):
    rng = np.random.default_rng(seed)
    pts = rng.random((npts, dim))
    tri = Delaunay(pts)
    cells = tri.simplices.astype(np.int64)
    # hull nodes and 10% of the others prescribed (the sink)
    fixed = np.zeros(npts, bool)
    fixed[np.unique(tri.convex_hull)] = True
    fixed[rng.random(npts) < 0.1] = True
    free = np.nonzero(~fixed)[0]
    neq = len(free)
    eq = np.full(npts, neq, np.int64)
    # equations numbered in a random order of the free nodes (within the
    # 15-bit offset fields at these sizes)
    eq[free] = rng.permutation(neq)
    lm = eq[cells].astype(np.int32).ravel()
    return neq, lm


def _synthetic(k, seed):
    """element lists (not a conforming mesh: ebemv takes any lm) whose 64-row
    slices have every width from 1 to 13: slice s's rows appear in 1..w_s
    elements, its first row exactly w_s times; about 5% of the other entries
    are the sink"""
    rng = np.random.default_rng(seed)
    widths = np.tile(np.arange(1, 14), 2)
    neq = 64 * len(widths)
    cnt = np.concatenate([np.append(w, rng.integers(1, w + 1, 63)) for w in widths])
    slots = np.repeat(np.arange(neq), cnt)
    rng.shuffle(slots)
    slots = np.concatenate([slots, np.full((-len(slots)) % k, neq)])
    L = slots.reshape(-1, k)
    # no dof twice in an element: swap each repeat with an entry of another
    # element where neither becomes a repeat
    for e in range(len(L)):
        for j in range(1, k):
            while L[e, j] < neq and L[e, j] in L[e, :j]:
                f, c = rng.integers(len(L)), rng.integers(k)
                a, b = L[e, j], L[f, c]
                if f != e and b not in L[e] and a not in L[f]:
                    L[e, j], L[f, c] = b, a
    sink = (rng.random(L.shape) < 0.05) & (L % 64 != 0)
    L[sink] = neq
    return neq, L.astype(np.int32).ravel()


def _kuhn(n, seed):
    """the tetrahedra of the n^3 cube cut into 6 per cell (at most 14
    neighbours per node) with 30% of them dropped at random (an element list,
    not a conforming mesh), random Dirichlet nodes; the free nodes numbered by
    decreasing element count (random among equals), so that the slices take
    widths from 24 down to a few entries"""
    rng = np.random.default_rng(seed)
    x, y, z, t, v = F.structured_mesh3d(n)
    nn = len(x)
    cells = np.asarray(v).reshape(-1, 4)
    cells = cells[rng.random(len(cells)) >= 0.3]
    fixed = (t <= 0) | (rng.random(nn) < 0.1)
    free = np.nonzero(~fixed)[0]
    neq = len(free)
    cnt = np.bincount(cells.ravel(), minlength=nn)[free]
    order = np.lexsort((rng.random(neq), -cnt))
    eq = np.full(nn, neq, np.int64)
    eq[free[order]] = np.arange(neq)
    return neq, eq[cells].astype(np.int32).ravel()


def _max_neighbours(neq, lm, k):
    L = lm.reshape(-1, k)
    nb = [set() for _ in range(neq)]
    for e in L:
        for a in range(k):
            if e[a] < neq:
                nb[e[a]].update(int(c) for c in e if c < neq and c != e[a])
    return max(len(s) for s in nb)


def _ebemv_f32(neq, lm, A, P, k):
    """ebemv.c:24-40 / ebemv3Dndof1.c:24-42 in float32: each element row
    summed left to right, the sums added to Q in element order"""
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


def _slice_widths(neq, lm):
    cnt = np.bincount(lm, minlength=neq + 1)[:neq]
    ns = (neq + 63) // 64
    c = np.zeros(ns * 64, np.int64)
    c[:neq] = cnt
    return c.reshape(ns, 64).max(1)


@pytest.mark.parametrize("dim,npts,seed", [(2, 6000, 1), (2, 3001, 2), (3, 3000, 3),
                                           (2, 0, 4), (3, 0, 5), (3, -14, 6), (3, -9, 7)])
def test_rows_unstructured_bitwise(dim, npts, seed):
    """(npts > 0: Delaunay of npts points; 0: synthetic element lists;
    -n: the n^3 Kuhn cube, whose tetrahedron rows use neighbour tables)"""
    k = dim + 1
    neq, lm = (_mesh(dim, npts, seed) if npts > 0 else _synthetic(k, seed) if npts == 0
               else _kuhn(-npts, seed))
    nel = lm.size // k
    w = _slice_widths(neq, lm)
    # (every interior vertex of a tetrahedralisation lies in an even number
    # of tetrahedra: its link is a triangulated sphere, F = 2V - 4)
    assert (dim == 3 and npts) or np.any(w % 2 == 1)
    assert np.any(w > 8) and len(np.unique(w)) > 2
    if not npts:
        assert set(range(1, 14)) <= set(w.tolist())
    # k = 4: neighbour tables where every equation has at most 15 distinct
    # other columns, else the offset rows
    form = 2 if k == 4 and _max_neighbours(neq, lm, k) <= 15 else 1
    rng = np.random.default_rng(seed + 10)
    ctx = F.Context(0)
    orc = O.Oracle()
    op = F.Operator.ebe(ctx, neq, lm, k)
    A = rng.standard_normal(k * k * nel)
    op.set_values(ctx.to_device(A))
    assert op.product_form(False) == form and op.product_form(True) == form
    P = rng.standard_normal(neq + 1)
    P[neq] = 0.0

    def check(vals):
        assert np.array_equal(op.matvec(P), orc.ebemv(neq, nel, k, vals, lm, P.copy()))
        P32 = P.astype(np.float32)
        assert np.array_equal(op.matvec(P32, np.float32), _ebemv_f32(neq, lm, vals, P32, k))

    check(A)
    # new values: the AoS scatter writes the blocked copy and both row copies
    A2 = rng.uniform(0.5, 1.5, k * k * nel)
    op.set_values(ctx.to_device(A2))
    check(A2)
    # Diag from the rows: the reference's element-order sums of the diagonal entries
    Fh = rng.standard_normal(neq + 1)
    Fh[neq] = 0.0
    pc = F.Preconditioner(op, F.PREC_DIAG)
    pc.setup(ctx.to_device(Fh))
    D, _ = pc.diag()
    Do, _, _ = orc.diag_setup_ebe(neq, nel, k, A2, lm, Fh.copy())
    assert np.array_equal(D[:neq], Do[:neq])
    # a LeftRight scaling rewrites the blocked values; the rows follow them
    pc.scaling(F.SCALING_LEFTRIGHT, ctx.to_device(np.ones(neq + 1)))
    scaled = op.get_values()
    assert not np.array_equal(scaled, A2)
    check(scaled)
    pc.close()
    op.close()
    ctx.close()
