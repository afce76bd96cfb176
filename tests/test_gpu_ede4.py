"""The Euler driver's edge-by-edge scheme on the GPU (fg_ede4 + a k = 8
operator) against the reference's fixtures (tests/golden/ede4_*, written by
ede_Initialization.c, ede_assembly.c, EDE/edemvNDOF4.c, EDE/ede2mvNDOF4.c and
BlockDiag_precond_EDE_setup.c themselves) and the oracle.

Tolerances: the edge numbering, vertex order and lm2 are integers: bitwise.
The edge assembly adds an edge's (at most two) element contributions in
element order, as the reference does: bitwise.  The node blocks, inverses
and preconditioned F keep the reference's order: bitwise on the nodes with a
free dof (the others are never read and are left 0).  The edge product adds
a dof's edge contributions in the blocked edge order, not in edge number
order: 1e-14 relative (max norm); ede2 inherits that.  FP32: 2e-6."""
import os

import numpy as np
import pytest

import fem_codes_amd as F
import oracle as O
from conftest import golden_files, load_golden

pytestmark = pytest.mark.gpu
CASES = golden_files("ede4_*.npz")


@pytest.fixture(scope="module")
def ctx():
    c = F.Context(0)
    yield c
    c.close()


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def _edges(ctx, d):
    g = F.Ede4(ctx, int(d["nnodes"]), d["vertex"], d["ids"], int(d["neq"]))
    return g, g.arrays()


@pytest.mark.parametrize("path", CASES, ids=os.path.basename)
def test_ede4_structure(ctx, path):
    d = load_golden(path)
    g, (order, ebe, lm2, en) = _edges(ctx, d)
    assert g.nedge == int(d["nedge"])
    assert np.array_equal(order, d["order"]) and np.array_equal(ebe, d["edge_by_elem"])
    assert np.array_equal(lm2, d["lm2"])
    _, _, _, en_o = O.Oracle().ede4_init(int(d["nnodes"]), d["vertex"], d["ids"], int(d["neq"]))
    assert np.array_equal(en, en_o)


@pytest.mark.parametrize("path", CASES, ids=os.path.basename)
def test_ede4_assembly(ctx, path):
    d = load_golden(path)
    g, _ = _edges(ctx, d)
    Aede = ctx.empty(64 * g.nedge)
    g.assemble(ctx.to_device(d["A12"]), Aede)
    assert np.array_equal(Aede.get(), d["Aede"])


def test_ede4_assembly_of_the_euler_build(ctx):
    """element matrices from the device Euler build, assembled on the device
    into edges: bitwise the oracle's assembly of the same element matrices"""
    d = load_golden(golden_files("eub2d_*.npz")[0])
    neq, nn = int(d["neq"]), len(d["x"])
    m = F.Euler2D(ctx, d["x"], d["y"], d["vertex"], d["lm"], neq)
    A = ctx.empty(144 * m.nel)
    R = ctx.empty(neq + 1)
    p = F.EulerParams(1.4, 3.0, int(d["dimless"]), 1.0, 0.01, 1e-10)
    m.build(p, ctx.to_device(d["U"]), ctx.to_device(d["dU"]), A, R, None, None)
    ids = d["ids"]
    g = F.Ede4(ctx, nn, d["vertex"], ids, neq)
    Aede = ctx.empty(64 * g.nedge)
    g.assemble(A, Aede)
    order, ebe, _, _ = g.arrays()
    ref = O.Oracle().ede4_assembly(g.nedge, A.get(), order, ebe)
    assert np.array_equal(Aede.get(), ref)


def _ops(ctx, d, dev_values=False):
    neq, nn = int(d["neq"]), int(d["nnodes"])
    g, (order, ebe, lm2, en) = _edges(ctx, d)
    op = F.Operator.ebe(ctx, neq, lm2, 8)
    if dev_values:
        Aede = ctx.empty(64 * g.nedge)
        g.assemble(ctx.to_device(d["A12"]), Aede)
        op.set_values(Aede)
    else:
        op.set_values(d["Aede"])
    pc = F.Preconditioner.blockdiag(op, nn, en, d["ids"])
    return g, op, pc


@pytest.mark.parametrize("path", CASES, ids=os.path.basename)
@pytest.mark.parametrize("dev_values", [False, True])
def test_edemv_ndof4(ctx, path, dev_values):
    d = load_golden(path)
    neq = int(d["neq"])
    _, op, _ = _ops(ctx, d, dev_values)
    Q = op.matvec(d["P"])
    assert _rel(Q[:neq], d["Q"][:neq]) <= 1e-14 and Q[neq] == 0.0
    Q32 = op.matvec(d["P"], np.float32)
    assert _rel(Q32[:neq].astype(np.float64), d["Q"][:neq]) <= 2e-6


@pytest.mark.parametrize("path", CASES, ids=os.path.basename)
def test_ede_blockdiag_setup(ctx, path):
    d = load_golden(path)
    _, op, pc = _ops(ctx, d)
    Fd = ctx.to_device(d["F"])
    pc.setup(Fd)
    assert np.array_equal(Fd.get(), d["F_bd"])
    BD, iBD = pc.blocks()
    act = np.repeat((d["ids"].reshape(-1, 4) >= 0).any(axis=1), 16)
    assert np.array_equal(BD[act], d["BD"][act])
    assert np.array_equal(iBD[act], d["invBD"][act])


@pytest.mark.parametrize("path", CASES, ids=os.path.basename)
def test_ede2mv_ndof4(ctx, path):
    d = load_golden(path)
    neq, nn, ne = int(d["neq"]), int(d["nnodes"]), int(d["nedge"])
    _, op, pc = _ops(ctx, d)
    pc.setup(None)
    e2 = F.Operator.ebe2(op, pc, F.EBE2_BLOCKDIAG)
    Q = e2.matvec(d["P"])
    assert _rel(Q[:neq], d["Q2"][:neq]) <= 1e-14 and Q[neq] == 0.0
    Q32 = e2.matvec(d["P"], np.float32)
    assert _rel(Q32[:neq].astype(np.float64), d["Q2"][:neq]) <= 2e-6
    # NOBlockDiag2 form: Q += BD P, against the oracle
    n2 = F.Operator.ebe2(op, pc, F.EBE2_NOBLOCKDIAG)
    Qn = n2.matvec(d["P"])
    Qo = O.Oracle().ebe2mv_items(neq, 8, d["Aede"], d["lm2"], nn, d["Id"], d["P"], BD=d["BD"])
    assert _rel(Qn[:neq], Qo[:neq]) <= 1e-14
    assert e2.algorithmic_bytes() > 0 and ne > 0


@pytest.mark.parametrize("path", CASES, ids=os.path.basename)
def test_gmres_edges_equal_elements(ctx, path):
    """the edge and the element operators are one global matrix: GMRES with
    the node block-diagonal on each gives the same solution (1e-9)"""
    d = load_golden(path)
    neq, nn = int(d["neq"]), int(d["nnodes"])
    _, op8, pc8 = _ops(ctx, d)
    op12 = F.Operator.ebe(ctx, neq, d["lm"], 12)
    op12.set_values(d["A12"])
    pc12 = F.Preconditioner.blockdiag(op12, nn, d["vertex"], d["ids"])
    xs = []
    for op, pc in ((op8, pc8), (op12, pc12)):
        B = ctx.to_device(d["F"])
        pc.setup(B)
        X = ctx.empty(neq + 1)
        info, _, _ = F.gmres(ctx, op, pc, B, X, kmax=30, lmax=300, tol=1e-12)
        assert info.status == F.FG_OK
        xs.append(X.get()[:neq])
    assert _rel(xs[0], xs[1]) <= 1e-9


def test_ede4_rejects_bad_input(ctx):
    d = load_golden(CASES[0])
    nn, neq = int(d["nnodes"]), int(d["neq"])
    bad = d["ids"].copy()
    bad[0] = -2
    with pytest.raises(F.FemGpuError, match="status 2"):
        F.Ede4(ctx, nn, d["vertex"], bad, neq)
    # three triangles on one edge
    v = np.array([0, 1, 2, 0, 1, 3, 0, 1, 4], np.int32)
    with pytest.raises(F.FemGpuError, match="status 2"):
        F.Ede4(ctx, 5, v, np.arange(20, dtype=np.int32), 20)
