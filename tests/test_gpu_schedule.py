"""Launch schedules of the blocked EBE/EDE product: the persistent pipelined
grid (fg_op_set_pipeline) must give bit-identical results to one block per
workgroup, on a mesh large enough that every workgroup walks several blocks
(512^2 cells: 524,288 elements, 2,048 blocks of 256 elements), and both must
match the oracle's ebemv/edemv (ebemv.c:4-43, edemv.c:4-39)."""
import numpy as np
import pytest

import fem_codes_amd as F
import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def system():
    ctx = F.Context(0)
    n = 512
    x, y, t, v = F.structured_mesh2d(n)
    mesh = F.Mesh2D(ctx, x, y, t, v)
    neq, nel = mesh.neq, mesh.nel
    ke = ctx.empty(9 * nel)
    Fv = ctx.empty(neq + 1)
    Fv.zero()
    pb = F.Supg2dProblem(F.P_ROTATING, 1e-5, 0.0, 1.0, 0.0, 0.0, 1.3, F.H_2SQRTAREA)
    mesh.build_supg(pb, ke, Fv)
    P = np.random.default_rng(11).standard_normal(neq + 1)
    P[neq] = 0.0
    yield ctx, neq, nel, mesh.lm(), ke, P
    ctx.close()


@pytest.mark.parametrize("epb", [256, 512, 1024])
def test_ebe_pipeline_identical(system, epb):
    ctx, neq, nel, lm, ke, P = system
    op = F.Operator.ebe(ctx, neq, lm, 3, epb)
    op.set_values(ke)
    ref64 = op.matvec(P)
    ref32 = op.matvec(P.astype(np.float32), np.float32)
    for wpc in (1, 2, 4):
        op.set_pipeline(wpc)
        assert np.array_equal(op.matvec(P), ref64)
        assert np.array_equal(op.matvec(P.astype(np.float32), np.float32), ref32)
    Qo = O.Oracle().ebemv(neq, nel, 3, ke.get(), lm, P.copy())
    assert np.max(np.abs(ref64[:neq] - Qo[:neq])) <= 1e-14 * np.max(np.abs(Qo[:neq]))
    op.close()


def test_ede_pipeline_identical(system):
    ctx, neq, nel, lm, ke, P = system
    nedge, order, ebe, lm2 = F.ede_pattern(neq, lm)
    op = F.Operator.ede(ctx, neq, lm2)
    F.assemble_ede(op, order, ebe, ke)
    ref = op.matvec(P)
    op.set_pipeline(3)
    assert np.array_equal(op.matvec(P), ref)
    op.close()
