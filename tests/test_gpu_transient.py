"""Transient 2D SUPG element build on the GPU (fg_supg2d_transient_build)
against the reference's fixtures (tests/golden/tr2d_*, written by
Build_M_K_R_SUPG.c itself) and the oracle: bitwise for the constant and
rotating problems (the same operation order, no FMA contraction, IEEE
division and square root); 1e-14 relative for VALIDACAO, whose source term
goes through exp and pow."""
import os

import numpy as np
import pytest

import fem_codes_amd as F
import oracle as O
from conftest import golden_files, load_golden

pytestmark = pytest.mark.gpu
CASES = golden_files("tr2d_*.npz")


@pytest.fixture(scope="module")
def ctx():
    c = F.Context(0)
    yield c
    c.close()


def _build(ctx, x, y, t, v, prob, alpha, dt, U, dU):
    mesh = F.Mesh2D(ctx, x, y, t, v)
    A = ctx.empty(9 * mesh.nel)
    R = ctx.empty(mesh.neq + 1)
    mesh.build_transient(prob, alpha, dt, ctx.to_device(U), ctx.to_device(dU), A, R)
    return mesh, A.get(), R.get()


@pytest.mark.parametrize("path", CASES, ids=os.path.basename)
def test_transient_golden(ctx, path):
    d = load_golden(path)
    p = d["prob"]
    prob = F.Supg2dProblem(int(p[0]), p[1], p[2], p[3], p[4], p[5], p[6], int(p[7]))
    mesh, A, R = _build(ctx, d["x"], d["y"], d["type"], d["vertex"], prob, float(d["alpha"]),
                        float(d["dt"]), d["U"], d["dU"])
    assert mesh.neq == int(d["neq"]) and np.array_equal(mesh.lm(), d["lm"])
    assert np.array_equal(A, d["A"])
    assert np.array_equal(R, d["R"])


@pytest.mark.parametrize("kind", [F.P_CONST, F.P_ROTATING, F.P_ROTATE_ABOUT, F.P_VALIDACAO])
def test_transient_vs_oracle(ctx, kind):
    o = O.Oracle()
    n = 40
    x, y, t, v = o.mesh2d(n)
    neq, idv = o.number_nodes(t)
    lm = o.fill_lm(neq, len(v) // 3, v, t, idv)
    rng = np.random.default_rng(kind + 10)
    U = np.where(t > 0, rng.standard_normal(len(x)), 0.3)
    dU = np.where(t > 0, rng.standard_normal(len(x)), 0.0)
    args = (kind, 1e-3, 0.2, 0.7, 0.5, 0.5, 1.5, 0)
    prob = F.Supg2dProblem(*args)
    _, A, R = _build(ctx, x, y, t, v, prob, 0.5, 0.02, U, dU)
    Ao, Ro = o.build_supg2d_transient(x, y, v, lm, neq, U, dU, O.OrcProb2d(*args), 0.5, 0.02)
    if kind == F.P_VALIDACAO:
        assert np.abs(A - Ao).max() <= 1e-14 * np.abs(Ao).max()
        assert np.abs(R - Ro).max() <= 1e-14 * np.abs(Ro).max()
    else:
        assert np.array_equal(A, Ao)
        assert np.array_equal(R, Ro)
    assert R[neq] == 0.0
