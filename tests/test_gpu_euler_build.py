"""Euler 2D SUPG element build on the GPU (fg_euler2d_build) against the
reference's fixtures (tests/golden/eub2d_*, written by Build_M_K_F_SUPG.c
itself) and the oracle.  Tolerance 1e-13 relative (max norm): the device
evaluates pow(v, 2) as v*v, the correctly rounded square, where the host's
pow differs in the last bit for about 0.1% of arguments, and its cos/sin
(slip-wall rotation) may differ from the host's by an ulp; every other
operation keeps the reference's order without FMA contraction."""
import os

import numpy as np
import pytest

import fem_codes_amd as F
import oracle as O
from conftest import golden_files, load_golden

pytestmark = pytest.mark.gpu
CASES = golden_files("eub2d_*.npz")
TOL = 1e-13


@pytest.fixture(scope="module")
def ctx():
    c = F.Context(0)
    yield c
    c.close()


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def _gpu(ctx, d, dimless, rot=True):
    m = F.Euler2D(ctx, d["x"], d["y"], d["vertex"], d["lm"], int(d["neq"]))
    A = ctx.empty(144 * m.nel)
    R = ctx.empty(m.neq + 1)
    p = F.EulerParams(1.4, 3.0, int(dimless), 1.0, 0.01, 1e-10)
    r = ctx.to_device((d["v1type"] == -1).astype(np.int32), np.int32) if rot else None
    th = ctx.to_device(d["theta"]) if rot else None
    m.build(p, ctx.to_device(d["U"]), ctx.to_device(d["dU"]), A, R, r, th)
    return A.get(), R.get()


@pytest.mark.parametrize("path", CASES, ids=os.path.basename)
def test_euler_build_golden(ctx, path):
    d = load_golden(path)
    A, R = _gpu(ctx, d, int(d["dimless"]))
    assert _rel(A, d["A_build"]) <= TOL
    assert _rel(R, d["R_build"]) <= TOL
    assert R[int(d["neq"])] == 0.0
    # most entries are bit-identical (only pow/cos/sin can differ)
    assert np.mean(A == d["A_build"]) > 0.5


@pytest.mark.parametrize("dimless", [0, 1])
def test_euler_build_vs_oracle(ctx, dimless):
    """a 40x40 mesh (3200 elements) with a seeded state, no slip walls"""
    import importlib.util
    spec = importlib.util.spec_from_file_location(
        "mk_eub", os.path.join(os.path.dirname(__file__), "golden", "make_golden_eub.py"))
    mk = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mk)
    d = mk.make_case(40, 33 + dimless)
    A, R = _gpu(ctx, d, dimless, rot=False)
    Ao, Ro = O.Oracle().build_euler2d(d["x"], d["y"], d["vertex"], d["lm"], int(d["neq"]),
                                      d["U"], d["dU"], None, None, 1.4, 3.0, dimless, 1.0, 0.01,
                                      1e-10)
    assert _rel(A, Ao) <= TOL
    assert _rel(R, Ro) <= TOL
