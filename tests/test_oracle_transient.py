"""Pins the oracle's transient 2D SUPG element build (Build_M_K_R_SUPG.c of
02_Transport_Equation) -- CPU only.

* against the reference's own routine: tests/golden/tr2d_* were written by
  the reference compiled from its sources (tests/golden/make_golden_tr.py);
  bitwise;
* against the mathematics and the steady build: with u = du = 0 the shock
  term vanishes, so alpha*dt*ke must equal the steady SUPG element matrix
  (Build_K_F_SUPG.c, pinned separately) and R the steady load vector
  exactly; every element mass matrix me (Ae at alpha*dt = 0) sums to the
  element area (the SUPG part tau*C_1i sums to zero over i); R is affine in
  du with slope -M (the shock term does not depend on du when u = 0).
"""
import os

import numpy as np
import pytest

import oracle as O
from conftest import golden_files, load_golden

CASES = golden_files("tr2d_*.npz")


@pytest.fixture(scope="module")
def o():
    return O.Oracle()


def _prob(d):
    p = d["prob"]
    return O.OrcProb2d(int(p[0]), p[1], p[2], p[3], p[4], p[5], p[6], int(p[7]))


@pytest.mark.parametrize("path", CASES, ids=os.path.basename)
def test_transient_golden(o, path):
    d = load_golden(path)
    A, R = o.build_supg2d_transient(d["x"], d["y"], d["vertex"], d["lm"], int(d["neq"]), d["U"],
                                    d["dU"], _prob(d), float(d["alpha"]), float(d["dt"]))
    assert np.array_equal(A, d["A"])
    assert np.array_equal(R, d["R"])


@pytest.mark.parametrize("path", CASES, ids=os.path.basename)
def test_transient_vs_steady(o, path):
    d = load_golden(path)
    neq, nel = int(d["neq"]), len(d["vertex"]) // 3
    Z = np.zeros(len(d["x"]))
    pb = _prob(d)
    A0, R0 = o.build_supg2d_transient(d["x"], d["y"], d["vertex"], d["lm"], neq, Z, Z, pb, 0.0, 1.0)
    A1, R1 = o.build_supg2d_transient(d["x"], d["y"], d["vertex"], d["lm"], neq, Z, Z, pb, 1.0, 1.0)
    ke, F = o.build_supg2d(d["x"], d["y"], d["type"], d["idv"], d["vertex"], neq,
                           np.zeros(neq + 1), pb)
    assert np.array_equal(R0, F) and np.array_equal(R1, F)
    scale = np.abs(ke).max()
    assert np.abs((A1 - A0) - ke).max() <= 1e-14 * scale
    # element mass matrices sum to the element area
    x, y, v = d["x"], d["y"], d["vertex"].reshape(nel, 3)
    area = 0.5 * ((x[v[:, 1]] - x[v[:, 0]]) * (y[v[:, 2]] - y[v[:, 0]]) -
                  (x[v[:, 2]] - x[v[:, 0]]) * (y[v[:, 1]] - y[v[:, 0]]))
    assert np.abs(A0.reshape(nel, 9).sum(axis=1) - area).max() <= 1e-14 * np.abs(area).max()


@pytest.mark.parametrize("path", CASES, ids=os.path.basename)
def test_transient_du_slope(o, path):
    """R(u=0, du) = R(0, 0) - M du, M assembled from the element mass matrices"""
    d = load_golden(path)
    neq, nel = int(d["neq"]), len(d["vertex"]) // 3
    Z = np.zeros(len(d["x"]))
    pb = _prob(d)
    me, R0 = o.build_supg2d_transient(d["x"], d["y"], d["vertex"], d["lm"], neq, Z, Z, pb, 0.0, 1.0)
    _, R1 = o.build_supg2d_transient(d["x"], d["y"], d["vertex"], d["lm"], neq, Z, d["dU"], pb,
                                     0.0, 1.0)
    want = R0.copy()
    me = me.reshape(nel, 3, 3)
    for e in range(nel):
        v = d["vertex"][3 * e:3 * e + 3]
        g = d["lm"][3 * e:3 * e + 3]
        want[g] -= me[e] @ d["dU"][v]
    want[neq] = 0
    assert np.abs(R1[:neq] - want[:neq]).max() <= 1e-13 * np.abs(want).max()
