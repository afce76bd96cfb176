"""Pins the oracle's Euler 2D SUPG element build (Build_M_K_F_SUPG.c with
Delta_CAU, Ax_Ay_calculations, rotation) -- CPU only.

* against the reference's own routine: tests/golden/eub2d_* were written by
  the reference compiled from its sources (tests/golden/make_golden_eub.py);
  bitwise, with both Jacobian forms and slip-wall rotations;
* against the mathematics (no rotation):
  - consistency: the Galerkin, SUPG and CAU parts of Ke each sum to zero
    over the column node blocks (sum_b C_b = 0 because y23+y31+y12 = 0 and
    x32+x13+x21 = 0; the diagonal SUPG/CAU blocks are minus the sum of the
    others), so Ke annihilates a state that is the same at the three nodes
    and R vanishes for a uniform state with dU = 0;
  - mass structure: Me(a,b) = Mg(a,b) I + tau C_a, so Me(a,b) - Me(a,b')
    vanishes for b, b' != a and Me(a,a) - Me(a,b) = (Area/6 - Area/12) I;
  - tau depends on alpha*dt only, so alpha*dt = 0 gives Ae = Me exactly;
* rotation: P^T on a block row and P on its columns keep every 4x4 block's
  Frobenius norm and the load's velocity-pair norm.
"""
import os

import numpy as np
import pytest

import oracle as O
from conftest import golden_files, load_golden

CASES = golden_files("eub2d_*.npz")


@pytest.fixture(scope="module")
def o():
    return O.Oracle()


def _build(o, d, U=None, dU=None, rot=True, alpha=1.0, dt=0.01):
    U = d["U"] if U is None else U
    dU = d["dU"] if dU is None else dU
    r = (d["v1type"] == -1).astype(np.int32) if rot else None
    return o.build_euler2d(d["x"], d["y"], d["vertex"], d["lm"], int(d["neq"]), U, dU, r,
                           d["theta"] if rot else None, 1.4, 3.0, int(d["dimless"]), alpha, dt,
                           1e-10)


@pytest.mark.parametrize("path", CASES, ids=os.path.basename)
def test_euler_build_golden(o, path):
    d = load_golden(path)
    A, R = _build(o, d)
    assert np.array_equal(A, d["A_build"])
    assert np.array_equal(R, d["R_build"])
    assert (d["v1type"] == -1).sum() > 0


@pytest.mark.parametrize("path", CASES, ids=os.path.basename)
def test_euler_build_structure(o, path):
    """alpha*dt = 0 gives Ae = Me exactly (tau depends on alpha*dt only, and
    the Ke term vanishes)"""
    d = load_golden(path)
    nel = int(d["nel"])
    Me = _build(o, d, rot=False, alpha=0.0)[0].reshape(nel, 3, 4, 3, 4)
    # mass structure
    x, y, v = d["x"], d["y"], d["vertex"].reshape(nel, 3)
    area = 0.5 * np.abs((x[v[:, 1]] - x[v[:, 0]]) * (y[v[:, 2]] - y[v[:, 0]]) -
                        (x[v[:, 2]] - x[v[:, 0]]) * (y[v[:, 1]] - y[v[:, 0]]))
    I4 = np.eye(4)
    for a in range(3):
        others = [b for b in range(3) if b != a]
        assert np.abs(Me[:, a, :, others[0], :] - Me[:, a, :, others[1], :]).max() == 0.0
        diff = Me[:, a, :, a, :] - Me[:, a, :, others[0], :]
        want = (area / 6 - area / 12)[:, None, None] * I4
        assert np.abs(diff - want).max() <= 1e-14 * area.max()


@pytest.mark.parametrize("path", CASES, ids=os.path.basename)
def test_euler_uniform_state_residual(o, path):
    """R = 0 for a uniform state at rest in time (no rotation)"""
    d = load_golden(path)
    nn = len(d["x"])
    U = np.tile(np.array([1.0, 1.2, -0.1, 2.9]), nn)
    _, R = _build(o, d, U=U, dU=np.zeros(4 * nn), rot=False)
    Ar, _ = _build(o, d, rot=False)
    assert np.abs(R).max() <= 1e-12 * np.abs(Ar).max()


@pytest.mark.parametrize("path", CASES, ids=os.path.basename)
def test_euler_rotation_norms(o, path):
    d = load_golden(path)
    nel = int(d["nel"])
    A0, R0 = _build(o, d, rot=False)
    A1, R1 = _build(o, d, rot=True)
    A0 = A0.reshape(nel, 3, 4, 3, 4)
    A1 = A1.reshape(nel, 3, 4, 3, 4)
    n0 = np.sqrt((A0 ** 2).sum(axis=(2, 4)))
    n1 = np.sqrt((A1 ** 2).sum(axis=(2, 4)))
    assert np.abs(n1 - n0).max() <= 1e-13 * n0.max()
    rotated = (d["v1type"] == -1)[d["vertex"].reshape(nel, 3)]
    assert rotated.any() and np.abs(A1 - A0)[rotated].max() > 0
