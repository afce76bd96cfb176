"""Pins the oracle's edge-by-edge scheme of the Euler driver -- CPU only.

* against the reference's own routines: tests/golden/ede4_* were written by
  ede_Initialization.c, ede_assembly.c (03_Euler_Equations), EDE/edemvNDOF4.c,
  EDE/ede2mvNDOF4.c with BlockDiag2_precond.c and
  BlockDiag_precond_EDE_setup.c compiled from their sources
  (tests/golden/make_golden_ede4.py); bitwise: order, EDGE_by_Element, lm2,
  the edge matrices, both products, the node blocks, their inverses and the
  preconditioned F;
* against what the mathematics fixes, independently of the oracle's loops:
  Euler's formula for the edge count of a triangulated square, the edge
  matrices scattered to the global matrix give the element matrices'
  global matrix (each node-diagonal block is shared by the element's two
  edges at that node, hence the halves), the edge node blocks equal the
  element node blocks, the edge product is that global matrix times P, and
  the ede2 product is P + invBD (K - Dblk) P.
"""
import os

import numpy as np
import pytest

import oracle as O
from conftest import golden_files, load_golden

CASES = golden_files("ede4_*.npz")


@pytest.fixture(scope="module")
def o():
    return O.Oracle()


def _glob(A, lm, k, neq):
    K = np.zeros((neq + 1, neq + 1))
    Ae = A.reshape(-1, k, k)
    L = lm.reshape(-1, k)
    for e in range(len(L)):
        K[np.ix_(L[e], L[e])] += Ae[e]
    return K[:neq, :neq]


@pytest.mark.parametrize("path", CASES, ids=os.path.basename)
def test_ede4_structure(o, path):
    d = load_golden(path)
    nn, nel, neq = int(d["nnodes"]), int(d["nel"]), int(d["neq"])
    order, ebe, lm2, en = o.ede4_init(nn, d["vertex"], d["ids"], neq)
    assert np.array_equal(order, d["order"]) and np.array_equal(ebe, d["edge_by_elem"])
    assert np.array_equal(lm2, d["lm2"])
    # a triangulated disk: V - E + F = 1
    assert len(lm2) // 8 == nn + nel - 1 == int(d["nedge"])
    V = d["vertex"].reshape(nel, 3)
    assert set(map(tuple, np.sort(V[np.arange(nel)[:, None], order.reshape(nel, 3)], 1))) == \
        set(map(tuple, np.sort(V, 1)))
    assert np.all(np.diff(V[np.arange(nel)[:, None], order.reshape(nel, 3)], axis=1) > 0)
    assert np.all(en[0::2] < en[1::2])
    ids = np.where(d["ids"] < 0, neq, d["ids"]).reshape(nn, 4)
    assert np.array_equal(lm2.reshape(-1, 2, 4), ids[en.reshape(-1, 2)])
    # every edge is numbered in order of first appearance
    first = np.unique(ebe, return_index=True)[1]
    assert np.all(np.diff(first) > 0)


@pytest.mark.parametrize("path", CASES, ids=os.path.basename)
def test_ede4_assembly(o, path):
    d = load_golden(path)
    neq, ne = int(d["neq"]), int(d["nedge"])
    Aede = o.ede4_assembly(ne, d["A12"], d["order"], d["edge_by_elem"])
    assert np.array_equal(Aede, d["Aede"])
    Ke = _glob(Aede, d["lm2"], 8, neq)
    K = _glob(d["A12"], d["lm"], 12, neq)
    assert np.max(np.abs(Ke - K)) <= 1e-13 * np.max(np.abs(K))


@pytest.mark.parametrize("path", CASES, ids=os.path.basename)
def test_edemv_and_ede2mv(o, path):
    d = load_golden(path)
    neq, nn, ne = int(d["neq"]), int(d["nnodes"]), int(d["nedge"])
    Q = o.ebemv(neq, ne, 8, d["Aede"], d["lm2"], d["P"].copy())
    assert np.array_equal(Q, d["Q"])
    K = _glob(d["Aede"], d["lm2"], 8, neq)
    assert np.max(np.abs(Q[:neq] - K @ d["P"][:neq])) <= 1e-13 * np.max(np.abs(Q))
    # node blocks from the edges
    _, _, _, en = o.ede4_init(nn, d["vertex"], d["ids"], neq)
    BD, iBD, Fb = o.blockdiag_setup_items(neq, 2, nn, en, d["Aede"], d["Id"], d["F"])
    assert np.array_equal(BD, d["BD"]) and np.array_equal(iBD, d["invBD"])
    assert np.array_equal(Fb, d["F_bd"])
    BDe, _, _ = o.blockdiag_setup_ebe(neq, int(d["nel"]), nn, d["vertex"], d["A12"], d["Id"])
    assert np.max(np.abs(BD - BDe)) <= 1e-14 * np.max(np.abs(BDe))
    Q2 = o.ebe2mv_items(neq, 8, d["Aede"], d["lm2"], nn, d["Id"], d["P"], invBD=iBD)
    assert np.array_equal(Q2, d["Q2"])
    # P + invBD (K - Dblk) P on the free dofs
    Id = d["Id"].reshape(nn, 4)
    Kf = np.zeros((neq + 1, neq + 1))
    Kf[:neq, :neq] = K
    Dblk = np.zeros_like(Kf)
    Binv = np.zeros_like(Kf)
    for i in range(nn):
        Dblk[np.ix_(Id[i], Id[i])] += BD[16 * i:16 * i + 16].reshape(4, 4)
        Binv[np.ix_(Id[i], Id[i])] = iBD[16 * i:16 * i + 16].reshape(4, 4)
    P = d["P"].copy()
    t = (Kf - Dblk) @ P
    t[neq] = 0.0   # the product leaves the sink slot 0 before the node pass
    ref = P + Binv @ t
    assert np.max(np.abs(Q2[:neq] - ref[:neq])) <= 1e-12 * np.max(np.abs(ref[:neq]))


@pytest.mark.skipif(not O.ReferenceEuler.available(), reason="reference build absent")
def test_ede4_vs_reference_live(o):
    """a mesh the fixtures do not hold, against the reference run here"""
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
    from make_golden_ede4 import make_case
    c = make_case(8, 21)
    R = O.ReferenceEuler()
    R.setup(c["nnodes"], c["nel"], c["neq"], c["vertex"], c["ids"], c["lm"], c["A12"],
            "BlockDiag")
    order, ebe, lm2 = R.ede_setup()
    oo, oe, ol, en = o.ede4_init(c["nnodes"], c["vertex"], c["ids"], c["neq"])
    assert np.array_equal(order, oo) and np.array_equal(ebe, oe) and np.array_equal(lm2, ol)
    Aede = R.ede_assemble(c["A12"])
    assert np.array_equal(Aede, o.ede4_assembly(len(lm2) // 8, c["A12"], oo, oe))
