"""Reverse Cuthill-McKee numbering (Reordering/symrcm.c REORDERING_SYMRCM):
the oracle restatement and the library's host implementation (fg_rcm) against
the reference's own permutation (fixture written by the reference compiled
from its sources, tests/golden/make_golden3d.py) and against what the
algorithm guarantees.  CPU only: fg_rcm is host code."""
import os

import numpy as np
import pytest

import fem_codes_amd as F
import oracle as O
from conftest import GOLDEN, load_golden


@pytest.fixture(scope="module")
def o():
    return O.Oracle()


def _permuted_pattern(neq, IA, JA, P):
    inv = np.empty(neq, np.int64)
    inv[P] = np.arange(neq)
    rows = np.repeat(np.arange(neq), np.diff(IA))
    nr, nc = inv[rows], inv[JA]
    order = np.lexsort((nc, nr))
    IA2 = np.concatenate([[0], np.cumsum(np.bincount(nr, minlength=neq))])
    return IA2.astype(np.int32), nc[order].astype(np.int32)


def test_rcm_golden(o):
    d = load_golden(os.path.join(GOLDEN, "ss3d_rcm_n6_seed3.npz"))
    neq = int(d["neq"])
    P = o.symrcm(neq, d["IA"], d["JA"])
    assert np.array_equal(P, d["P"])
    assert np.array_equal(F.rcm(neq, d["IA"], d["JA"]), d["P"])
    IA2, JA2 = _permuted_pattern(neq, d["IA"], d["JA"], P)
    assert np.array_equal(IA2, d["IA_rcm"]) and np.array_equal(JA2, d["JA_rcm"])
    # the renumbered mesh gives the reordered pattern directly
    lm2 = F.renumber_lm(neq, d["lm"], P)
    nnz, IA3, JA3, _ = F.csr_pattern_k(neq, lm2, 4)
    assert np.array_equal(IA3, d["IA_rcm"]) and np.array_equal(JA3, d["JA_rcm"])


@pytest.mark.parametrize("n,seed", [(5, 1), (9, 2), (14, 4)])
def test_rcm_library_matches_oracle(o, n, seed):
    x, y, z, t, v = F.structured_mesh3d(n)
    nn = len(x)
    perm = np.random.default_rng(seed).permutation(nn)
    inv = np.empty(nn, np.int64)
    inv[perm] = np.arange(nn)
    t, v = t[perm], inv[v].astype(np.int32)
    neq, idv = o.number_nodes(t)
    lm = o.fill_lm(neq, len(v) // 4, v, t, idv, nnoel=4)
    nnz, IA, JA, _ = F.csr_pattern_k(neq, lm, 4)
    P = F.rcm(neq, IA, JA)
    assert np.array_equal(P, o.symrcm(neq, IA, JA))
    assert np.array_equal(np.sort(P), np.arange(neq))
    IA2, JA2 = _permuted_pattern(neq, IA, JA, P)
    assert F.csr_bandwidth(neq, IA2, JA2) < F.csr_bandwidth(neq, IA, JA)
    if O.Reference.available():
        import ctypes as C
        ref = C.CDLL(O.Reference.PATH)
        ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
        ref.ref_symrcm.argtypes = [C.c_int, C.c_int, ip, ip, ip, ip]
        Pr = np.zeros(neq, np.int32)
        ref.ref_symrcm(neq, nnz, IA.copy(), JA.copy(), Pr, np.zeros(nnz, np.int32))
        assert np.array_equal(P, Pr)


def test_rcm_recovers_a_path(o):
    """a path graph numbered at random: RCM finds an end (pseudo-peripheral
    node) and numbers the path consecutively -> bandwidth 1"""
    m = 60
    order = np.random.default_rng(9).permutation(m)
    pos = np.empty(m, np.int64)
    pos[order] = np.arange(m)
    rows = [[] for _ in range(m)]
    for a in range(m - 1):
        i, j = order[a], order[a + 1]
        rows[i].append(j)
        rows[j].append(i)
    IA = [0]
    JA = []
    for i in range(m):
        JA += sorted(rows[i] + [i])
        IA.append(len(JA))
    IA, JA = np.array(IA, np.int32), np.array(JA, np.int32)
    P = F.rcm(m, IA, JA)
    assert np.array_equal(P, o.symrcm(m, IA, JA))
    IA2, JA2 = _permuted_pattern(m, IA, JA, P)
    assert F.csr_bandwidth(m, IA2, JA2) == 1
    assert P[0] in (order[0], order[-1]) and P[-1] in (order[0], order[-1])


def test_rcm_components_and_isolated(o):
    """two components and an isolated equation (a diagonal-only row)"""
    IA = np.array([0, 2, 4, 5, 7, 9], np.int32)
    JA = np.array([0, 1, 0, 1, 2, 3, 4, 3, 4], np.int32)
    P = F.rcm(5, IA, JA)
    assert np.array_equal(P, o.symrcm(5, IA, JA))
    assert sorted(P.tolist()) == [0, 1, 2, 3, 4]
