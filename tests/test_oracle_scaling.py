"""Pins the oracle's diagonal scalings (scaling.c: Left_scaling_EBE/EDE/CSR,
LeftRight_scaling_*, Left_unscaling) -- CPU only.

* against the reference's own routines: tests/golden/scal_* were written by
  scaling.c and the Diag setups compiled from their sources
  (tests/golden/make_golden_scal.py) on the SS-2D fixtures' matrices;
  bitwise: the scaled matrix, F, Diag, invDiag, the unscaled vector;
* against the mathematics: the assembled Left-scaled matrix is D^{-1} K, the
  LeftRight-scaled one D^{-1/2} K D^{-1/2} (unit diagonal) with F scaled by
  D^{-1/2}, and unscaling multiplies by D^{-1/2}.
"""
import os

import numpy as np
import pytest

import oracle as O
from conftest import golden_files, load_golden

CASES = golden_files("scal_*.npz")


@pytest.fixture(scope="module")
def o():
    return O.Oracle()


def _scale(o, d):
    neq, sch = int(d["neq"]), str(d["scheme"])
    lr = str(d["mode"]) == "LeftRight"
    F = d["F_in"].copy()
    if sch == "ebe":
        D, iD, F = o.diag_setup_ebe(neq, len(d["lm"]) // 3, 3, d["A_in"], d["lm"], F)
        A, F = o.scaling_items(neq, 3, d["A_in"], d["lm"], D, iD, lr, F)
    elif sch == "ede":
        ne = len(d["lm2"]) // 2
        D, iD, F = o.diag_setup_ede(neq, ne, d["A_in"], d["lm2"], F)
        A, F = o.scaling_items(neq, 2, d["A_in"], d["lm2"], D, iD, lr, F)
    else:
        D, iD, F = o.diag_setup_csr(neq, d["A_in"], d["JA"], d["IA"], F)
        A, F = o.scaling_csr(neq, d["A_in"], d["JA"], d["IA"], D, iD, lr, F)
    return A, F, D, iD


def _glob(d, A):
    neq, sch = int(d["neq"]), str(d["scheme"])
    K = np.zeros((neq + 1, neq + 1))
    if sch == "csr":
        IA, JA = d["IA"], d["JA"]
        for i in range(neq):
            K[i, JA[IA[i]:IA[i + 1]]] += A[IA[i]:IA[i + 1]]
    else:
        k = 3 if sch == "ebe" else 2
        L = (d["lm"] if sch == "ebe" else d["lm2"]).reshape(-1, k)
        Ae = A[:len(L) * k * k].reshape(-1, k, k)
        for e in range(len(L)):
            K[np.ix_(L[e], L[e])] += Ae[e]
    return K[:neq, :neq]


@pytest.mark.parametrize("path", CASES, ids=os.path.basename)
def test_scaling_golden(o, path):
    d = load_golden(path)
    neq = int(d["neq"])
    A, F, D, iD = _scale(o, d)
    assert np.array_equal(D[:neq], d["Diag"][:neq]) and np.array_equal(iD[:neq], d["invDiag"][:neq])
    n = len(A) if str(d["scheme"]) == "ebe" else len(A) - (4 if str(d["scheme"]) == "ede" else 1)
    assert np.array_equal(A[:n], d["A"][:n])
    assert np.array_equal(F[:neq], d["F"][:neq])
    x = o.left_unscaling(neq, iD, d["xv"])
    assert np.array_equal(x[:neq], d["x_unscaled"][:neq])


@pytest.mark.parametrize("path", CASES, ids=os.path.basename)
def test_scaling_mathematics(o, path):
    d = load_golden(path)
    neq = int(d["neq"])
    A, F, D, iD = _scale(o, d)
    K0, K1 = _glob(d, d["A_in"]), _glob(d, A)
    Dg = np.diag(K0)
    assert np.allclose(Dg, D[:neq], rtol=1e-14, atol=0)
    s = np.abs(K0).max()
    if str(d["mode"]) == "Left":
        assert np.abs(K1 - K0 / Dg[:, None]).max() <= 1e-14 * s / np.abs(Dg).min()
        assert np.allclose(F[:neq], d["F_in"][:neq] / Dg, rtol=1e-14, atol=0)
    else:
        r = 1.0 / np.sqrt(Dg)
        assert np.abs(K1 - r[:, None] * K0 * r[None, :]).max() <= 1e-13
        assert np.allclose(np.diag(K1), 1.0, rtol=1e-14, atol=0)
        assert np.allclose(F[:neq], d["F_in"][:neq] * r, rtol=1e-14, atol=0)
    x = o.left_unscaling(neq, iD, d["xv"])
    assert np.allclose(x[:neq], d["xv"][:neq] / np.sqrt(Dg), rtol=1e-15, atol=0)
