"""The oracle against the reference's own outputs (golden fixtures written by
tests/golden/make_golden.py from the reference routines) -- bitwise."""
import os

import numpy as np
import pytest

import oracle as O
from conftest import golden_files, load_golden

FILES = golden_files()


def _prob(d):
    k = d["prob"]
    return O.prob2d(int(k[0]), k[1], k[2], k[3], k[4], k[5], k[6], int(k[7]))


def test_fixtures_present():
    assert len(FILES) == 18


@pytest.mark.parametrize("path", FILES, ids=os.path.basename)
def test_structures_and_assembly(orc, path):
    d = load_golden(path)
    n, neq = int(d["n"]), int(d["neq"])
    x, y, t, v = orc.mesh2d(n)
    assert np.array_equal(x, d["x"]) and np.array_equal(y, d["y"])
    assert np.array_equal(t, d["type"]) and np.array_equal(v, d["vertex"])
    neq2, idv = orc.number_nodes(t)
    assert neq2 == neq and np.array_equal(idv, d["ids"])
    nel = len(v) // 3
    lm = orc.fill_lm(neq, nel, v, t, idv)
    assert np.array_equal(lm, d["lm"])
    p = _prob(d)
    for tag in ("u0", "u1"):
        u = np.zeros(neq + 1) if tag == "u0" else d["u1"]
        ke, F = orc.build_supg2d(x, y, t, idv, v, neq, u, p)
        assert np.array_equal(F, d["F_" + tag]), tag
        scheme = str(d["scheme"])
        if scheme == "EBE":
            A = ke
        elif scheme == "CSR":
            nnz, IA, JA, cbe = orc.csr_init(neq, nel, lm)
            assert np.array_equal(IA, d["IA"]) and np.array_equal(JA, d["JA"])
            assert np.array_equal(cbe, d["csr_by_elem"])
            A = orc.csr_assembly(nel, nnz, ke, cbe)
        else:
            nedge, order, ebe, lm2 = orc.ede_init(neq, nel, lm)
            assert np.array_equal(order, d["order"]) and np.array_equal(ebe, d["edge_by_elem"])
            assert np.array_equal(lm2, d["lm2"])
            A = orc.ede_assembly(nel, nedge, ke, order, ebe)
        assert np.array_equal(A, d["A_" + tag]), tag


def _ops(orc, d):
    scheme = str(d["scheme"])
    neq, nel = int(d["neq"]), len(d["vertex"]) // 3
    A = d["A_u0"]
    if scheme == "EBE":
        return (lambda P: orc.ebemv(neq, nel, 3, A, d["lm"], P)), ("EBE", (A, d["lm"], 3, nel))
    if scheme == "CSR":
        return (lambda P: orc.csrmv(neq, A, d["JA"], d["IA"], P)), ("CSR", (A, d["JA"], d["IA"]))
    nedge = len(d["lm2"]) // 2
    return (lambda P: orc.edemv(neq, nedge, A, d["lm2"], P)), ("EDE", (A, d["lm2"], nedge))


@pytest.mark.parametrize("path", FILES, ids=os.path.basename)
def test_matvec(orc, path):
    d = load_golden(path)
    mv, _ = _ops(orc, d)
    assert np.array_equal(mv(d["P"].copy()), d["Q"])


@pytest.mark.parametrize("path", FILES, ids=os.path.basename)
def test_diag_setup(orc, path):
    d = load_golden(path)
    scheme, neq = str(d["scheme"]), int(d["neq"])
    nel = len(d["vertex"]) // 3
    A, F = d["A_u0"], d["F_u0"]
    if scheme == "EBE":
        D, iD, Fs = orc.diag_setup_ebe(neq, nel, 3, A, d["lm"], F)
    elif scheme == "EDE":
        D, iD, Fs = orc.diag_setup_ede(neq, len(d["lm2"]) // 2, A, d["lm2"], F)
    else:
        D, iD, Fs = orc.diag_setup_csr(neq, A, d["JA"], d["IA"], F)
    assert np.array_equal(D[:neq], d["Diag"][:neq])
    assert np.array_equal(iD, d["invDiag"])
    assert np.array_equal(Fs, d["F_diag"])


@pytest.mark.parametrize("path", golden_files("ss2d_*_ebe.npz"), ids=os.path.basename)
def test_jacobi_ebe(orc, path):
    d = load_golden(path)
    neq, nel = int(d["neq"]), len(d["vertex"]) // 3
    invDe, Fj = orc.jacobi_ebe_setup(neq, nel, d["A_u0"], d["lm"], d["F_u0"])
    assert np.array_equal(invDe, d["invDe"]) and np.array_equal(Fj, d["F_jac"])
    assert np.array_equal(orc.jacobi_ebe_apply(neq, nel, invDe, d["lm"], d["P"]), d["zjac"])


@pytest.mark.parametrize("path", FILES, ids=os.path.basename)
def test_pgmres(orc, path):
    d = load_golden(path)
    neq = int(d["neq"])
    _, (scheme, mats) = _ops(orc, d)
    X, steps, _, rho, _ = orc.pgmres(neq, 30, 200, 1e-8, d["F_u0"], scheme, mats)
    assert steps == int(d["it_not"]) and np.array_equal(X, d["X_not"])
    for k in range(1, d["Xk_not"].shape[0] + 1):
        Xk, _, _, _, _ = orc.pgmres(neq, k, 1, 1e-30, d["F_u0"], scheme, mats)
        assert np.array_equal(Xk, d["Xk_not"][k - 1]), k
    X, steps, _, _, _ = orc.pgmres(neq, 30, 200, 1e-8, d["F_diag"], scheme, mats,
                                   invDiag=d["invDiag"])
    assert steps == int(d["it_diag"]) and np.array_equal(X, d["X_diag"])
