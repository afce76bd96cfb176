"""CPU-only checks of the C ABI: the library loads, exports every symbol that
include/femgpu.h declares, and its host-side structure builders (which need
no GPU) reproduce the reference's numbering."""
import os
import re

import numpy as np
import pytest

import fem_codes_amd as F
from conftest import ROOT, golden_files, load_golden


def _declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(fg_[a-z0-9_]+)\s*\(", txt)))


def test_every_declared_symbol_is_exported_and_bound():
    names = _declared("femgpu.h")
    assert len(names) > 30
    for n in names:
        assert hasattr(F.lib, n), n
        assert n in F.SIGNATURES, f"{n} has no ctypes signature"


def test_version():
    assert F.lib.fg_version().decode().startswith("femgpu")


@pytest.mark.parametrize("path", golden_files("ss2d_n*_const_*.npz"), ids=os.path.basename)
def test_structure_builders_match_reference(path):
    d = load_golden(path)
    x, y, t, v = F.structured_mesh2d(int(d["n"]))
    assert np.array_equal(x, d["x"]) and np.array_equal(y, d["y"])
    assert np.array_equal(t, d["type"]) and np.array_equal(v, d["vertex"])
    neq, idv = F.number_nodes(t)
    assert neq == int(d["neq"]) and np.array_equal(idv, d["ids"])
    lm = F.fill_lm(neq, v, t, idv)
    assert np.array_equal(lm, d["lm"])
    scheme = str(d["scheme"])
    if scheme == "CSR":
        nnz, IA, JA, cbe = F.csr_pattern(neq, lm)
        assert np.array_equal(IA, d["IA"]) and np.array_equal(JA, d["JA"])
        assert np.array_equal(cbe, d["csr_by_elem"])
    if scheme == "EDE":
        nedge, order, ebe, lm2 = F.ede_pattern(neq, lm)
        assert np.array_equal(order, d["order"]) and np.array_equal(ebe, d["edge_by_elem"])
        assert np.array_equal(lm2, d["lm2"])


@pytest.mark.parametrize("n", [2, 3, 64, 100])
def test_structure_sizes_closed_form(orc, n):
    """nnz and nedge of the structured mesh against the closed forms
    neq=(n-1)^2, nnz=neq+2[2(n-1)(n-2)+(n-2)^2], nedge=2(n-1)(n-2)+(n-2)^2+4(n-2)
    and against the oracle's restatement of csr_/ede_Initialization."""
    x, y, t, v = F.structured_mesh2d(n)
    neq, idv = F.number_nodes(t)
    lm = F.fill_lm(neq, v, t, idv)
    assert neq == (n - 1) ** 2
    nnz = F.csr_pattern(neq, lm)[0]
    assert nnz == neq + 2 * (2 * (n - 1) * (n - 2) + (n - 2) ** 2)
    nedge, order, ebe, lm2 = F.ede_pattern(neq, lm)
    if n > 2:
        assert nedge == 2 * (n - 1) * (n - 2) + (n - 2) ** 2 + 4 * (n - 2)
    on, oo, oe, ol = orc.ede_init(neq, len(v) // 3, lm)
    assert on == nedge and np.array_equal(oo, order) and np.array_equal(oe, ebe)
    assert np.array_equal(ol, lm2)
    onz, oIA, oJA, ocbe = orc.csr_init(neq, len(v) // 3, lm)
    _, IA, JA, cbe = F.csr_pattern(neq, lm)
    assert onz == nnz and np.array_equal(oIA, IA) and np.array_equal(oJA, JA)
    assert np.array_equal(ocbe, cbe)


def test_invalid_arguments_fail_loudly():
    z = np.zeros(3, np.int32)
    rc = F.lib.fg_fill_lm(3, 1, 0, z, z, z, z)
    assert rc == F.FG_ERR_ARG
    assert "invalid argument" in F.lib.fg_last_error().decode()


def test_operator_sizes_beyond_32bit_offsets_are_rejected():
    """the blocked products address P and Q through 32-bit byte offsets:
    fg_op_ebe_create / fg_op_ede_create refuse neq >= 2^28 (and nitems * k
    beyond an int) before any device work, with or without a context"""
    lm = np.zeros(3, np.int32)
    out = F._vp()
    for neq in (1 << 28, (1 << 29) + 5):
        rc = F.lib.fg_op_ebe_create(None, neq, 1, 3, lm, 0, F.C.byref(out))
        assert rc == F.FG_ERR_ARG
        assert "2^28" in F.lib.fg_last_error().decode()
        rc = F.lib.fg_op_ede_create(None, neq, 1, lm[:2].copy(), 0, F.C.byref(out))
        assert rc == F.FG_ERR_ARG
        assert "2^28" in F.lib.fg_last_error().decode()
    rc = F.lib.fg_op_ebe_create(None, 100, 1 << 28, 12, lm, 0, F.C.byref(out))
    assert rc == F.FG_ERR_ARG
    assert "fit in an int" in F.lib.fg_last_error().decode()
    # the largest admissible size passes the size checks and stops at the
    # missing context
    rc = F.lib.fg_op_ebe_create(None, (1 << 28) - 1, 1, 3, lm, 0, F.C.byref(out))
    assert rc == F.FG_ERR_ARG
    assert "ctx" in F.lib.fg_last_error().decode()


def test_ordering_knob_arguments():
    """fg_op_ebe_create_ex: an unknown ordering, or an edge operator with
    k != 2, is refused before any device work"""
    lm = np.zeros(3, np.int32)
    out = F._vp()
    rc = F.lib.fg_op_ebe_create_ex(None, 10, 1, 3, lm, 0, 7, 0, F.C.byref(out))
    assert rc == F.FG_ERR_ARG and "ordering" in F.lib.fg_last_error().decode()
    rc = F.lib.fg_op_ebe_create_ex(None, 10, 1, 3, lm, 0, F.ORDER_MESH, 1, F.C.byref(out))
    assert rc == F.FG_ERR_ARG and "k = 2" in F.lib.fg_last_error().decode()
    rc = F.lib.fg_op_ebe_create_ex(None, 10, 1, 3, lm, 0, F.ORDER_MESH, 0, F.C.byref(out))
    assert rc == F.FG_ERR_ARG and "ctx" in F.lib.fg_last_error().decode()
