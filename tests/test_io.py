"""The drivers' on-disk formats (fg_param_file_read, fg_mesh_file_read) --
CPU only: the library's readers against the oracle, which reads the same
files with the fscanf formats of each driver's Preprocess.c, and against the
arrays the files were written from.

Files: tests/golden/files/ss3d_params_GM1_DD_DIFREA1.dat is one of the
reference's own parameter files (01_SS_Transport_Equation/3D/01_CommonFiles/
0_Parameters/DIFREA); the 2D scalar and Euler parameter files and the mesh
files are written here in the documented formats, from the golden meshes
(coordinates with %.17g and with the 6 decimals a %lf writer gives)."""
import os

import numpy as np
import pytest

import fem_codes_amd as F
import oracle as O
from conftest import load_golden

HERE = os.path.dirname(os.path.abspath(__file__))
FILES = os.path.join(HERE, "golden", "files")


@pytest.fixture(scope="module")
def o():
    return O.Oracle()


def _cmp_params(o, path, driver):
    fields = F.PARAM_FIELDS[driver]
    kinds = "".join(k for _, k in fields)
    got = F.read_param_file(path, driver)
    ref = o.read_params(path, kinds)
    assert len(ref) == len(fields)
    for (name, kind), rv in zip(fields, ref):
        v = got[name]
        if kind == "d":
            assert float(rv) == v, name
        elif kind == "i":
            assert int(rv) == v, name
        else:
            assert rv == v, name
    return got


def test_reference_parameter_file(o):
    p = _cmp_params(o, os.path.join(FILES, "ss3d_params_GM1_DD_DIFREA1.dat"), "ss3d")
    assert p["ProblemTitle"] == "DIFREA" and p["Preconditioner"] == "ILU1"
    assert p["KrylovBasisVectorsQuantity"] == 50 and p["SolverTolerance"] == 1e-5


def test_written_parameter_files(o, tmp_path):
    rng = np.random.default_rng(3)
    for driver in ("ss2d", "euler2d"):
        lines = []
        for name, kind in F.PARAM_FIELDS[driver]:
            v = {"d": f"{rng.uniform(-5, 5):.10g}", "i": str(int(rng.integers(0, 10 ** 6))),
                 "s": "VAL_" + name[:6], "p": "BlockDiag"}[kind]
            lines.append(f"{v}\t\t:{name} label, with spaces: and colons\n")
        path = tmp_path / f"{driver}.dat"
        path.write_text("".join(lines))
        _cmp_params(o, str(path), driver)
    short = tmp_path / "short.dat"
    short.write_text("EXP\t:Experiments\n")
    with pytest.raises(F.FemGpuError, match="status 2"):
        F.read_param_file(str(short), "ss2d")


def _write_mesh(path, xyz, ni, v, ei, fmt):
    with open(path, "w") as f:
        f.write(f"{len(xyz)}\n")
        for c, t in zip(xyz, ni):
            f.write("\t".join(fmt % a for a in c) + "\t" + "\t".join(str(a) for a in t) + "\n")
        nv = len(v) // len(ei) if len(ei) else 3
        f.write(f"{len(ei)}\n")
        V = v.reshape(-1, nv)
        for a, b in zip(V, ei):
            f.write("\t".join(str(x) for x in list(a) + list(b)) + "\n")


@pytest.mark.parametrize("fmt", ["%.17g", "%lf"])
def test_mesh_files(o, tmp_path, fmt):
    d = load_golden(os.path.join(HERE, "golden", "ss2d_n8_rot_ebe.npz"))
    e = load_golden(os.path.join(HERE, "golden", "eub2d_n6_dimless.npz"))
    nel2, nele = len(d["vertex"]) // 3, len(e["vertex"]) // 3
    cases = {
        "ss2d": (np.stack([d["x"], d["y"]], 1), d["type"].reshape(-1, 1), d["vertex"],
                 np.ones((nel2, 1), np.int32)),
        "euler2d": (np.stack([e["x"], e["y"]], 1),
                    np.stack([e["v1type"] * 0 + 1, e["v1type"], e["v1type"], e["v1type"] * 0],
                             1).astype(np.int32), e["vertex"], np.zeros((nele, 1), np.int32)),
    }
    for driver, (xyz, ni, v, ei) in cases.items():
        path = str(tmp_path / f"{driver}_mesh.dat")
        _write_mesh(path, xyz, ni, v, ei, fmt)
        got = F.read_mesh_file(path, driver)
        ref = o.read_mesh(path, *F.MESH_FORMATS[driver])
        for g, r in zip(got, ref):
            assert np.array_equal(g, r), driver
        assert np.array_equal(got[1], ni) and np.array_equal(got[2], v)
        assert np.array_equal(got[3], ei)
        if fmt == "%.17g":
            assert np.array_equal(got[0], xyz)
        else:
            assert np.max(np.abs(got[0] - xyz)) <= 5e-7
    bad = tmp_path / "bad.dat"
    bad.write_text("3\n0 0 1\n1 0 1\n0 1 1\n1\n0 1 7 1\n")
    with pytest.raises(ValueError):
        F.read_mesh_file(str(bad), "ss2d")
    trunc = tmp_path / "trunc.dat"
    trunc.write_text("3\n0 0 1\n1 0 1\n")
    with pytest.raises(F.FemGpuError, match="status 2"):
        F.read_mesh_file(str(trunc), "ss2d")


def test_mesh_file_3d_gmsh_numbering(o, tmp_path):
    """x y z Type nodes, v0..v3 Type elements, 1-based vertices (GMSH origin)"""
    d = load_golden(os.path.join(HERE, "golden", "ss3d_kuhn4.npz"))
    xyz = np.stack([d["x"], d["y"], d["z"]], 1)
    nel = len(d["vertex"]) // 4
    ni = np.ones((len(xyz), 1), np.int32)
    ei = np.zeros((nel, 1), np.int32)
    path = str(tmp_path / "m3d.dat")
    _write_mesh(path, xyz, ni, d["vertex"] + 1, ei, "%.17g")
    got = F.read_mesh_file(path, "ss3d", one_based=True)
    assert np.array_equal(got[0], xyz) and np.array_equal(got[2], d["vertex"])
    ref = o.read_mesh(path, *F.MESH_FORMATS["ss3d"])
    assert np.array_equal(ref[2] - 1, got[2]) and np.array_equal(ref[0], got[0])
