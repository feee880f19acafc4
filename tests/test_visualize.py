"""Visualize::PostProcess mirror (nuto_amd.visualize, PostProcess.h:13-93 with AverageHandler and
XMLWriter semantics): the written .vtu files are read back with an XML parser (ASCII and the
appended raw binary form must hold the same numbers) and checked against the inputs: one point
per cell corner, dof values at the corners, IP fields averaged per cell in VTK's Voigt order, and
a .pvd that lists every step with its time. Host only; the IP fields come from the oracle's
Cell::Eval on a small gradient-damage mesh."""
import os
import xml.etree.ElementTree as ET

import numpy as np
import pytest

import nuto_amd as N
import oracle as O
from nuto_amd.visualize import PostProcess
from pairs import Pair, make_mesh

GDM = O.Params(kind=1, dim=3, E=30000.0, nu=0.2, k=10.0, kappa0=1e-4, beta=100.0, alpha=0.99, c=4.0)


def read_vtu(path):
    raw = open(path, "rb").read()
    head, _, appended = raw.partition(b'<AppendedData encoding="raw">')
    root = ET.fromstring(head.decode() + "</VTKFile>" if appended else raw.decode())
    data = appended[appended.index(b"_") + 1:] if appended else b""
    types = {"Float32": np.float32, "UInt32": np.uint32, "UInt8": np.uint8}
    out = {}
    for da in root.iter("DataArray"):
        dt = types[da.get("type")]
        if da.get("format") == "appended":
            off = int(da.get("offset"))
            n = int(np.frombuffer(data[off:off + 8], np.uint64)[0])
            a = np.frombuffer(data[off + 8:off + 8 + n], dt)
        else:
            a = np.array(da.text.split(), dtype=np.float64).astype(dt)
        nc = int(da.get("NumberOfComponents", "0"))
        out[da.get("Name")] = a.reshape(-1, nc) if nc > 1 else a
    piece = root.find("UnstructuredGrid/Piece")
    return out, int(piece.get("NumberOfPoints")), int(piece.get("NumberOfCells"))


@pytest.fixture(scope="module")
def problem():
    X, C = make_mesh(N.HEX8, (3, 2, 2), (3.0, 2.0, 2.0), jitter=0.1)
    p = Pair(N.HEX8, N.HEX8, X, C, N.gradient_damage_params(), GDM, True)
    u, e, kap = p.inputs()
    return p, u, e, kap


@pytest.mark.parametrize("binary", [False, True])
def test_vtu_contents(tmp_path, problem, binary):
    p, u, e, kap = problem
    damage = p.orc.eval_ip(3, u, e, kap)      # (cells, ips, 1)
    stress = p.orc.eval_ip(1, u, e, kap)      # (cells, ips, 6)
    pp = PostProcess(str(tmp_path / "out"))
    pp.define_visualizer("GDM", N.HEX8, p.X, p.C)
    pp.add_dof("GDM", "Displacements", u, N.cell_dofs(p.ranks, 3), 3)
    pp.add_dof("GDM", "NonlocalEquivalentStrains", e, N.cell_dofs(p.ranks, 1), 1)
    pp.add_cell_function("GDM", lambda: damage, "Damage")
    pp.add_cell_function("GDM", lambda: stress, "Stress")
    pp.plot(0.0, as_binary=binary)
    pp.plot(0.5, as_binary=binary)
    d, npts, ncells = read_vtu(str(tmp_path / "out" / "GDM1.vtu"))
    nc = p.ncells
    assert ncells == nc and npts == 8 * nc
    np.testing.assert_allclose(d["points"], p.X[p.C].reshape(-1, 3), rtol=1e-6)
    assert np.array_equal(d["connectivity"], np.arange(8 * nc))
    assert np.array_equal(d["offsets"], 8 * np.arange(1, nc + 1))
    assert np.all(d["types"] == 12)
    ud = u[N.cell_dofs(p.ranks, 3)].reshape(-1, 3)
    np.testing.assert_allclose(d["Displacements"], ud, rtol=1e-6, atol=1e-12)
    np.testing.assert_allclose(d["NonlocalEquivalentStrains"], e[p.ranks].ravel(), rtol=1e-6, atol=1e-12)
    np.testing.assert_allclose(d["Damage"], damage.mean(1)[:, 0], rtol=1e-6, atol=1e-9)
    s = stress.mean(1)
    np.testing.assert_allclose(d["Stress"], s[:, [0, 1, 2, 5, 3, 4]], rtol=1e-5, atol=1e-6 * np.abs(s).max())
    pvd = ET.parse(str(tmp_path / "out" / "GDM.pvd")).getroot()
    sets = [(float(x.get("timestep")), x.get("file")) for x in pvd.iter("DataSet")]
    assert sets == [(0.0, "GDM0.vtu"), (0.5, "GDM1.vtu")]


def test_quad_points_are_3d_and_2d_vectors_padded(tmp_path):
    X, C = make_mesh(N.QUAD4, (3, 2))
    ranks, nn = N.dof_layer(N.QUAD4, C, N.QUAD4)
    u = np.arange(2 * nn, dtype=float)
    pp = PostProcess(str(tmp_path))
    pp.define_visualizer("LE", N.QUAD4, X, C)
    pp.add_dof("LE", "Displacements", u, N.cell_dofs(ranks, 2), 2)
    pp.plot(1.0, as_binary=True)
    d, npts, _ = read_vtu(str(tmp_path / "LE0.vtu"))
    assert d["points"].shape == (npts, 3) and np.all(d["points"][:, 2] == 0)
    assert d["Displacements"].shape == (npts, 3) and np.all(d["Displacements"][:, 2] == 0)
    np.testing.assert_allclose(d["Displacements"][:, :2], u[N.cell_dofs(ranks, 2)].reshape(-1, 2))
    assert np.all(d["types"] == 9)


def test_errors(tmp_path):
    pp = PostProcess(str(tmp_path))
    with pytest.raises(KeyError):
        pp.add_cell_function("missing", lambda: None, "x")
    pp.define_visualizer("A", N.TET4, np.zeros((4, 3)), np.array([[0, 1, 2, 3]]))
    with pytest.raises(ValueError):
        pp.define_visualizer("A", N.TET4, np.zeros((4, 3)), np.array([[0, 1, 2, 3]]))
    with pytest.raises(RuntimeError):
        PostProcess("").plot(0.0)
