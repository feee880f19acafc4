"""CPU-side checks of the product library: it loads, exports every symbol include/nuto_amd.h
declares, and its host set-up (structured meshes, first-touch dof layers) matches the oracle's
restatement of UnitMeshFem / AddDofInterpolation / DofNumbering bit for bit."""
import os
import re

import numpy as np
import pytest

import nuto_amd as N
import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "nuto_amd.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*|void)\s+(nuto_\w+)\s*\(", txt, re.M)))


def test_library_exports_every_declared_symbol():
    L = N.lib()
    syms = declared_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(L, s), f"missing export {s}"


def test_structured_bricks_match_unit_mesh_transform():
    X, C = N.structured_mesh(N.HEX8, 5, 4, 3, (100.0, 100.0, 100.0))
    m = O.Mesh.bricks(5, 4, 3)
    m.scale(100.0)
    assert np.array_equal(X, m.coords())  # double(i)/n then * 100, bit for bit
    assert np.array_equal(C, m.conn())


def test_structured_quads_match():
    X, C = N.structured_mesh(N.QUAD4, 7, 5)
    m = O.Mesh.quads(7, 5)
    assert np.array_equal(X, m.coords()) and np.array_equal(C, m.conn())


@pytest.mark.parametrize("ct,dt,n", [(N.QUAD4, N.QUAD4, (6, 5, 0)), (N.HEX8, N.HEX8, (4, 3, 5)),
                                     (N.HEX8, N.HEX20, (3, 2, 2)), (N.TET4, N.TET4, (2, 3, 2)),
                                     (N.TET4, N.TET10, (2, 2, 3))])
def test_first_touch_layers_match_oracle(ct, dt, n):
    X, C = N.structured_mesh(ct, *n, (1.0, 2.0, 3.0), 0.08, 99)
    ranks, nn = N.dof_layer(ct, C, dt)
    m = O.Mesh.from_arrays(ct, X, C)
    L = m.add_dof(dt, 1)
    lay = m.dof_layer(L)
    assert nn == lay["n_nodes"]
    assert np.array_equal(lay["numbers"][lay["conn"], 0], ranks)


def test_tet_split_is_conforming_and_positive():
    X, C = N.structured_mesh(N.TET4, 3, 3, 3, (1.0, 1.0, 1.0), 0.1, 5)
    v = X[C[:, 1:]] - X[C[:, :1]]
    det = np.linalg.det(v.transpose(0, 2, 1))
    assert (det > 0).all()
    # each face is shared by exactly two tetrahedra or lies on the boundary
    faces = np.sort(np.concatenate([C[:, [0, 1, 2]], C[:, [0, 1, 3]], C[:, [0, 2, 3]], C[:, [1, 2, 3]]]), axis=1)
    _, cnt = np.unique(faces, axis=0, return_counts=True)
    assert cnt.max() == 2 and (cnt == 1).sum() == 6 * 9 * 2
