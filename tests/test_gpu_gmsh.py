"""Assembly on a mesh read from a gmsh file: the device path against the oracle on the same arrays.

MeshCompanionGmsh.msh (applications/integrationtests/mechanics/meshes, binary MSH 2.2: 1637
ten-node tetrahedra of a 24 mm cube with inclusions, 2768 nodes, two physical groups) is read by
nuto_gmsh_read. It runs twice: with its corner nodes as tet4 coordinate elements and the ten-node
dof interpolation (the straight-sided geometry), and isoparametrically, with the ten-node
coordinate elements (its edges are curved by up to 0.41 mm, so J varies over each cell). Bar as in
the other parity tests: pattern, numbering bit-equal, values within 1e-12 * sum|contributions|.
"""
import os

import numpy as np
import pytest

import nuto_amd as N
from pairs import Pair, check_values, to_dev
from test_gpu_parity import ld_params, le_params, run_pair

pytestmark = pytest.mark.gpu
MESH = os.path.join(os.path.dirname(__file__), "golden", "gmsh", "MeshCompanionGmsh.msh")


def companion(group=None):
    m = N.GmshMesh(MESH)
    X, C = m.cells(N.TET10, group)
    return X, np.ascontiguousarray(C[:, :4])


def test_linear_elastic_tet10_dofs_on_gmsh_mesh():
    X, C = companion()
    assert C.shape == (1637, 4)
    p = Pair(N.TET4, N.TET10, X, C, *le_params(3), False)
    u, _, _ = p.inputs(u_range=1e-2)
    run_pair(p, u, None, None)


@pytest.mark.parametrize("law", ["linear_elastic", "local_damage"])
def test_isoparametric_tet10_on_gmsh_mesh(law):
    m = N.GmshMesh(MESH)
    X, C = m.cells(N.TET10)
    assert C.shape == (1637, 10)
    params = le_params(3) if law == "linear_elastic" else ld_params()
    p = Pair(N.TET10, N.TET10, X, C, *params, False)
    u, _, kap = p.inputs(u_range=5e-3, k_range=3e-4)
    # J = sum_k X_k dN_k over ten nodes: terms of ~|X| |dN| (coordinates up to 24 mm) cancel to
    # entries of the element size (~1 mm), and device (FMA) and oracle round that sum differently,
    # by ~1e-16 of the terms. Where a gradient component vanishes in exact arithmetic (faces in
    # coordinate planes, straight edges) that noise reaches K entries that are themselves near zero
    # (up to ~1e-7 max|K|), which the magnitude bound, built from the computed J, cannot see. Their
    # differences stay below 6e-17 max|K| (measured); the absolute floor 1e-15 max|K| is ~20x that.
    run_pair(p, u, None, kap, k_floor=1e-15)


def test_local_damage_on_gmsh_groups():
    m = N.GmshMesh(MESH)
    groups = sorted(set(m.physical.tolist()))
    assert sum(len(m.physical_group(g)) for g in groups) == m.n_elements  # the groups partition the mesh
    for gid in groups:
        X, C = companion(gid)
        assert 0 < C.shape[0] < 1637
        p = Pair(N.TET4, N.TET10, X, C, *ld_params(), False)
        u, _, kap = p.inputs(u_range=5e-3, k_range=3e-4)
        run_pair(p, u, None, kap)


def test_morton_renumbered_cells_same_residual():
    """Reordering the cells along a Morton curve of their centroids changes the summation order
    only: R agrees within the parity bar, K entry by entry after the pattern is mapped."""
    X, C = companion()
    cen = X[C].mean(1)
    order = N.morton_order(cen)
    assert sorted(order) == list(range(len(C)))
    p0 = Pair(N.TET4, N.TET4, X, C, *le_params(3), False)
    p1 = Pair(N.TET4, N.TET4, X, np.ascontiguousarray(C[order]), *le_params(3), False)
    # the same displacement field: value per coordinate node, mapped through each numbering
    field = np.random.default_rng(1).uniform(-1e-2, 1e-2, (X.shape[0], 3))
    def dofs(p):
        u = np.zeros(p.n_u)
        for c in range(p.ncells):
            for a in range(4):
                u[3 * p.ranks[c, a]:3 * p.ranks[c, a] + 3] = field[p.C[c, a]]
        return u
    R0 = p0.context().build_vector(to_dev(dofs(p0)), None, None).cpu().numpy()
    R1 = p1.context().build_vector(to_dev(dofs(p1)), None, None).cpu().numpy()
    node_of = lambda p: {int(p.ranks[c, a]): int(p.C[c, a]) for c in range(p.ncells) for a in range(4)}
    n0, n1 = node_of(p0), node_of(p1)
    F0 = np.zeros_like(field)
    F1 = np.zeros_like(field)
    for r, n in n0.items():
        F0[n] = R0[3 * r:3 * r + 3]
    for r, n in n1.items():
        F1[n] = R1[3 * r:3 * r + 3]
    ref = p0.orc.assemble(dofs(p0), None, None)
    bound = np.zeros_like(field)
    for r, n in n0.items():
        bound[n] = ref["R_bound"][3 * r:3 * r + 3]
    check_values(F1.ravel(), F0.ravel(), 2 * bound.ravel(), "R after Morton renumbering")
