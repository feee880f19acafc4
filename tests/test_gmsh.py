"""Mesh input (nuto_gmsh_*, MeshGmsh) and Morton codes (nuto_morton_*, MortonOrder.h) on the host.

The cases of test/mechanics/mesh/MeshGmsh.cpp:7-125 on the same mesh files (copied to
tests/golden/gmsh/, data fixtures of the reference's own tests): physical groups by id and name,
non-contiguous node ids, the binary import, node counts of every element type, and the rejected
inverted triangle. CheckMesh's sign test of det J at every node's natural coordinates
(MeshGmsh.cpp:23-45) is restated here with shape functions built independently of the reader,
as the unique interpolants of each element's polynomial space through its natural node
coordinates (a Vandermonde solve), so a wrong node reordering fails it. Further cases cover what
the reader shares with the format rather than with the test files (binary blocks of several
elements, unknown sections and element types, old versions), and the Morton codes against
MortonOrder's own example (EncodeMorton3D(2, 1, 1) == 14) and bitwise restatements.
"""
import itertools
import os
import struct

import numpy as np
import pytest

import nuto_amd as N

MESHES = os.path.join(os.path.dirname(__file__), "golden", "gmsh")


def path(name):
    return os.path.join(MESHES, name)


# ---- natural node coordinates and polynomial spaces of the interpolations -------------------------
LOB = [-1.0, 0.0, 1.0]
NODES = {  # gmsh type -> natural coordinates of the interpolation's nodes, in its node order
    1: [(-1.0,), (1.0,)],
    8: [(x,) for x in LOB],
    2: [(0, 0), (1, 0), (0, 1)],
    9: [(0, 0), (1, 0), (0, 1), (0.5, 0), (0.5, 0.5), (0, 0.5)],
    3: [(-1, -1), (1, -1), (1, 1), (-1, 1)],
    16: [(-1, -1), (1, -1), (1, 1), (-1, 1), (0, -1), (1, 0), (0, 1), (-1, 0)],
    10: [(x, y) for y in LOB for x in LOB],
    4: [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1)],
    5: [(-1, -1, -1), (1, -1, -1), (1, 1, -1), (-1, 1, -1), (-1, -1, 1), (1, -1, 1), (1, 1, 1), (-1, 1, 1)],
    12: [(x, y, z) for z in LOB for y in LOB for x in LOB],
    6: [(0, 0, -1), (1, 0, -1), (0, 1, -1), (0, 0, 1), (1, 0, 1), (0, 1, 1)],
}


def monomials(t):
    """Exponent tuples spanning the interpolation's polynomial space."""
    if t in (1, 8):
        return [(i,) for i in range(2 if t == 1 else 3)]
    if t == 2:
        return [(0, 0), (1, 0), (0, 1)]
    if t == 9:
        return [(i, j) for i in range(3) for j in range(3) if i + j <= 2]
    if t == 3:
        return [(i, j) for i in range(2) for j in range(2)]
    if t == 16:
        return [(i, j) for i in range(3) for j in range(3) if (i, j) != (2, 2)]
    if t == 10:
        return [(i, j) for i in range(3) for j in range(3)]
    if t == 4:
        return [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1)]
    if t == 5:
        return list(itertools.product(range(2), repeat=3))
    if t == 12:
        return list(itertools.product(range(3), repeat=3))
    if t == 6:
        return [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (0, 1, 1)]
    raise KeyError(t)


def shape_derivatives(t, xi):
    """dN/dxi (n x dim) at xi: the Lagrange basis of the polynomial space through NODES[t]."""
    P = np.array(NODES[t], float)
    E = np.array(monomials(t))
    V = np.prod(P[:, None, :] ** E[None, :, :], axis=2)  # V[a, m] = p_m(node a)
    coef = np.linalg.solve(V, np.eye(len(P)))  # N_b = sum_m coef[m, b] p_m
    xi = np.asarray(xi, float)
    dP = np.zeros((len(E), P.shape[1]))
    for m, e in enumerate(E):
        for d in range(P.shape[1]):
            if e[d]:
                ee = e.copy()
                ee[d] -= 1
                dP[m, d] = e[d] * np.prod(xi ** ee)
    return coef.T @ dP


def det_measure(X, dN):
    J = X.T @ dN  # dim x edim
    if J.shape[0] == J.shape[1]:
        return np.linalg.det(J)
    return np.sqrt(np.linalg.det(J.T @ J))  # length / area of a line / surface element


@pytest.mark.parametrize("t", sorted(NODES))
def test_shape_functions_interpolate(t):
    """The restated bases are nodal (N_b(node a) = delta_ab) and reproduce linear fields."""
    P = np.array(NODES[t], float)
    E = np.array(monomials(t))
    V = np.prod(P[:, None, :] ** E[None, :, :], axis=2)
    coef = np.linalg.solve(V, np.eye(len(P)))
    assert np.allclose(V @ coef, np.eye(len(P)))
    for xi in P[:2] * 0.3:
        assert np.allclose(shape_derivatives(t, xi).sum(0), 0.0, atol=1e-12)
        assert np.allclose(P.T @ shape_derivatives(t, xi), np.eye(P.shape[1]), atol=1e-12)


# ---- MeshGmsh.cpp test cases ---------------------------------------------------------------------
def test_physical_groups():
    m = N.GmshMesh(path("quad.msh"))
    assert len(m.physical_group(1)) > 0
    with pytest.raises(N.NutoError, match="Couldn't find group with ID -42"):
        m.physical_group(-42)
    assert np.array_equal(m.physical_group("domain"), m.physical_group("DoMaIn"))
    with pytest.raises(N.NutoError, match="physicalName ARE_YOU_THERE"):
        m.physical_group("are_you_there?")


def test_non_contiguous_node_numbering():
    a = N.GmshMesh(path("quad.msh"))
    b = N.GmshMesh(path("quadNoncontiguous.msh"))
    assert b.n_nodes == a.n_nodes and b.n_elements == a.n_elements


CASES = [("quad1.msh", 4), ("quadSerendipity2.msh", 8), ("triangle1.msh", 3), ("triangle2.msh", 6),
         ("line1.msh", 2), ("line2.msh", 3), ("tet1.msh", 4), ("hex1.msh", 8), ("prism1.msh", 6),
         ("pyramid1.msh", 5), ("quadL2.msh", 9), ("brickL2.msh", 27)]


@pytest.mark.parametrize("name,n", CASES)
def test_check_mesh(name, n):
    m = N.GmshMesh(path(name))
    assert m.n_nodes == n
    nodes = m.element_nodes(0)
    assert len(nodes) == n
    t = int(m.gmsh_type[0])
    if t not in NODES:  # the pyramid's rational basis is not restated; node count only
        assert t == 7
        return
    X = m.coords[nodes]
    dets = np.array([det_measure(X, shape_derivatives(t, xi)) for xi in NODES[t]])
    # CheckMesh: one sign at every node (det > -1e-10); the reader already required det > 0 at the
    # natural origin, so here every node must be positive
    assert np.all(dets > -1e-10), dets


def test_binary_import():
    m = N.GmshMesh(path("binary.msh"))
    assert m.n_nodes == 8 and m.n_elements == 7
    assert len(m.element_nodes(0)) == 3


def test_inverted_triangle():
    with pytest.raises(N.NutoError, match="Negative Jacobian not allowed"):
        N.GmshMesh(path("invertedTriangle.msh"))


# ---- behaviour of the format ---------------------------------------------------------------------
def test_dimension_from_coordinates():
    assert N.GmshMesh(path("line2.msh")).dim == 1
    assert N.GmshMesh(path("line1.msh")).dim == 2
    assert N.GmshMesh(path("quad.msh")).dim == 2
    assert N.GmshMesh(path("hex1.msh")).dim == 3


def read_ascii(name):
    """(nodes {id: xyz}, elements [(id, type, tags, nodes)], names) of an ASCII MSH 2 file."""
    lines = open(path(name)).read().split("\n")
    nodes, elems, names = {}, [], []
    i = 0
    while i < len(lines):
        if lines[i].startswith("$PhysicalNames"):
            k = int(lines[i + 1])
            names = lines[i + 2:i + 2 + k]
            i += 2 + k
        elif lines[i].startswith("$Nodes"):
            k = int(lines[i + 1])
            for ln in lines[i + 2:i + 2 + k]:
                f = ln.split()
                nodes[int(f[0])] = [float(v) for v in f[1:4]]
            i += 2 + k
        elif lines[i].startswith("$Elements"):
            k = int(lines[i + 1])
            for ln in lines[i + 2:i + 2 + k]:
                f = [int(v) for v in ln.split()]
                elems.append((f[0], f[1], f[3:3 + f[2]], f[3 + f[2]:]))
            i += 2 + k
        else:
            i += 1
    return nodes, elems, names


def write_binary(fn, nodes, elems, names, block):
    """MSH 2.2 binary: elements of one type and tag count grouped in headers of up to `block`."""
    with open(fn, "wb") as f:
        f.write(b"$MeshFormat\n2.2 1 8\n" + struct.pack("<i", 1) + b"\n$EndMeshFormat\n")
        if names:
            f.write(("$PhysicalNames\n%d\n" % len(names) + "\n".join(names) + "\n$EndPhysicalNames\n").encode())
        f.write(b"$Nodes\n%d\n" % len(nodes))
        for nid, x in nodes.items():
            f.write(struct.pack("<i3d", nid, *x))
        f.write(b"\n$EndNodes\n$Elements\n%d\n" % len(elems))
        i = 0
        while i < len(elems):
            j = i
            while (j < len(elems) and j - i < block and elems[j][1] == elems[i][1]
                   and len(elems[j][2]) == len(elems[i][2])):
                j += 1
            f.write(struct.pack("<3i", elems[i][1], j - i, len(elems[i][2])))
            for e in elems[i:j]:
                f.write(struct.pack("<%di" % (1 + len(e[2]) + len(e[3])), e[0], *e[2], *e[3]))
            i = j
        f.write(b"\n$EndElements\n")


@pytest.mark.parametrize("name", ["quad.msh", "brickL2.msh", "quadL2.msh"])
@pytest.mark.parametrize("block", [1, 3, 1000])
def test_binary_blocks_equal_ascii(tmp_path, name, block):
    nodes, elems, names = read_ascii(name)
    fn = str(tmp_path / "b.msh")
    write_binary(fn, nodes, elems, names, block)
    a, b = N.GmshMesh(path(name)), N.GmshMesh(fn)
    assert a.dim == b.dim
    for k in ("coords", "gmsh_type", "offsets", "conn", "physical"):
        assert np.array_equal(getattr(a, k), getattr(b, k)), k


def test_reordering_of_lobatto_quad(tmp_path):
    """gmsh numbers the 9-node quad corners, edges, centre; the interpolation numbers the 3 x 3
    grid row by row (MeshGmsh.cpp:402-411)."""
    m = N.GmshMesh(path("quadL2.msh"))
    X = m.coords[m.element_nodes(0)]
    lo, hi = X.min(0), X.max(0)
    grid = np.round((X - lo) / (hi - lo) * 2).astype(int)  # the file's element is an axis-aligned square
    assert [tuple(g) for g in grid] == [(i, j) for j in range(3) for i in range(3)]


def write_ascii(fn, body, version="2.2"):
    with open(fn, "w") as f:
        f.write(f"$MeshFormat\n{version} 0 8\n$EndMeshFormat\n" + body)


NODES_TRI = "$Nodes\n3\n1 0 0 0\n2 1 0 0\n3 0 1 0\n$EndNodes\n"


def test_format_errors(tmp_path):
    fn = str(tmp_path / "m.msh")
    write_ascii(fn, NODES_TRI + "$Elements\n1\n1 2 2 1 1 1 2 3\n$EndElements\n", version="1.0")
    with pytest.raises(N.NutoError, match="version 2.0 or higher"):
        N.GmshMesh(fn)
    write_ascii(fn, NODES_TRI + "$Periodic\n0\n$EndPeriodic\n")
    with pytest.raises(N.NutoError, match="Unhandled gmsh section type"):
        N.GmshMesh(fn)
    write_ascii(fn, NODES_TRI + "$Elements\n1\n1 21 2 1 1 1 2 3 1 2 3 1 2 3 1\n$EndElements\n")
    with pytest.raises(N.NutoError, match="Unhandled gmsh element type 21"):
        N.GmshMesh(fn)
    write_ascii(fn, NODES_TRI + "$Elements\n1\n1 2 2 1 1 1 2 4\n$EndElements\n")
    with pytest.raises(N.NutoError, match="unknown node"):
        N.GmshMesh(fn)
    write_ascii(fn, NODES_TRI + "$Elements\n1\n1 2 2 1 1 1 2 3\n$Elementz\n")
    with pytest.raises(N.NutoError, match="Expected \\$EndElements"):
        N.GmshMesh(fn)
    with pytest.raises(N.NutoError, match="Error opening input file"):
        N.GmshMesh(str(tmp_path / "missing.msh"))


def test_points_skipped_and_groups(tmp_path):
    fn = str(tmp_path / "m.msh")
    write_ascii(fn, '$PhysicalNames\n2\n2 7 "Plate"\n0 9 "corner"\n$EndPhysicalNames\n' + NODES_TRI +
                "$Elements\n3\n1 15 2 9 1 1\n2 2 2 7 1 1 2 3\n3 1 2 4 1 2 3\n$EndElements\n")
    m = N.GmshMesh(fn)
    assert m.n_elements == 2 and list(m.gmsh_type) == [2, 1]
    assert list(m.physical_group("plate")) == [0] and list(m.physical_group(4)) == [1]
    with pytest.raises(N.NutoError):  # the point element (and its group) is dropped
        m.physical_group("corner")


def test_cells_for_assembly():
    m = N.GmshMesh(path("hex1.msh"))
    X, C = m.cells(N.HEX8)
    assert C.shape == (1, 8) and np.array_equal(C[0], m.element_nodes(0))
    q = N.GmshMesh(path("quad1.msh"))
    X, C = q.cells(N.QUAD4, int(q.physical[0]))
    assert C.shape == (1, 4) and X.shape == (4, 2)
    X, C = N.GmshMesh(path("quad.msh")).cells(N.QUAD4, "domain")  # triangles only
    assert C.shape == (0, 4)
    with pytest.raises(N.NutoError):  # a 3D element type from a 2D mesh
        N.GmshMesh(path("quad.msh")).cells(N.HEX8)
    t = N.GmshMesh(path("MeshCompanionGmsh.msh"))
    X, C = t.cells(N.TET10)
    assert C.shape == (t.n_elements, 10) and X.shape == (2768, 3)


def test_isoparametric_tet10_numbering():
    """Ten-node coordinate elements with the ten-node dof interpolation: every dof node sits on a
    coordinate node (Kronecker property), so AddDofInterpolation's coordinate matching
    (MeshFemDofConvert.cpp:164-210) numbers the file's nodes by first touch in cell order."""
    import oracle as O

    X, C = N.GmshMesh(path("MeshCompanionGmsh.msh")).cells(N.TET10)
    ranks, nn = N.dof_layer(N.TET10, C, N.TET10)
    assert nn == 2768  # every node of the file is used
    om = O.Mesh.from_arrays(N.TET10, X, C)
    lay = om.dof_layer(om.add_dof(N.TET10, 3))
    assert np.array_equal(lay["numbers"][lay["conn"], 0] // 3, ranks)
    # first touch: the ranks in cell order are 0, 1, 2, ... at their first appearance
    flat = ranks.ravel()
    _, first = np.unique(flat, return_index=True)
    assert np.array_equal(flat[np.sort(first)], np.arange(nn))
    with pytest.raises(N.NutoError):  # hex20 dofs over tet10 coordinates
        N.dof_layer(N.TET10, C, N.HEX20)


# ---- Morton codes ---------------------------------------------------------------------------------
def spread(v, step, bits):
    return sum(((v >> b) & 1) << (step * b) for b in range(bits))


def test_morton_example_and_round_trip():
    assert N.morton_encode3d(2, 1, 1) == 14  # MortonOrder.cpp example (x = 2, y = z = 1)
    rng = np.random.default_rng(3)
    for x, y, z in rng.integers(0, 1024, (200, 3)).tolist():
        c = N.morton_encode3d(x, y, z)
        assert c == spread(x, 3, 10) | spread(y, 3, 10) << 1 | spread(z, 3, 10) << 2
        assert N.morton_decode3d(c) == (x, y, z)
    for x, y in rng.integers(0, 65536, (200, 2)).tolist():
        c = N.morton_encode2d(x, y)
        assert c == spread(x, 2, 16) | spread(y, 2, 16) << 1
        assert N.morton_decode2d(c) == (x, y)
    assert N.morton_encode3d(1024 + 5, 0, 0) == N.morton_encode3d(5, 0, 0)  # 10 bits per axis


def test_morton_neighbor():
    k = N.morton_encode3d(4, 7, 9)
    assert N.morton_neighbor3d(k, 1, -2, 3) == N.morton_encode3d(5, 5, 12)
    assert N.morton_neighbor3d(k, -5, 0, 0) == 0xFFFFFFFF
    assert N.morton_neighbor3d(k, 0, 0, 0) == k


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_morton_order(dim):
    rng = np.random.default_rng(dim)
    X = rng.uniform(-2.0, 5.0, (500, dim))
    X[7] = X[3]  # a tie keeps the input order
    o = N.morton_order(X)
    steps = 1024 if dim == 3 else 65536
    lo, hi = X.min(0), X.max(0)
    q = np.clip(((X - lo) / (hi - lo) * steps).astype(np.int64), 0, steps - 1)
    bits = 10 if dim == 3 else 16
    key = [sum(spread(int(q[i, d]), dim if dim > 1 else 2, bits) << d for d in range(dim)) for i in range(len(X))]
    ref = sorted(range(len(X)), key=lambda i: (key[i], i))
    assert list(o) == ref
    assert sorted(o) == list(range(len(X)))


def test_cpp_mesh_gmsh_mirror():
    """The C++ MeshGmsh mirror (nuto_amd/host) through its own test program (host only)."""
    import subprocess

    exe = os.path.join(os.path.dirname(__file__), "cpp", "test_host_gmsh")
    assert os.path.exists(exe), "build tests/cpp first (__graft_entry__.build())"
    r = subprocess.run([exe, MESHES], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), r.stdout + r.stderr
