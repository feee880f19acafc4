"""Gradient damage on a hex8 mesh with a node beyond the node kernel's bounds: five hexahedra fan
around a vertical axis in two layers, so the axis node of the middle level has 10 incident cells
and 33 neighbours (the node kernel takes at most 8 and 32). That node's rows come from the generic
row-owner kernel, every other node's from the node kernel; R and K against the oracle, the
outside node's rows bit-identical to the all-generic assembly, and the allocation-free entry point
on the same mesh. The adjoint and batches still need every node in bounds (refused here)."""
import numpy as np
import pytest
import torch

import nuto_amd as N
from nuto_amd import binding as B
from pairs import Pair, check_values, to_dev
from test_gpu_parity import gdm_params

pytestmark = pytest.mark.gpu


def fan_mesh(sectors=5, layers=2, outer=1.2):
    per = 1 + 2 * sectors  # axis node, P_k on the unit circle, Q_k between P_k and P_k+1
    X = []
    for lev in range(layers + 1):
        z = float(lev)
        X.append((0.0, 0.0, z))
        for k in range(sectors):
            t = 2 * np.pi * k / sectors
            X.append((np.cos(t), np.sin(t), z))
        for k in range(sectors):
            t = 2 * np.pi * (k + 0.5) / sectors
            X.append((outer * np.cos(t), outer * np.sin(t), z))
    C = []
    for lay in range(layers):
        b0, b1 = lay * per, (lay + 1) * per
        for k in range(sectors):
            kn = (k + 1) % sectors
            quad = [0, 1 + k, 1 + sectors + k, 1 + kn]  # counterclockwise seen from +z
            C.append([b0 + v for v in quad] + [b1 + v for v in quad])
    return np.array(X), np.array(C, np.int32)


def test_outside_node_by_the_generic_kernel(monkeypatch):
    X, C = fan_mesh()
    p = Pair(N.HEX8, N.HEX8, X, C, *gdm_params(), True)
    u, e, kap = p.inputs(u_range=4e-4, k_range=1.5e-4)
    ctx = p.context()
    R, K = ctx.build_vector_matrix(to_dev(u), to_dev(e), to_dev(kap))
    ref = p.orc.assemble(u, e, kap)
    rp, ci = ctx.pattern()
    assert np.array_equal(rp, ref["row_ptr"]) and np.array_equal(ci, ref["col"])
    check_values(R.cpu().numpy(), ref["R"], ref["R_bound"], "R")
    check_values(K.cpu().numpy(), ref["val"], ref["bound"], "K")
    # the axis node of the middle level (dof rank of coordinate node 11) is the one outside
    axis = int(p.ranks[np.argwhere(C == 11)[0][0], np.argwhere(C == 11)[0][1]])
    monkeypatch.setenv("NUTO_AMD_GENERIC", "1")
    Rg, Kg = ctx.build_vector_matrix(to_dev(u), to_dev(e), to_dev(kap))
    monkeypatch.delenv("NUTO_AMD_GENERIC")
    Rg, Kg, Rn, Kn = Rg.cpu().numpy(), Kg.cpu().numpy(), R.cpu().numpy(), K.cpu().numpy()
    rows = [3 * axis, 3 * axis + 1, 3 * axis + 2, p.n_u + axis]
    for r in rows:
        assert np.array_equal(Kn[rp[r]:rp[r + 1]], Kg[rp[r]:rp[r + 1]]), r
        assert Rn[r] == Rg[r], r
    # no node kernel for the adjoint on this mesh
    with pytest.raises(N.NutoError) as ei:
        ctx.vjp(to_dev(u), to_dev(e), to_dev(kap), to_dev(np.ones(ctx.n_rows)))
    assert ei.value.code == 5


def test_outside_node_allocation_free():
    X, C = fan_mesh()
    p = Pair(N.HEX8, N.HEX8, X, C, *gdm_params(), True)
    u, e, kap = (to_dev(a) for a in p.inputs(u_range=4e-4, k_range=1.5e-4))
    ctx = p.context()
    R = torch.full((ctx.n_rows,), float("nan"), dtype=torch.float64, device="cuda")
    K = torch.full((ctx.nnz,), float("nan"), dtype=torch.float64, device="cuda")
    assert B.lib().nuto_build_vector_matrix_noalloc(ctx.h, *B._ptrs(u, e, kap, R, K)) == 0
    Rr, Kr = ctx.build_vector_matrix(u, e, kap)
    torch.cuda.synchronize()
    assert torch.equal(R, Rr) and torch.equal(K, Kr)
