"""Gradient damage on a hex8 mesh with a node beyond the node kernel's bounds: five hexahedra fan
around a vertical axis in two layers, so the axis node of the middle level has 10 incident cells
and 33 neighbours (the node kernel takes at most 8 and 32). That node's rows come from the generic
row-owner kernel, every other node's from the node kernel; R and K against the oracle, the
outside node's rows bit-identical to the all-generic assembly, and the allocation-free entry point
on the same mesh. The adjoint of that node (and d_theta of the cells whose local node 0 it is) comes
from the adjoint's list kernel: d_state against K^T lambda of the oracle's K, d_theta against
central finite differences of the oracle's R. Batches holding this mesh: tests/test_gpu_batch.py."""
import numpy as np
import pytest
import torch

import nuto_amd as N
from nuto_amd import binding as B
from nuto_amd import synthetic as S
from pairs import Pair, check_values, to_dev
from test_gpu_adjoint import kt_lambda
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


@pytest.mark.parametrize("app", [N.FULL, N.UNILATERAL])
def test_outside_node_adjoint(app):
    X, C = fan_mesh()
    p = Pair(N.HEX8, N.HEX8, X, C, *gdm_params(app), True)
    th = S.material_fields(p.ncells)
    u, e, kap = p.inputs(u_range=4e-4, k_range=1.5e-4)
    p.orc.set_theta(*S.perturbed_fields(*th))
    r_obs = p.orc.assemble(u, e, kap, want_k=False)["R"]
    p.orc.set_theta(*th)
    ref = p.orc.assemble(u, e, kap)
    lam = ref["R"] - r_obs
    ctx = p.context()
    ctx.set_material_fields(*th)
    d_state, d_theta = ctx.vjp(to_dev(u), to_dev(e), to_dev(kap), to_dev(lam))
    got = d_state.cpu().numpy()
    kt, kb = kt_lambda(ref, lam, len(lam))
    err = np.abs(got - kt)
    assert (err <= 1e-12 * kb + 1e-300).all(), f"max err/bound {np.max(err / (kb + 1e-300)):.3e}"
    # the cells of the upper layer have the middle axis node (beyond the bounds) as local node 0
    assert all(C[c][0] == 11 for c in range(5, 10))

    def L(E, ft, gf):
        p.orc.set_theta(E, ft, gf)
        R = p.orc.assemble(u, e, kap, want_k=False)["R"]
        return 0.5 * np.sum((R - r_obs) ** 2)

    got_t = d_theta.cpu().numpy()
    base = [np.array(v, dtype=np.float64) for v in th]
    fd = np.zeros_like(got_t)
    for k in range(3):
        for c in range(p.ncells):
            h = 1e-6 * base[k][c]
            plus = [v.copy() for v in base]
            minus = [v.copy() for v in base]
            plus[k][c] += h
            minus[k][c] -= h
            fd[k, c] = (L(*plus) - L(*minus)) / (2 * h)
    p.orc.set_theta(*th)
    scale = np.abs(fd).max(axis=1, keepdims=True)
    assert (np.abs(got_t - fd) <= 1e-6 * scale).all(), f"max rel {np.max(np.abs(got_t - fd) / scale):.3e}"


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
