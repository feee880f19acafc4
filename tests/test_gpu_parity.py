"""Parity of the HIP path (through the C ABI) with the oracle and with the reference's golden data.

Bar: pattern, row_ptr, col and dof numbering bit-equal; R and K within 1e-12 * sum|contributions|
(the tolerance of the path; FMA contraction and summation order differ from the reference build).
"""
import os

import numpy as np
import pytest

import nuto_amd as N
import oracle as O
from nuto_amd import synthetic as S
from pairs import Pair, check_values, make_mesh, to_dev

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def gdm_params(app=N.FULL):
    dev = N.gradient_damage_params(damage_application=app)
    orc = O.Params(kind=1, dim=3, E=30000.0, nu=0.2, k=10.0, kappa0=1e-4, beta=100.0, alpha=0.99, c=4.0,
                   damage_application=app)
    return dev, orc


def ld_params(app=N.FULL):
    dev = N.local_damage_params(damage_application=app)
    orc = O.Params(kind=2, dim=3, E=30000.0, nu=0.2, k=10.0, kappa0=1e-4, beta=100.0, alpha=0.99,
                   damage_application=app)
    return dev, orc


def le_params(dim, E=30000.0, nu=0.2):
    return N.linear_elastic_params(dim, E, nu), O.Params(kind=0, dim=dim, E=E, nu=nu, plane_state=0)


def run_pair(p, u, e, kap, want=("R", "K"), k_floor=0.0):
    ctx = p.context()
    R, K = ctx.build_vector_matrix(to_dev(u), to_dev(e), to_dev(kap))
    ctx.synchronize()
    rp, ci = ctx.pattern()
    ref = p.orc.assemble(u, e, kap)
    assert p.numbering_matches(), "dof numbering differs from the oracle"
    assert np.array_equal(rp, ref["row_ptr"]), "row_ptr differs"
    assert np.array_equal(ci, ref["col"]), "col differs"
    mr = check_values(R.cpu().numpy(), ref["R"], ref["R_bound"], "R")
    kf = k_floor * np.abs(ref["val"]).max(initial=0.0)
    mk = check_values(K.cpu().numpy(), ref["val"], ref["bound"], "K", kf)
    # BuildVector / BuildMatrix alone meet the same bar (separate kernel variants; FMA contraction
    # may differ in the last bit from the fused variant)
    R2 = ctx.build_vector(to_dev(u), to_dev(e), to_dev(kap))
    K2 = ctx.build_matrix(to_dev(u), to_dev(e), to_dev(kap))
    check_values(R2.cpu().numpy(), ref["R"], ref["R_bound"], "R (BuildVector)")
    check_values(K2.cpu().numpy(), ref["val"], ref["bound"], "K (BuildMatrix)", kf)
    return ctx, R, K, mr, mk


def test_quad4_linear_elastic_config1_material():
    X, C = make_mesh(N.QUAD4, (8, 6))
    dev, orc = le_params(2)
    p = Pair(N.QUAD4, N.QUAD4, X, C, dev, orc, False)
    u = S.uniform(S.SEED_U, p.n_u, -1e-3, 1e-3)
    run_pair(p, u, None, None)


def test_quad4_linear_elastic_config1_full_size():
    """The plane-stress workload at its full size: 200 x 200 quad4 cells of the unit square, 2x2
    Gauss, E = 30000, nu = 0.2, u uniform in [-1e-3, 1e-3] (80,802 dofs, 1,444,804 nonzeros)."""
    X, C = make_mesh(N.QUAD4, (200, 200))
    dev, orc = le_params(2)
    p = Pair(N.QUAD4, N.QUAD4, X, C, dev, orc, False)
    assert p.n_u == 80802
    u = S.uniform(S.SEED_U, p.n_u, -1e-3, 1e-3)
    ctx, R, K, mr, mk = run_pair(p, u, None, None)
    assert K.numel() == 1444804


def test_hex8_gradient_damage_headline_cells():
    """16^3 cells of 1 mm (the cell size of the 100 mm cube workload) with its material and state
    distributions: most nodes are interior, so the node kernel's full 8-cell neighbourhoods, the
    Morton traversal and the affine mode are exercised against the oracle."""
    n = 16
    X, C = make_mesh(N.HEX8, (n, n, n), (float(n),) * 3)
    dev, orc = gdm_params(N.FULL)
    p = Pair(N.HEX8, N.HEX8, X, C, dev, orc, True)
    u, e, kap = p.inputs()
    run_pair(p, u, e, kap)


def test_quad4_fixture_4x4():
    f = np.load(os.path.join(GOLD, "quad4_le_4x4.npz"))
    X, C = make_mesh(N.QUAD4, (4, 4))
    assert np.array_equal(X, f["coords"]) and np.array_equal(C, f["conn"])
    ranks, nn = N.dof_layer(N.QUAD4, C, N.QUAD4)
    assert np.array_equal(f["dof_numbers"][f["dof_conn"], 0] // 2, ranks)
    ctx = N.Context()
    ctx.set_cells(N.QUAD4, X, C)
    ctx.set_dof_type(0, N.QUAD4, 2, 2 * nn, N.cell_dofs(ranks, 2))
    ctx.set_integrand(N.linear_elastic_params(2, 30000.0, 0.2))
    ctx.set_rule(0, 2)
    ctx.build_pattern()
    R, K = ctx.build_vector_matrix(to_dev(f["u"]))
    rp, ci = ctx.pattern()
    assert np.array_equal(rp, f["row_ptr"]) and np.array_equal(ci, f["col"])
    check_values(R.cpu().numpy(), f["R"], f["R_bound"], "R")
    check_values(K.cpu().numpy(), f["val"], f["bound"], "K")


@pytest.mark.parametrize("app", [N.FULL, N.UNILATERAL])
def test_hex8_gradient_damage(app):
    X, C = make_mesh(N.HEX8, (4, 3, 3), (4.0, 3.0, 3.0))
    dev, orc = gdm_params(app)
    p = Pair(N.HEX8, N.HEX8, X, C, dev, orc, True)
    u, e, kap = p.inputs()
    run_pair(p, u, e, kap)


@pytest.mark.parametrize("app", [N.FULL, N.UNILATERAL])
def test_hex8_gradient_damage_decreasing_interaction(app):
    """NonlocalInteraction::Decreasing(R = 0.005, eta = 5) on a jittered mesh, damaged and undamaged
    IPs mixed: the generic row-owner kernel (the node kernel takes Constant only) against the
    oracle; the adjoint and batches refuse it."""
    X, C = make_mesh(N.HEX8, (4, 3, 3), (4.0, 3.0, 3.0), jitter=0.1)
    dev = N.gradient_damage_params(damage_application=app, interaction=N.DECREASING, interaction_r=0.005,
                                   interaction_eta=5.0)
    orc = O.Params(kind=1, dim=3, E=30000.0, nu=0.2, k=10.0, kappa0=1e-4, beta=100.0, alpha=0.99, c=4.0,
                   damage_application=app, interaction=1, interaction_r=0.005, interaction_eta=5.0)
    p = Pair(N.HEX8, N.HEX8, X, C, dev, orc, True)
    u, e, kap = p.inputs()
    ctx, R, K, mr, mk = run_pair(p, u, e, kap)
    with pytest.raises(N.NutoError):
        ctx.vjp(to_dev(u), to_dev(e), to_dev(kap), to_dev(np.ones(p.n_u + p.n_e)))
    b = N.Batch()
    with pytest.raises(N.NutoError):
        b.set_integrand(dev)


def test_hex8_gradient_damage_fixture():
    f = np.load(os.path.join(GOLD, "hex8_gdm_3x3x3.npz"))
    X, C = make_mesh(N.HEX8, (3, 3, 3), (3.0, 3.0, 3.0))
    assert np.array_equal(X, f["coords"])
    ranks, nn = N.dof_layer(N.HEX8, C, N.HEX8)
    ctx = N.Context()
    ctx.set_cells(N.HEX8, X, C)
    ctx.set_dof_type(0, N.HEX8, 3, 3 * nn, N.cell_dofs(ranks, 3))
    ctx.set_dof_type(1, N.HEX8, 1, nn, N.cell_dofs(ranks, 1))
    ctx.set_integrand(N.gradient_damage_params())
    ctx.set_rule(0, 2)
    ctx.build_pattern()
    R, K = ctx.build_vector_matrix(to_dev(f["u"]), to_dev(f["e"]), to_dev(f["kappa"]))
    rp, ci = ctx.pattern()
    assert np.array_equal(rp, f["row_ptr"]) and np.array_equal(ci, f["col"])
    check_values(R.cpu().numpy(), f["R"], f["R_bound"], "R")
    check_values(K.cpu().numpy(), f["val"], f["bound"], "K")


def test_hex8_gradient_damage_material_fields():
    X, C = make_mesh(N.HEX8, (3, 3, 2), (3.0, 3.0, 2.0))
    dev, orc = gdm_params()
    p = Pair(N.HEX8, N.HEX8, X, C, dev, orc, True)
    n = p.ncells
    E = S.lognormal_mean_cv(S.SEED_E_MOD, n, 30000.0, 0.1)
    ft = S.lognormal_mean_cv(S.SEED_FT, n, 3.0, 0.1)
    gf = S.lognormal_mean_cv(S.SEED_GF, n, 0.1, 0.15)
    p.orc.set_theta(E, ft, gf)
    u, e, kap = p.inputs()
    ctx = p.context()
    ctx.set_material_fields(E, ft, gf)
    R, K = ctx.build_vector_matrix(to_dev(u), to_dev(e), to_dev(kap))
    ref = p.orc.assemble(u, e, kap)
    check_values(R.cpu().numpy(), ref["R"], ref["R_bound"], "R")
    check_values(K.cpu().numpy(), ref["val"], ref["bound"], "K")


def test_brick8n_golden_through_device():
    # Brick8N.py:18-49 golden stiffness/force (reference result files), via the C ABI
    X = O.local_coords(O.HEX8)
    C = np.arange(8, dtype=np.int32)[None, :]
    ranks = np.arange(8, dtype=np.int32)[None, :]
    ctx = N.Context()
    ctx.set_cells(N.HEX8, X, C)
    ctx.set_dof_type(0, N.HEX8, 3, 24, N.cell_dofs(ranks, 3))
    ctx.set_integrand(N.linear_elastic_params(3, 10.0, 0.25))
    ctx.set_rule(0, 2)
    ctx.build_pattern()
    u = np.zeros(24)
    for n in (1, 2, 5, 6):
        u[3 * n:3 * n + 3] = 0.2
    R, K = ctx.build_vector_matrix(to_dev(u))
    rp, ci = ctx.pattern()
    Kd = O.csr_to_dense(rp, ci, K.cpu().numpy(), 24)
    Kg = np.loadtxt(os.path.join(GOLD, "ref_results", "Brick8NStiffness.txt"), skiprows=1)
    Rg = np.loadtxt(os.path.join(GOLD, "ref_results", "Brick8NInternalforce.txt"), skiprows=1)
    np.testing.assert_allclose(Kd, Kg, rtol=0, atol=5e-12 * np.abs(Kg).max())
    np.testing.assert_allclose(R.cpu().numpy(), Rg, rtol=0, atol=5e-12 * np.abs(Rg).max())


@pytest.mark.parametrize("ct,dt", [(N.TET4, N.TET4), (N.TET4, N.TET10), (N.HEX8, N.HEX8), (N.HEX8, N.HEX20)])
def test_local_damage_mixed_types_jittered(ct, dt):
    X, C = make_mesh(ct, (2, 2, 2), (2.0, 2.0, 2.0), jitter=0.1)
    dev, orc = ld_params()
    p = Pair(ct, dt, X, C, dev, orc, False)
    u, _, kap = p.inputs(u_range=5e-4, k_range=1e-4)
    run_pair(p, u, None, kap)


@pytest.mark.parametrize("ct,dt", [(N.TET4, N.TET10), (N.HEX8, N.HEX20)])
def test_linear_elastic_quadratic(ct, dt):
    X, C = make_mesh(ct, (2, 1, 2), (2.0, 1.0, 2.0), jitter=0.05)
    dev, orc = le_params(3)
    p = Pair(ct, dt, X, C, dev, orc, False)
    u, _, _ = p.inputs()
    run_pair(p, u, None, None)


def test_history_update_gradient_damage():
    X, C = make_mesh(N.HEX8, (3, 2, 2), (3.0, 2.0, 2.0))
    dev, orc = gdm_params()
    p = Pair(N.HEX8, N.HEX8, X, C, dev, orc, True)
    u, e, kap = p.inputs()
    ctx = p.context()
    k_dev = to_dev(kap.copy())
    ctx.update_history(to_dev(u), to_dev(e), k_dev)
    xi, _ = O.rule(0, 3, 2)
    Nip = np.stack([O.shape(O.HEX8, x)[0] for x in xi])  # nip x 8
    eip = (e[p.ranks] @ Nip.T)  # cells x nip
    ref = np.maximum(kap.reshape(p.nip, p.ncells).T, eip).T.ravel()
    np.testing.assert_allclose(k_dev.cpu().numpy(), ref, rtol=1e-14, atol=1e-20)


def test_deterministic_bits():
    X, C = make_mesh(N.HEX8, (4, 4, 4), (4.0, 4.0, 4.0))
    dev, orc = gdm_params()
    p = Pair(N.HEX8, N.HEX8, X, C, dev, orc, True)
    u, e, kap = p.inputs()
    ctx = p.context()
    a = ctx.build_vector_matrix(to_dev(u), to_dev(e), to_dev(kap))
    b = ctx.build_vector_matrix(to_dev(u), to_dev(e), to_dev(kap))
    assert np.array_equal(a[0].cpu().numpy(), b[0].cpu().numpy())
    assert np.array_equal(a[1].cpu().numpy(), b[1].cpu().numpy())


def test_error_paths():
    ctx = N.Context()
    with pytest.raises(N.NutoError) as ei:
        ctx.build_vector(to_dev(np.zeros(3)))
    assert ei.value.code == 2  # state: no cells / numbering (SimpleAssembler.cpp:16-23)
    X, C = make_mesh(N.HEX8, (1, 1, 1))
    ctx.set_cells(N.HEX8, X, C)
    with pytest.raises(N.NutoError) as ei:
        ctx.set_dof_type(0, N.QUAD4, 3, 24, np.zeros((1, 12), np.int32))
    assert ei.value.code == 1
    bad = N.cell_dofs(np.arange(8, dtype=np.int32)[None, :], 3)
    bad[:, [0, 1]] = bad[:, [1, 0]]  # components of node 0 swapped
    with pytest.raises(N.NutoError) as ei:
        ctx.set_dof_type(0, N.HEX8, 3, 24, bad)
    assert ei.value.code == 5  # not node-blocked


def test_hex8_node_kernel_matches_generic_kernel(monkeypatch):
    # the specialised hex8 gradient-damage kernel against the generic row-owner kernel and the oracle
    X, C = make_mesh(N.HEX8, (6, 5, 4), (6.0, 5.0, 4.0))
    dev, orc = gdm_params()
    p = Pair(N.HEX8, N.HEX8, X, C, dev, orc, True)
    u, e, kap = p.inputs()
    ctx = p.context()
    Rf, Kf = ctx.build_vector_matrix(to_dev(u), to_dev(e), to_dev(kap))
    monkeypatch.setenv("NUTO_AMD_GENERIC", "1")
    Rg, Kg = ctx.build_vector_matrix(to_dev(u), to_dev(e), to_dev(kap))
    monkeypatch.delenv("NUTO_AMD_GENERIC")
    ref = p.orc.assemble(u, e, kap)
    for R, K in ((Rf, Kf), (Rg, Kg)):
        check_values(R.cpu().numpy(), ref["R"], ref["R_bound"], "R")
        check_values(K.cpu().numpy(), ref["val"], ref["bound"], "K")
