"""Batched assembly: every mesh of a batch against the oracle (pattern bit-equal, R and K within
1e-12 of the contribution magnitudes), all launch modes bit-identical, graph replays identical to
the eager launch, and the batched gradient-damage residual + adjoint against single meshes.
"""
import numpy as np
import pytest
import torch

import nuto_amd as N
import oracle as O
from nuto_amd import synthetic as S
from nuto_amd import workloads as W
from pairs import check_values, to_dev
from test_gpu_gdm_outliers import fan_mesh

pytestmark = pytest.mark.gpu

RULE = {(N.TET4, N.TET4): (1, 1), (N.TET4, N.TET10): (1, 2), (N.HEX8, N.HEX8): (0, 2), (N.HEX8, N.HEX20): (0, 3)}


def ld_oracle_params():
    return O.Params(kind=2, dim=3, E=30000.0, nu=0.2, k=10.0, kappa0=1e-4, beta=100.0, alpha=0.99)


def build(meshes, params):
    b = N.Batch()
    b.set_integrand(params)
    for ct, dt, X, C, R, nn in meshes:
        b.add_mesh(ct, dt, X, C, R, nn)
    sizes = b.build()
    return b, sizes


def mesh_history(kap, info):
    """History of one mesh in NuTo's layout (cell + ip * n_cells) from the batch array."""
    c = np.arange(info["cells"])
    return np.concatenate([kap[info["history"] + c + i * info["history_ld"]] for i in range(info["nip"])])


def batch_inputs(sizes, seed_shift=0):
    u = S.uniform(S.SEED_U + seed_shift, sizes["u"], -5e-3, 5e-3)
    e = S.uniform(S.SEED_E + seed_shift, sizes["e"], 0.0, 5e-4) if sizes["e"] else None
    kap = S.uniform(S.SEED_KAPPA + seed_shift, sizes["history"], 0.0, 3e-4)
    return u, e, kap


@pytest.fixture(scope="module")
def mixed():
    meshes = W.mixed_batch(7, 60.0, 400.0)
    assert {(m[0], m[1]) for m in meshes} == set(W.TYPES)  # every element pair present
    b, sizes = build(meshes, N.local_damage_params())
    u, _, kap = batch_inputs(sizes)
    return meshes, b, sizes, u, kap


@pytest.mark.parametrize("mode", [0, 1, 2, 3, 4, 5, 6])
def test_every_mesh_matches_oracle(mixed, mode):
    meshes, b, sizes, u, kap = mixed
    R, K = b.assemble(to_dev(u), None, to_dev(kap), mode=mode)
    R, K = R.cpu().numpy(), K.cpu().numpy()
    for m, (ct, dt, X, C, ranks, nn) in enumerate(meshes):
        info = b.mesh_info(m)
        om = O.Mesh.from_arrays(ct, X, C)
        Ld = om.add_dof(dt, 3)
        lay = om.dof_layer(Ld)
        assert np.array_equal(lay["numbers"][lay["conn"], 0] // 3, ranks), "numbering differs"
        P = O.Problem(om, ld_oracle_params(), RULE[(ct, dt)][0], RULE[(ct, dt)][1], Ld, -1)
        um = u[info["u"]:info["u"] + 3 * nn]
        ref = P.assemble(um, None, mesh_history(kap, info))
        rp, ci = b.pattern(m)
        assert np.array_equal(rp, ref["row_ptr"]) and np.array_equal(ci, ref["col"]), f"pattern of mesh {m}"
        check_values(R[info["R_u"]:info["R_u"] + info["rows"]], ref["R"], ref["R_bound"], f"R of mesh {m}")
        check_values(K[info["K_u"]:info["K_u"] + info["nnz_u"]], ref["val"], ref["bound"], f"K of mesh {m}")


def test_launch_modes_are_deterministic(mixed):
    """Each mode reproduces its own bits; the per-type and per-mesh launches share one kernel per
    element type and agree bit for bit, for the row-owner (3, 4) and the two-phase form (5, 6) (the
    fused kernel is compiled separately, FMA contraction may differ in the last bit; every mode is
    held to the oracle bar above)."""
    _, b, sizes, u, kap = mixed
    per_mesh = {m: [t.cpu().numpy() for t in b.assemble(to_dev(u), None, to_dev(kap), mode=m)] for m in (4, 6)}
    for mode in (0, 1, 2, 3, 5):
        a = [t.cpu().numpy() for t in b.assemble(to_dev(u), None, to_dev(kap), mode=mode)]
        c = [t.cpu().numpy() for t in b.assemble(to_dev(u), None, to_dev(kap), mode=mode)]
        assert np.array_equal(a[0], c[0]) and np.array_equal(a[1], c[1]), f"mode {mode}"
        if mode in (3, 5):
            ref = per_mesh[mode + 1]
            assert np.array_equal(a[0], ref[0]) and np.array_equal(a[1], ref[1]), f"mode {mode}"


def test_graph_replays_match_eager(mixed):
    _, b, sizes, u, kap = mixed
    du, dk = to_dev(u), to_dev(kap)
    R = torch.empty(sizes["R"], dtype=torch.float64, device="cuda")
    K = torch.empty(sizes["K"], dtype=torch.float64, device="cuda")
    eager = [t.cpu().numpy() for t in b.assemble(du, None, dk, mode=3)]  # the untuned graph replays mode 3
    for rep in range(4):
        R.fill_(float("nan"))
        K.fill_(float("nan"))
        du.mul_(1.0)  # eager work on torch's stream between replays
        b.assemble(du, None, dk, R, K, mode=-1)
        torch.cuda.synchronize()
        assert np.array_equal(R.cpu().numpy(), eager[0]) and np.array_equal(K.cpu().numpy(), eager[1]), rep


def test_autotune_picks_a_batched_mode(mixed):
    """A batch of its own (tuning changes each group's cell kernel and staging chunk): the tuned graph
    replays the chosen mode bit for bit, and every group reports its choices."""
    meshes, _, _, u, kap = mixed
    b, sizes = build(meshes, N.local_damage_params())
    ms, best = b.autotune(to_dev(u), None, to_dev(kap))
    assert best in (0, 1, 2, 3, 5) and all(v > 0 for v in ms)
    for g in range(b.group_info(0)["groups"]):
        gi = b.group_info(g)
        assert gi["cell_kernel"] in (0, 1, 2, 3) and gi["stage_cells"] >= 0
    for rep in range(3):
        R, K = b.assemble(to_dev(u), None, to_dev(kap), mode=-1)
        ref = b.assemble(to_dev(u), None, to_dev(kap), mode=best)
        assert torch.equal(R, ref[0]) and torch.equal(K, ref[1])


def test_staging_chunks_same_results(mixed):
    """Two-phase launches over chunks of meshes (bounded staging) reproduce the whole-group launch
    bit for bit: every mesh alone (1 byte), a few meshes per chunk, and the whole group."""
    meshes, _, _, u, kap = mixed
    b, sizes = build(meshes, N.local_damage_params())
    b.set_stage_bytes(0)
    ref = [t.cpu().numpy() for t in b.assemble(to_dev(u), None, to_dev(kap), mode=5)]
    for nbytes in (1, 200_000, 2_000_000, 0):
        b.set_stage_bytes(nbytes)
        chunks = {b.group_info(g)["stage_cells"] for g in range(b.group_info(0)["groups"])}
        assert (chunks == {0}) == (nbytes == 0)
        R, K = (t.cpu().numpy() for t in b.assemble(to_dev(u), None, to_dev(kap), mode=5))
        assert np.array_equal(R, ref[0]) and np.array_equal(K, ref[1]), nbytes


def test_group_limit_splits_types_same_results(mixed):
    """With a small group limit every element type spreads over several groups (super-meshes): each
    mesh's R and K slices are bit-identical to the one-group-per-type batch: fused (mode 0), per
    group (mode 3), two-phase (mode 5), and graph replay (mode -1)."""
    meshes, b, sizes, u, kap = mixed
    b2 = N.Batch()
    b2.set_integrand(N.local_damage_params())
    b2.set_group_limit(600)  # tet10: 245 + 343 nodes, then 245; hex20: 515, then 624
    for ct, dt, X, C, R, nn in meshes:
        b2.add_mesh(ct, dt, X, C, R, nn)
    sizes2 = b2.build()
    assert sizes2["R"] == sizes["R"] and sizes2["K"] == sizes["K"]
    groups = {b2.mesh_info(m)["group"] for m in range(len(meshes))}
    assert len(groups) > len(W.TYPES)
    u2, kap2 = np.empty_like(u), np.empty_like(kap)
    for m, (ct, dt, X, C, ranks, nn) in enumerate(meshes):  # the same per-mesh inputs in b2's layout
        i, j = b.mesh_info(m), b2.mesh_info(m)
        u2[j["u"]:j["u"] + 3 * nn] = u[i["u"]:i["u"] + 3 * nn]
        h = mesh_history(kap, i)
        c = np.arange(j["cells"])
        for ip in range(j["nip"]):
            kap2[j["history"] + c + ip * j["history_ld"]] = h[ip * j["cells"] + c]
    for mode in (0, 3, 5, -1):
        R, K = (t.cpu().numpy() for t in b.assemble(to_dev(u), None, to_dev(kap), mode=mode if mode >= 0 else 3))
        R2, K2 = (t.cpu().numpy() for t in b2.assemble(to_dev(u2), None, to_dev(kap2), mode=mode))
        for m in range(len(meshes)):
            i, j = b.mesh_info(m), b2.mesh_info(m)
            assert np.array_equal(R[i["R_u"]:i["R_u"] + i["rows"]], R2[j["R_u"]:j["R_u"] + j["rows"]]), (mode, m)
            assert np.array_equal(K[i["K_u"]:i["K_u"] + i["nnz_u"]], K2[j["K_u"]:j["K_u"] + j["nnz_u"]]), (mode, m)


@pytest.mark.parametrize("limit,groups,fan", [(0, 1, False), (64, 2, False), (30, 3, False), (30, 4, True),
                                              (0, 1, True)])
def test_gdm_batch_residual_adjoint_matches_single_meshes(limit, groups, fan):
    """Also with the meshes split over several groups (27, 36 and 60 dof nodes: a limit of 64 nodes
    packs the first two together, 30 gives each its own group), and with the fan mesh, whose axis
    node is beyond the node kernel's bounds (10 cells, 33 neighbours): in a group of its own it
    matches the single-mesh calls bit for bit; sharing one group with the bricks it makes the group
    non-affine (the bricks then run the node kernel's general form, which sums the IPs in another
    order than the affine form of their single-mesh calls), so there the match is to 1e-12 of the
    magnitudes."""
    dims = [(2, 2, 2), (3, 2, 2), (2, 3, 4)]
    meshes = []
    for n in dims:
        X, C = N.structured_mesh(N.HEX8, *n, tuple(float(v) for v in n))
        ranks, nn = N.dof_layer(N.HEX8, C, N.HEX8)
        meshes.append((N.HEX8, N.HEX8, X, C, ranks, nn))
    if fan:
        X, C = fan_mesh()
        ranks, nn = N.dof_layer(N.HEX8, C, N.HEX8)
        meshes.append((N.HEX8, N.HEX8, X, C, ranks, nn))
    exact = not (fan and groups == 1)

    def same(a, b):
        if exact:
            return np.array_equal(a, b)
        return np.allclose(a, b, rtol=1e-12, atol=1e-12 * np.abs(b).max())

    b = N.Batch()
    b.set_integrand(N.gradient_damage_params())
    b.set_group_limit(limit)
    for ct, dt, X, C, R, nn in meshes:
        b.add_mesh(ct, dt, X, C, R, nn)
    sizes = b.build()
    assert len({b.mesh_info(m)["group"] for m in range(len(meshes))}) == groups
    n_cells = sum(m[3].shape[0] for m in meshes)
    E, ft, gf = S.material_fields(n_cells)
    b.set_material_fields(E, ft, gf)
    u, e, kap = batch_inputs(sizes)
    r_obs = S.uniform(0x5151, sizes["R"], -1e-2, 1e-2)
    R, K, ds, dt = b.residual_adjoint(to_dev(u), to_dev(e), to_dev(kap), to_dev(r_obs))
    R2, K2, ds2, dt2 = b.residual_adjoint(to_dev(u), to_dev(e), to_dev(kap), to_dev(r_obs))  # replay
    assert torch.equal(R, R2) and torch.equal(ds, ds2) and torch.equal(dt, dt2)
    if fan:
        # per-mesh launches (mode 4: the node kernel per mesh, then the group's outside nodes)
        R4 = torch.full_like(R, float("nan"))
        K4 = torch.full_like(K, float("nan"))
        b.assemble(to_dev(u), to_dev(e), to_dev(kap), R4, K4, mode=4)
        for a, c in ((R4, R), (K4, K)):
            a, c = a.cpu().numpy(), c.cpu().numpy()
            assert np.allclose(a, c, rtol=1e-12, atol=1e-12 * np.abs(c).max())
    R, K, ds, dt = (t.cpu().numpy() for t in (R, K, ds, dt))
    c0 = 0
    for m, (ct, dtp, X, C, ranks, nn) in enumerate(meshes):
        info = b.mesh_info(m)
        nc = C.shape[0]
        ctx = N.Context()
        ctx.set_cells(ct, X, C)
        ctx.set_dof_type(0, dtp, 3, 3 * nn, N.cell_dofs(ranks, 3))
        ctx.set_dof_type(1, dtp, 1, nn, N.cell_dofs(ranks, 1))
        ctx.set_integrand(N.gradient_damage_params())
        ctx.set_rule(0, 2)
        ctx.build_pattern()
        ctx.set_material_fields(E[c0:c0 + nc], ft[c0:c0 + nc], gf[c0:c0 + nc])
        um = to_dev(u[info["u"]:info["u"] + 3 * nn])
        em = to_dev(e[info["e"]:info["e"] + nn])
        km = to_dev(mesh_history(kap, info))
        Rm, Km = ctx.build_vector_matrix(um, em, km)
        rows = np.r_[info["R_u"]:info["R_u"] + 3 * nn, info["R_e"]:info["R_e"] + nn]
        assert same(R[rows], Rm.cpu().numpy()), m
        kk = np.r_[info["K_u"]:info["K_u"] + info["nnz_u"], info["K_e"]:info["K_e"] + info["nnz_e"]]
        assert same(K[kk], Km.cpu().numpy()), m
        lam = to_dev(R[rows] - r_obs[rows])
        dsm, dtm = ctx.vjp(um, em, km, lam)
        assert same(ds[rows], dsm.cpu().numpy()), m
        th = np.concatenate([dt[info["theta"] + k * info["theta_stride"] + np.arange(nc)] for k in range(3)])
        assert same(th, dtm.cpu().numpy().ravel()), m
        c0 += nc
