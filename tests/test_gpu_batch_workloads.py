"""The batched forms at their workload sizes.

Batched residual + adjoint (hex8 gradient damage, element counts log-uniform in [1e4, 2e5], per-cell
material fields E ~ LN(30000, 0.1), f_t ~ LN(3, 0.1), G_f ~ LN(0.1, 0.15), r_obs = R at a seeded
perturbation of the fields): the first 8 meshes of the benchmark's 256-mesh batch, in several groups.
Per mesh:
* nnz = 16 prod_d (3 n_d + 1) (structured closed form);
* d_state = K^T (R - r_obs) formed from the device's own K, within 1e-12 of |K|^T |lambda|;
* rigid translations in the null space of K (every row, 1e-12 of the absolute row sums);
two meshes bit-identical to single-mesh calls (R, K, d_state, d_theta); the smallest mesh
(1.0e4 cells) against the oracle (pattern bit-equal, R and K within the parity bound), and its
d_theta against central differences of the device's R (relative step 1e-6).

Mixed-type batch (local isotropic damage, jittered grids): one mesh per element pair with 1e3-5e4
cells, each against the oracle, in the tuned (graph-replayed) mode and the two-phase mode.
"""
import numpy as np
import pytest
import torch

import nuto_amd as N
import oracle as O
from nuto_amd import synthetic as S
from nuto_amd import workloads as W
from pairs import check_values, to_dev

pytestmark = pytest.mark.gpu


def mesh_history(kap, info):
    c = np.arange(info["cells"])
    return np.concatenate([kap[info["history"] + c + i * info["history_ld"]] for i in range(info["nip"])])


def grid_dims(X):
    return [len(np.unique(np.round(X[:, d], 9))) - 1 for d in range(3)]


@pytest.fixture(scope="module")
def gdm_batch():
    meshes = W.gdm_batch(8)
    cells = [m[3].shape[0] for m in meshes]
    assert min(cells) >= 0.9e4 and max(cells) <= 2.2e5
    b = N.Batch()
    b.set_integrand(N.gradient_damage_params())
    b.set_group_limit(2 * max(m[5] for m in meshes))  # several groups
    for ct, dt, X, C, R, nn in meshes:
        b.add_mesh(ct, dt, X, C, R, nn)
    sizes = b.build()
    assert len({b.mesh_info(m)["group"] for m in range(len(meshes))}) > 1
    n_cells = sum(cells)
    E, ft, gf = S.material_fields(n_cells)
    u = S.uniform(S.SEED_U, sizes["u"], -5e-3, 5e-3)
    e = S.uniform(S.SEED_E, sizes["e"], 0.0, 5e-4)
    kap = S.uniform(S.SEED_KAPPA, sizes["history"], 0.0, 3e-4)
    du, de, dk = to_dev(u), to_dev(e), to_dev(kap)
    b.set_material_fields(*S.perturbed_fields(E, ft, gf))
    r_obs = b.assemble(du, de, dk)[0].clone()
    b.set_material_fields(E, ft, gf)
    R, K, ds, dth = b.residual_adjoint(du, de, dk, r_obs)
    torch.cuda.synchronize()
    return dict(meshes=meshes, b=b, fields=(E, ft, gf), u=u, e=e, kap=kap, r_obs=r_obs, R=R, K=K, ds=ds, dth=dth)


def _mesh_csr(B, m):
    info = B["b"].mesh_info(m)
    rp, ci = B["b"].pattern(m)
    K = B["K"]
    vals = torch.cat([K[info["K_u"]:info["K_u"] + info["nnz_u"]], K[info["K_e"]:info["K_e"] + info["nnz_e"]]])
    rows = torch.repeat_interleave(torch.arange(len(rp) - 1, device="cuda"), torch.from_numpy(np.diff(rp).astype(np.int64)).cuda())
    nn = B["meshes"][m][5]
    rix = torch.from_numpy(np.r_[info["R_u"]:info["R_u"] + 3 * nn, info["R_e"]:info["R_e"] + nn]).cuda()
    return info, rp, torch.from_numpy(ci).cuda().long(), vals, rows, rix


@pytest.mark.parametrize("m", range(8))
def test_gdm_batch_mesh_properties(gdm_batch, m):
    B = gdm_batch
    ct, dt, X, C, ranks, nn = B["meshes"][m]
    info, rp, ci, vals, rows, rix = _mesh_csr(B, m)
    dims = grid_dims(X)
    assert info["nnz_u"] + info["nnz_e"] == 16 * np.prod([3 * n + 1 for n in dims])
    assert info["nnz_u"] == 12 * np.prod([3 * n + 1 for n in dims])
    # adjoint: d_state = K^T lambda, lambda = R - r_obs on the mesh's rows
    lam = (B["R"] - B["r_obs"])[rix]
    kt = torch.zeros(4 * nn, dtype=torch.float64, device="cuda").index_add_(0, ci, vals * lam[rows])
    mag = torch.zeros(4 * nn, dtype=torch.float64, device="cuda").index_add_(0, ci, (vals * lam[rows]).abs())
    err = (B["ds"][rix] - kt).abs()
    assert bool((err <= 1e-12 * mag + 1e-300).all()), float((err / (mag + 1e-300)).max())
    # rigid translations: K t = 0 row by row
    for d in range(3):
        t = torch.zeros(4 * nn, dtype=torch.float64, device="cuda")
        t[d:3 * nn:3] = 1.0
        prod = vals * t[ci]
        y = torch.zeros(4 * nn, dtype=torch.float64, device="cuda").index_add_(0, rows, prod)
        a = torch.zeros(4 * nn, dtype=torch.float64, device="cuda").index_add_(0, rows, prod.abs())
        assert bool((y.abs() <= 1e-12 * a + 1e-300).all()), (m, d)


def _single(B, m):
    ct, dt, X, C, ranks, nn = B["meshes"][m]
    c0 = sum(B["meshes"][k][3].shape[0] for k in range(m))
    nc = C.shape[0]
    E, ft, gf = (f[c0:c0 + nc] for f in B["fields"])
    ctx = N.Context()
    ctx.set_cells(ct, X, C)
    ctx.set_dof_type(0, dt, 3, 3 * nn, N.cell_dofs(ranks, 3))
    ctx.set_dof_type(1, dt, 1, nn, N.cell_dofs(ranks, 1))
    ctx.set_integrand(N.gradient_damage_params())
    ctx.set_rule(0, 2)
    ctx.build_pattern()
    ctx.set_material_fields(E, ft, gf)
    info = B["b"].mesh_info(m)
    um = B["u"][info["u"]:info["u"] + 3 * nn]
    em = B["e"][info["e"]:info["e"] + nn]
    km = mesh_history(B["kap"], info)
    return ctx, info, (E, ft, gf), um, em, km


@pytest.mark.parametrize("m", [0, 1])
def test_gdm_batch_bit_identical_to_single_mesh(gdm_batch, m):
    B = gdm_batch
    ctx, info, _, um, em, km = _single(B, m)
    nn = B["meshes"][m][5]
    nc = B["meshes"][m][3].shape[0]
    Rm, Km = ctx.build_vector_matrix(to_dev(um), to_dev(em), to_dev(km))
    rix = np.r_[info["R_u"]:info["R_u"] + 3 * nn, info["R_e"]:info["R_e"] + nn]
    kk = np.r_[info["K_u"]:info["K_u"] + info["nnz_u"], info["K_e"]:info["K_e"] + info["nnz_e"]]
    assert torch.equal(B["R"][torch.from_numpy(rix).cuda()], Rm)
    assert torch.equal(B["K"][torch.from_numpy(kk).cuda()], Km)
    lam = Rm - B["r_obs"][torch.from_numpy(rix).cuda()]
    dsm, dtm = ctx.vjp(to_dev(um), to_dev(em), to_dev(km), lam)
    assert torch.equal(B["ds"][torch.from_numpy(rix).cuda()], dsm)
    th = torch.cat([B["dth"][info["theta"] + k * info["theta_stride"]:][:nc] for k in range(3)])
    assert torch.equal(th, dtm.reshape(-1))


def test_gdm_batch_smallest_mesh_against_oracle_and_fd(gdm_batch):
    B = gdm_batch
    m = int(np.argmin([mm[3].shape[0] for mm in B["meshes"]]))
    ct, dt, X, C, ranks, nn = B["meshes"][m]
    ctx, info, (E, ft, gf), um, em, km = _single(B, m)
    om = O.Mesh.from_arrays(O.HEX8, X, C)
    Ld, Le = om.add_dof(O.HEX8, 3), om.add_dof(O.HEX8, 1)
    p = O.Params(kind=1, dim=3, E=30000.0, nu=0.2, k=10.0, kappa0=1e-4, beta=100.0, alpha=0.99, c=4.0)
    P = O.Problem(om, p, 0, 2, Ld, Le)
    P.set_theta(E, ft, gf)
    ref = P.assemble(um, em, km, nthreads=8)
    rp, ci = B["b"].pattern(m)
    assert np.array_equal(rp, ref["row_ptr"]) and np.array_equal(ci, ref["col"])
    rix = np.r_[info["R_u"]:info["R_u"] + 3 * nn, info["R_e"]:info["R_e"] + nn]
    kk = np.r_[info["K_u"]:info["K_u"] + info["nnz_u"], info["K_e"]:info["K_e"] + info["nnz_e"]]
    R = B["R"].cpu().numpy()
    check_values(R[rix], ref["R"], ref["R_bound"], "R")
    check_values(B["K"][torch.from_numpy(kk).cuda()].cpu().numpy(), ref["val"], ref["bound"], "K")
    # d_theta against central differences of the device's R
    lam = torch.from_numpy(R[rix] - B["r_obs"].cpu().numpy()[rix]).cuda()
    nc = C.shape[0]
    dth = np.stack([B["dth"][info["theta"] + k * info["theta_stride"]:][:nc].cpu().numpy() for k in range(3)])
    du, de, dk = to_dev(um), to_dev(em), to_dev(km)
    h = 1e-6
    rng = np.random.default_rng(5)
    for c in rng.choice(nc, 3, replace=False):
        for k in range(3):
            f = [E.copy(), ft.copy(), gf.copy()]
            step = h * f[k][c]
            f[k][c] += step
            ctx.set_material_fields(*f)
            Rp = ctx.build_vector(du, de, dk)
            f[k][c] -= 2 * step
            ctx.set_material_fields(*f)
            Rm = ctx.build_vector(du, de, dk)
            fd = float(lam @ (Rp - Rm)) / (2 * step)
            assert dth[k, c] == pytest.approx(fd, rel=1e-5, abs=1e-9 * abs(float(lam.abs().max()))), (c, k)
    ctx.set_material_fields(E, ft, gf)


RULE = {(N.TET4, N.TET4): (1, 1), (N.TET4, N.TET10): (1, 2), (N.HEX8, N.HEX8): (0, 2), (N.HEX8, N.HEX20): (0, 3)}
MIXED_CELLS = {(N.TET4, N.TET4): 4.0e4, (N.TET4, N.TET10): 3.0e3, (N.HEX8, N.HEX8): 1.2e4, (N.HEX8, N.HEX20): 1.0e3}


@pytest.fixture(scope="module")
def mixed_batch():
    meshes = []
    for i, (ct, dt) in enumerate(W.TYPES):
        meshes.append((ct, dt) + W.mesh(ct, dt, MIXED_CELLS[(ct, dt)], 0.1, W.SEED_JITTER + 100 + i))
    cells = [m[3].shape[0] for m in meshes]
    assert min(cells) >= 1e3 and max(cells) <= 5e4, cells
    b = N.Batch()
    b.set_integrand(N.local_damage_params())
    for ct, dt, X, C, R, nn in meshes:
        b.add_mesh(ct, dt, X, C, R, nn)
    sizes = b.build()
    u = S.uniform(S.SEED_U, sizes["u"], -5e-3, 5e-3)
    kap = S.uniform(S.SEED_KAPPA, sizes["history"], 0.0, 3e-4)
    return meshes, b, u, kap


@pytest.mark.parametrize("mode", [-1, 5])
def test_mixed_batch_in_range_against_oracle(mixed_batch, mode):
    meshes, b, u, kap = mixed_batch
    if mode == -1:
        b.autotune(to_dev(u), None, to_dev(kap))
    R, K = (t.cpu().numpy() for t in b.assemble(to_dev(u), None, to_dev(kap), mode=mode))
    p = O.Params(kind=2, dim=3, E=30000.0, nu=0.2, k=10.0, kappa0=1e-4, beta=100.0, alpha=0.99)
    for m, (ct, dt, X, C, ranks, nn) in enumerate(meshes):
        info = b.mesh_info(m)
        om = O.Mesh.from_arrays(ct, X, C)
        Ld = om.add_dof(dt, 3)
        P = O.Problem(om, p, RULE[(ct, dt)][0], RULE[(ct, dt)][1], Ld, -1)
        ref = P.assemble(u[info["u"]:info["u"] + 3 * nn], None, mesh_history(kap, info), nthreads=8)
        rp, ci = b.pattern(m)
        assert np.array_equal(rp, ref["row_ptr"]) and np.array_equal(ci, ref["col"]), m
        check_values(R[info["R_u"]:info["R_u"] + info["rows"]], ref["R"], ref["R_bound"], f"R of mesh {m}")
        check_values(K[info["K_u"]:info["K_u"] + info["nnz_u"]], ref["val"], ref["bound"], f"K of mesh {m}")
