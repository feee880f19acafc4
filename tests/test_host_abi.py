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


# ---- the binding refuses tensors the C ABI would overrun (checked before any library call) ------
def _ctx_without_device(n_u=30, n_e=10, n_cells=5, n_ip=8, nnz=100):
    """A Context whose sizes are set as after set_cells / set_dof_type / set_rule / build_pattern,
    with no device handle: every refusal below must happen before the C ABI is reached."""
    ctx = object.__new__(N.Context)
    ctx._init_sizes()
    ctx.h = None
    ctx._n_dofs = {0: n_u, 1: n_e}
    ctx.n_rows, ctx.nnz, ctx.n_cells, ctx.n_ip = n_u + n_e, nnz, n_cells, n_ip
    ctx.n_dofs0 = n_u
    ctx.dim = 3
    return ctx


def test_binding_refuses_wrong_dtype_and_short_tensors():
    import torch

    ctx = _ctx_without_device()
    f64 = lambda n: torch.zeros(n, dtype=torch.float64)  # noqa: E731
    u, e, kap, R, K = f64(30), f64(10), f64(40), f64(40), f64(100)
    cases = [
        (lambda: ctx.build_vector(u.float(), e, kap, out=R), "expected float64"),
        (lambda: ctx.build_vector(f64(29), e, kap, out=R), "needs 30"),
        (lambda: ctx.build_vector(u, f64(9), kap, out=R), "needs 10"),
        (lambda: ctx.build_vector(u, e, f64(39), out=R), "needs 40"),
        (lambda: ctx.build_vector(u, e, kap, out=f64(39)), "needs 40"),
        (lambda: ctx.build_matrix(u, e, kap, out=f64(99)), "needs 100"),
        (lambda: ctx.build_vector_matrix(u, e, kap, R=R, K=K.float()), "expected float64"),
        (lambda: ctx.vjp(u, e, kap, f64(39)), "needs 40"),
        (lambda: ctx.update_history(u, e, f64(7)), "needs 40"),
        (lambda: ctx.build_lumped_mass(1.0, out=f64(29)), "needs 30"),
        (lambda: ctx.solve(f64(99), f64(40), out=f64(40)), "needs 100"),
        (lambda: ctx.eval_ip(0, u, e, kap, out=f64(5 * 8 * 6 - 1)), "needs 240"),
        (lambda: ctx.quasistatic_step(f64(39), kap), "needs 40"),
    ]
    for call, msg in cases:
        with pytest.raises(ValueError, match=msg):
            call()
    # a wider history (set_history_ld) needs a longer kappa
    ctx._history_ld = 7
    with pytest.raises(ValueError, match="needs 56"):
        ctx.build_vector(u, e, kap, out=R)
    ctx._history_ld = None
    # correct dtypes and lengths get as far as the device check
    with pytest.raises(ValueError, match="device tensor"):
        ctx.build_vector_matrix(u, e, kap, R=R, K=K)


def test_batch_binding_refuses_short_tensors():
    import torch

    b = object.__new__(N.Batch)
    b.h = None
    b.sizes = dict(u=30, e=10, R=40, K=100, history=16, theta=6)
    f64 = lambda n: torch.zeros(n, dtype=torch.float64)  # noqa: E731
    u, e, kap = f64(30), f64(10), f64(16)
    with pytest.raises(ValueError, match="needs 40"):
        b.residual_adjoint(u, e, kap, f64(39), R=f64(40), K=f64(100), d_state=f64(40), d_theta=f64(6))
    with pytest.raises(ValueError, match="needs 6"):
        b.residual_adjoint(u, e, kap, f64(40), R=f64(40), K=f64(100), d_state=f64(40), d_theta=f64(5))
    with pytest.raises(ValueError, match="needs 16"):
        b.assemble(u, e, f64(15), R=f64(40), K=f64(100))
    b.sizes = None
    with pytest.raises(ValueError, match="build"):
        b.assemble(u, e, kap, R=f64(40), K=f64(100))


def test_cpp_mirror_containers():
    """The C++ mirror's DofMatrixSparse blocks and ToEigen on the reference's container cases
    (test/mechanics/dofs/DofMatrixSparseConvertEigen.cpp:4-49): host code, run without a device."""
    import subprocess

    exe = os.path.join(ROOT, "tests", "cpp", "test_host_shim")
    if not os.path.exists(exe):
        pytest.skip("tests/cpp not built (__graft_entry__.build())")
    r = subprocess.run([exe, "containers"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), r.stdout + r.stderr
