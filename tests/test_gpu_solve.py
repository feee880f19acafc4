"""The constrained solve on the device (nuto_set_constraints, nuto_solve, nuto_solve_trial_state)
against the oracle's direct solve of the same system (NuTo::Solve / SolveTrialState, Solve.cpp:8-78,
with Constraints::BuildUnitConstraintMatrix, Constraints.cpp:89-137).

Both solve with the K that the device assembled (copied back; the pattern is bit-equal to the
oracle's): the comparison isolates the solve. The device solves iteratively to a relative reduced
residual of 1e-14; the bar is a relative error of 1e-9 against the direct solution (the systems
here have condition numbers below 1e4). The linear patch test is exact to 1e-10.
"""
import numpy as np
import pytest

import nuto_amd as N
import oracle as O
from pairs import Pair, make_mesh, to_dev
from test_gpu_parity import gdm_params, le_params

pytestmark = pytest.mark.gpu
TOL = 1e-14


def nodes_where(p, pred):
    """Dof-node ranks whose (first-touch) coordinate satisfies pred(x)."""
    dim = N.DIM[p.coord_type]
    xs = np.zeros((p.nn, dim))
    xs[p.ranks.ravel()] = p.X[p.C.ravel(), :dim]
    return np.nonzero(pred(xs))[0], xs


def rel_err(x, ref):
    return np.abs(x - ref).max() / np.abs(ref).max()


@pytest.mark.parametrize("method", [N.SOLVER_CG, N.SOLVER_BICGSTAB])
def test_linear_elastic_clamped_edge(method):
    X, C = make_mesh(N.QUAD4, (10, 6), (2.0, 1.0, 1.0))
    p = Pair(N.QUAD4, N.QUAD4, X, C, *le_params(2), two_fields=False)
    ctx = p.context()
    K = ctx.build_matrix(to_dev(np.zeros(p.n_u)))
    rp, ci = ctx.pattern()
    left, _ = nodes_where(p, lambda xs: np.abs(xs[:, 0]) < 1e-12)
    eqs = [(int(2 * r + c), []) for r in left for c in range(2)]
    ctx.set_constraints(0, eqs)
    f = np.random.default_rng(11).uniform(-1.0, 1.0, p.n_u)
    x, info = ctx.solve(K, to_dev(f), method=method, tol=TOL)
    x = x.cpu().numpy()
    ref = O.solve_constrained([p.n_u], [eqs], rp, ci, K.cpu().numpy(), f)
    assert info["independent_dofs"] == p.n_u - len(eqs)
    assert info["relative_residual"] <= TOL and info["iterations"] > 0
    assert rel_err(x, ref) <= 1e-9, rel_err(x, ref)
    assert np.all(x[[d for d, _ in eqs]] == 0.0)
    # bit-reproducible
    x2, _ = ctx.solve(K, to_dev(f), method=method, tol=TOL)
    assert np.array_equal(x2.cpu().numpy(), x)


@pytest.mark.parametrize("ct,n", [(N.QUAD4, (6, 4)), (N.HEX8, (3, 3, 2))])
def test_trial_state_linear_patch(ct, n):
    dim = N.DIM[ct]
    X, C = make_mesh(ct, n, (2.0, 1.0, 1.0))
    p = Pair(ct, ct, X, C, *le_params(dim), two_fields=False)
    ctx = p.context()
    u0 = to_dev(np.zeros(p.n_u))
    K = ctx.build_matrix(u0)
    R = ctx.build_vector(u0)
    bnd, xs = nodes_where(p, lambda xs: np.any((np.abs(xs - xs.min(0)) < 1e-12) | (np.abs(xs - xs.max(0)) < 1e-12),
                                               axis=1))
    A = np.array([[1.0, -0.5, 0.25], [0.75, 0.5, -1.0], [0.2, 0.1, 0.3]])[:dim, :dim] * 1e-3
    field = xs @ A.T
    eqs = [(int(dim * r + c), []) for r in bnd for c in range(dim)]
    rhs = np.array([field[r, c] for r in bnd for c in range(dim)])
    ctx.set_constraints(0, eqs)
    x, info = ctx.solve_trial_state(K, R, to_dev(np.zeros(len(rhs))), to_dev(rhs), method=N.SOLVER_CG, tol=TOL)
    expect = -field.ravel()  # dof = dim * rank + comp
    assert np.allclose(x.cpu().numpy(), expect, rtol=0, atol=1e-10 * np.abs(expect).max())


def test_gradient_damage_two_dof_types_bicgstab():
    X, C = make_mesh(N.HEX8, (3, 3, 3), (1.0, 1.0, 1.0), jitter=0.1)
    p = Pair(N.HEX8, N.HEX8, X, C, *gdm_params(), two_fields=True)
    ctx = p.context()
    # a mildly loaded state (strains near kappa0): K non-symmetric (K_eu) but well conditioned
    u, e, kap = p.inputs(u_range=2e-5, e_range=1.2e-4, k_range=5e-5)
    K = ctx.build_matrix(to_dev(u), to_dev(e), to_dev(kap))
    R = ctx.build_vector(to_dev(u), to_dev(e), to_dev(kap))
    rp, ci = ctx.pattern()
    X0, _ = make_mesh(N.HEX8, (3, 3, 3), (1.0, 1.0, 1.0))
    xs0 = np.zeros((p.nn, 3))
    xs0[p.ranks.ravel()] = X0[p.C.ravel()]
    bottom = np.nonzero(np.abs(xs0[:, 2]) < 1e-12)[0]
    eq_u = [(int(3 * r + c), []) for r in bottom for c in range(3)]
    top = np.nonzero(np.abs(xs0[:, 2] - 1.0) < 1e-12)[0]
    eq_u += [(int(3 * r + 2), []) for r in top]  # prescribed vertical displacement (rhs below)
    eq_e = [(int(bottom[0]), [])]
    ctx.set_constraints(0, eq_u)
    ctx.set_constraints(1, eq_e)
    rhs_new = np.zeros(len(eq_u) + len(eq_e))
    rhs_new[len(bottom) * 3:len(eq_u)] = 1e-3
    x, info = ctx.solve_trial_state(K, R, None, to_dev(rhs_new), method=N.SOLVER_BICGSTAB, tol=1e-13)
    ref = O.solve_constrained([p.n_u, p.n_e], [eq_u, eq_e], rp, ci, K.cpu().numpy(), R.cpu().numpy(),
                              rhs_old=np.zeros(len(rhs_new)), rhs_new=rhs_new)
    assert info["relative_residual"] <= 1e-13
    assert rel_err(x.cpu().numpy(), ref) <= 1e-8, rel_err(x.cpu().numpy(), ref)


def test_coupled_equations():
    X, C = make_mesh(N.QUAD4, (8, 4), (2.0, 1.0, 1.0))
    p = Pair(N.QUAD4, N.QUAD4, X, C, *le_params(2), two_fields=False)
    ctx = p.context()
    K = ctx.build_matrix(to_dev(np.zeros(p.n_u)))
    rp, ci = ctx.pattern()
    left, _ = nodes_where(p, lambda xs: np.abs(xs[:, 0]) < 1e-12)
    right, _ = nodes_where(p, lambda xs: np.abs(xs[:, 0] - 2.0) < 1e-12)
    master = int(2 * right[0])
    eqs = [(int(2 * r + c), []) for r in left for c in range(2)]
    eqs += [(int(2 * r), [(master, -0.8)]) for r in right[1:]]
    ctx.set_constraints(0, eqs)
    f = np.random.default_rng(5).uniform(-1.0, 1.0, p.n_u)
    x, _ = ctx.solve(K, to_dev(f), method=N.SOLVER_CG, tol=TOL)
    ref = O.solve_constrained([p.n_u], [eqs], rp, ci, K.cpu().numpy(), f)
    assert rel_err(x.cpu().numpy(), ref) <= 1e-9


def test_constraint_errors():
    X, C = make_mesh(N.QUAD4, (3, 2), (1.0, 1.0, 1.0))
    p = Pair(N.QUAD4, N.QUAD4, X, C, *le_params(2), two_fields=False)
    ctx = p.context()
    with pytest.raises(N.NutoError):  # ConstraintDoubleConstrained
        ctx.set_constraints(0, [(0, []), (0, [])])
    with pytest.raises(N.NutoError):  # ConstraintTwoDependentDofsInOneEquation
        ctx.set_constraints(0, [(0, []), (1, [(0, 0.4)])])
    with pytest.raises(N.NutoError):  # dof number beyond the dof type
        ctx.set_constraints(0, [(p.n_u, [])])
    with pytest.raises(N.NutoError):  # not numbered
        ctx.set_constraints(0, [(-1, [])])
    ctx.set_constraints(0, [(0, []), (1, [(2, 0.4)])])  # valid
