"""Pins of the oracle's constraint matrix and constrained solve (Constraints::BuildUnitConstraintMatrix,
Constraints.cpp:89-137; NuTo::Solve / SolveTrialState, Solve.cpp:8-78):

* the reference's own known-answer test ConstraintCMatrixInteracting
  (test/mechanics/constraints/Constraints.cpp:52-98): the 6 x 4 matrix with the -42 coupling entry;
* no equations give the identity, a dependent dof out of range is rejected;
* the linear patch test through SolveTrialState: every boundary dof prescribed on a linear field
  u = A x, zero load and zero start state: the returned negative increment is -A x at every node
  (linear elements reproduce a linear field exactly);
* a solve with coupled equations satisfies its equations and has a zero reduced residual
  C^T (f - K x).
"""
import numpy as np
import pytest

import nuto_amd as N
import oracle as O

from pairs import make_mesh


def test_unit_constraint_matrix_reference_kat():
    C = O.unit_constraint_matrix(6, [(1, []), (4, [(2, 42.0)])])
    E = np.zeros((6, 4))
    E[0, 0] = 1
    E[2, 1] = 1
    E[3, 2] = 1
    E[4, 1] = -42
    E[5, 3] = 1
    assert np.array_equal(C, E)


def test_unit_constraint_matrix_identity_and_range():
    assert np.array_equal(O.unit_constraint_matrix(5, []), np.eye(5))
    with pytest.raises(ValueError):
        O.unit_constraint_matrix(4, [(4, [])])
    # a coefficient below 1e-18 is dropped (Constraints.cpp:128)
    C = O.unit_constraint_matrix(3, [(0, [(1, 1e-19), (2, 0.5)])])
    assert C[0, 0] == 0.0 and C[0, 1] == -0.5


def linear_elastic_problem(coord_type, n):
    X, C = make_mesh(coord_type, n, (2.0, 1.0, 1.0), jitter=0.1)
    X0, _ = make_mesh(coord_type, n, (2.0, 1.0, 1.0))  # the same lattice unjittered (boundary flags)
    dim = N.DIM[coord_type]
    om = O.Mesh.from_arrays(coord_type, X, C)
    L = om.add_dof(coord_type, dim)
    P = O.Problem(om, O.Params(kind=0, dim=dim, E=30000.0, nu=0.2, plane_state=0), 0, 2, L)
    lay = om.dof_layer(L)
    # dof node coordinates (linear elements: the coordinate nodes), numbers[node, comp]
    xs = np.zeros((lay["numbers"].shape[0], dim))
    xs[lay["conn"].ravel()] = X[np.asarray(C).ravel(), :dim]
    x0 = np.zeros_like(xs)
    x0[lay["conn"].ravel()] = X0[np.asarray(C).ravel(), :dim]
    lo, hi = x0.min(axis=0), x0.max(axis=0)
    boundary = np.any((np.abs(x0 - lo) < 1e-12) | (np.abs(x0 - hi) < 1e-12), axis=1)
    return X, P, lay["numbers"], xs, boundary


@pytest.mark.parametrize("ct,n", [(N.QUAD4, (4, 3)), (N.HEX8, (3, 2, 2))])
def test_trial_state_linear_patch(ct, n):
    X, P, numbers, xs, boundary = linear_elastic_problem(ct, n)
    dim = N.DIM[ct]
    out = P.assemble(np.zeros(P.ndof[0]))
    A = np.array([[1.0, -0.5, 0.25], [0.75, 0.5, -1.0], [0.2, 0.1, 0.3]])[:dim, :dim] * 1e-3
    field = xs @ A.T  # u(x) = A x per node
    eqs, rhs = [], []
    for node in np.nonzero(boundary)[0]:
        for c in range(dim):
            eqs.append((int(numbers[node, c]), []))
            rhs.append(field[node, c])
    x = O.solve_constrained([P.ndof[0]], [eqs], out["row_ptr"], out["col"], out["val"], out["R"],
                            rhs_old=np.zeros(len(rhs)), rhs_new=np.array(rhs))
    expect = np.zeros(P.ndof[0])
    for node in range(xs.shape[0]):
        for c in range(dim):
            expect[numbers[node, c]] = -field[node, c]
    assert np.allclose(x, expect, rtol=0, atol=1e-12 * np.abs(expect).max())


def test_coupled_equations_hold_and_reduced_residual_vanishes():
    X, P, numbers, xs, _ = linear_elastic_problem(N.QUAD4, (5, 3))
    out = P.assemble(np.zeros(P.ndof[0]))
    left = np.nonzero(np.abs(xs[:, 0] - xs[:, 0].min()) < 1e-12)[0]
    right = np.nonzero(np.abs(xs[:, 0] - xs[:, 0].max()) < 1e-12)[0]
    master = int(numbers[right[0], 0])
    eqs = [(int(numbers[nd, c]), []) for nd in left for c in range(2)]  # clamped left edge
    # the other right-edge x dofs follow the master with a factor: 1 x_d - 0.8 x_master = 0
    eqs += [(int(numbers[nd, 0]), [(master, -0.8)]) for nd in right[1:]]
    f = np.random.default_rng(3).uniform(-1.0, 1.0, P.ndof[0])
    x = O.solve_constrained([P.ndof[0]], [eqs], out["row_ptr"], out["col"], out["val"], f)
    for d, terms in eqs:
        assert x[d] + sum(c * x[t] for t, c in terms) == pytest.approx(0.0, abs=1e-15)
    C = O.unit_constraint_matrix(P.ndof[0], eqs)
    K = O.csr_to_dense(out["row_ptr"], out["col"], out["val"], P.ndof[0])
    red = C.T @ (f - K @ x)
    assert np.abs(red).max() <= 1e-10 * np.abs(C.T @ f).max()
