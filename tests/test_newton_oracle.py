"""The oracle's load step (QuasistaticSolver::DoStep: trial state, Newton-Raphson with line search,
history update; QuasistaticSolver.cpp:116-147, NewtonRaphson.h:62-94, LineSearch.h:39-63), pinned
by closed forms.

Homogeneous gradient-damage patch: every boundary node's displacement follows u = A x and the
nonlocal strains are free. The exact solution is the linear displacement field everywhere, the
uniform nonlocal strain e = eeq(A) (then R_e = N (e - eeq) + c B^T grad e vanishes at every IP)
and kappa = max(kappa_old, e); with A beyond kappa0 the step damages the whole block and needs
several Newton iterations. Linear elasticity: the trial state is already the solution (no Newton
iteration). A step that cannot converge raises NoConvergence.
"""
import numpy as np
import pytest

import nuto_amd as N
import oracle as O
from pairs import Pair, make_mesh

GDM = O.Params(kind=1, dim=3, E=30000.0, nu=0.2, k=10.0, kappa0=1e-4, beta=100.0, alpha=0.99, c=4.0)
A = np.array([[2.0, 0.5, 0.0], [0.5, -0.4, 0.3], [0.0, 0.3, -0.6]]) * 1e-4  # symmetric: no rotation


def boundary_setup(p, xs):
    dim = xs.shape[1]
    lo, hi = xs.min(0), xs.max(0)
    bnd = np.nonzero(np.any((np.abs(xs - lo) < 1e-12) | (np.abs(xs - hi) < 1e-12), axis=1))[0]
    eqs = [(int(dim * r + c), []) for r in bnd for c in range(dim)]
    field = xs @ A[:dim, :dim].T
    rhs = np.array([field[r, c] for r in bnd for c in range(dim)])
    return eqs, rhs, field


def node_coords(p):
    dim = N.DIM[p.coord_type]
    xs = np.zeros((p.nn, dim))
    xs[p.ranks.ravel()] = p.X[p.C.ravel(), :dim]
    return xs


def eeq_closed_form(eps, nu=0.2, k=10.0):
    """Modified Mises norm of a strain tensor: K1 I1 + sqrt(K1^2 I1^2 + K2 J2)
    (ModifiedMisesStrainNorm.h:70-84), J2 = dev(eps) : dev(eps) / 2."""
    K1 = (k - 1.0) / (2.0 * k * (1.0 - 2.0 * nu))
    K2 = 3.0 / (k * (1.0 + nu) ** 2)
    I1 = np.trace(eps)
    dev = eps - I1 / 3.0 * np.eye(3)
    J2 = 0.5 * np.sum(dev * dev)
    return K1 * I1 + np.sqrt(K1 ** 2 * I1 ** 2 + K2 * J2)


def test_gdm_homogeneous_patch_closed_form():
    X, C = make_mesh(N.HEX8, (3, 2, 2), (3.0, 2.0, 2.0))
    p = Pair(N.HEX8, N.HEX8, X, C, N.gradient_damage_params(), GDM, True)
    xs = node_coords(p)
    eqs, rhs, field = boundary_setup(p, xs)
    x = np.zeros(p.n_u + p.n_e)
    kappa = np.zeros(p.ncells * p.nip)
    xn, kn, it = O.quasistatic_step(p.orc, [p.n_u, p.n_e], [eqs, []], x, kappa, np.zeros(len(rhs)), rhs)
    eeq = eeq_closed_form(A)
    assert eeq > 1e-4 and it >= 1  # damaging, nonlinear: the trial state alone does not converge
    np.testing.assert_allclose(xn[:p.n_u], field.ravel(), rtol=0, atol=1e-12 * np.abs(field).max())
    np.testing.assert_allclose(xn[p.n_u:], eeq, rtol=1e-9)
    np.testing.assert_allclose(kn, eeq, rtol=1e-9)


def test_gdm_history_kept_when_larger():
    X, C = make_mesh(N.HEX8, (2, 2, 2), (2.0, 2.0, 2.0))
    p = Pair(N.HEX8, N.HEX8, X, C, N.gradient_damage_params(), GDM, True)
    eqs, rhs, field = boundary_setup(p, node_coords(p))
    kappa = np.full(p.ncells * p.nip, 5e-4)  # beyond the step's eeq: the history stays
    xn, kn, _ = O.quasistatic_step(p.orc, [p.n_u, p.n_e], [eqs, []], np.zeros(p.n_u + p.n_e), kappa,
                                   np.zeros(len(rhs)), rhs)
    assert np.array_equal(kn, kappa)
    eeq = eeq_closed_form(A)
    np.testing.assert_allclose(xn[p.n_u:], eeq, rtol=1e-9)


def test_linear_elastic_step_is_the_linear_solve():
    X, C = make_mesh(N.QUAD4, (5, 3), (2.0, 1.0, 1.0))
    orc = O.Params(kind=0, dim=2, E=30000.0, nu=0.2, plane_state=0)
    p = Pair(N.QUAD4, N.QUAD4, X, C, N.linear_elastic_params(2, 30000.0, 0.2), orc, False)
    eqs, rhs, field = boundary_setup(p, node_coords(p))
    xn, kn, it = O.quasistatic_step(p.orc, [p.n_u], [eqs], np.zeros(p.n_u), None, np.zeros(len(rhs)), rhs)
    assert it == 0 and kn is None
    np.testing.assert_allclose(xn, field.ravel(), rtol=0, atol=1e-12 * np.abs(field).max())


def test_no_convergence():
    X, C = make_mesh(N.HEX8, (2, 1, 1), (2.0, 1.0, 1.0))
    p = Pair(N.HEX8, N.HEX8, X, C, N.gradient_damage_params(), GDM, True)
    eqs, rhs, _ = boundary_setup(p, node_coords(p))
    with pytest.raises(O.NoConvergence):
        O.quasistatic_step(p.orc, [p.n_u, p.n_e], [eqs, []], np.zeros(p.n_u + p.n_e), np.zeros(p.ncells * p.nip),
                           np.zeros(len(rhs)), 20 * rhs, tol=1e-30, max_iter=1)


def test_adaptive_solve_time_stepping():
    """AdaptiveSolve::Solve (AdaptiveSolve.cpp:18-67) with a step that fails above dt = 0.3 and
    reports 1 iteration: dt shrinks by 0.5 on failure, grows by 1.5 (to dt_max) after easy steps,
    and the last step is cut to end at t_end."""
    calls, post = [], []

    def do_step(t):
        dt = t - (post[-1] if post else 0.0)
        calls.append(round(t, 12))
        if dt > 0.3 + 1e-12:
            raise N.NutoError(N.NO_CONVERGENCE, "No convergence after 6 iterations.")
        return 1

    times = N.adaptive_solve(do_step, 1.0, post_process=post.append, dt=0.4, dt_max=1.0)
    # 0.4 fails -> 0.2 ok (t = 0.2), dt 0.3 ok (0.5), dt 0.45 fails -> 0.225 ok (0.725), dt 0.3375 is
    # cut to 0.275 before it is tried and ends at 1.0
    assert np.allclose(times, [0.2, 0.5, 0.725, 1.0])
    assert np.allclose(calls, [0.4, 0.2, 0.5, 0.95, 0.725, 1.0])
    assert post[0] == 0.0 and np.allclose(post[1:], times)
    with pytest.raises(N.NutoError, match="minimal time step"):
        N.adaptive_solve(lambda t: (_ for _ in ()).throw(N.NutoError(N.NO_CONVERGENCE, "x")), 1.0, dt_min=0.01)
