"""The load step on the device (nuto_quasistatic_step: QuasistaticSolver::DoStep with
NewtonRaphson::Solve and LineSearch, QuasistaticSolver.cpp:116-147, NewtonRaphson.h:62-94) against
the closed-form homogeneous patch and against the oracle's restatement of the same step
(oracle.quasistatic_step, direct linear solves) on a localising gradient-damage bar.

The device solves every linear system iteratively (BiCGSTAB, relative residual 1e-14); Newton then
reaches the same state as the oracle's direct solves up to that accuracy: the bar is 1e-8 relative
on x and kappa, and the same number of Newton iterations.
"""
import numpy as np
import pytest
import torch

import nuto_amd as N
import oracle as O
from pairs import Pair, make_mesh, to_dev
from test_newton_oracle import A, GDM, boundary_setup, eeq_closed_form, node_coords

pytestmark = pytest.mark.gpu
LIN = dict(method=N.SOLVER_BICGSTAB, linear_tol=1e-14)


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def test_gdm_homogeneous_patch_on_device():
    X, C = make_mesh(N.HEX8, (3, 2, 2), (3.0, 2.0, 2.0))
    p = Pair(N.HEX8, N.HEX8, X, C, N.gradient_damage_params(), GDM, True)
    eqs, rhs, field = boundary_setup(p, node_coords(p))
    ctx = p.context()
    ctx.set_constraints(0, eqs)
    x = to_dev(np.zeros(p.n_u + p.n_e))
    kappa = to_dev(np.zeros(p.ncells * p.nip))
    info = ctx.quasistatic_step(x, kappa, to_dev(np.zeros(len(rhs))), to_dev(rhs), **LIN)
    _, _, it = O.quasistatic_step(p.orc, [p.n_u, p.n_e], [eqs, []], np.zeros(p.n_u + p.n_e),
                                  np.zeros(p.ncells * p.nip), np.zeros(len(rhs)), rhs)
    xs = x.cpu().numpy()
    eeq = eeq_closed_form(A)
    assert info["iterations"] == it >= 1 and info["residual_norm"] < 1e-10
    np.testing.assert_allclose(xs[:p.n_u], field.ravel(), rtol=0, atol=1e-11 * np.abs(field).max())
    np.testing.assert_allclose(xs[p.n_u:], eeq, rtol=1e-9)
    np.testing.assert_allclose(kappa.cpu().numpy(), eeq, rtol=1e-9)


def bar(nx=6):
    """Bar [0, nx] x [0, 2] x [0, 2], clamped at x = 0, pulled along x at x = nx; one weak cell."""
    X, C = make_mesh(N.HEX8, (nx, 2, 2), (float(nx), 2.0, 2.0))
    p = Pair(N.HEX8, N.HEX8, X, C, N.gradient_damage_params(), GDM, True)
    xs = node_coords(p)
    left = np.nonzero(np.abs(xs[:, 0]) < 1e-12)[0]
    right = np.nonzero(np.abs(xs[:, 0] - nx) < 1e-12)[0]
    eqs = [(int(3 * r + c), []) for r in left for c in range(3)] + [(int(3 * r), []) for r in right]
    kappa = np.zeros(p.ncells * p.nip)
    cx = X[C].mean(1)[:, 0]
    weak = np.nonzero(np.abs(cx - (nx // 2 + 0.5)) < 1e-9)[0][0]  # a cell of the middle slice
    kappa[weak + np.arange(p.nip) * p.ncells] = 3e-4
    return p, eqs, len(left) * 3, kappa


def test_localising_bar_matches_oracle():
    p, eqs, n_fixed, kappa0 = bar()
    ctx = p.context()
    ctx.set_constraints(0, eqs)
    x = to_dev(np.zeros(p.n_u + p.n_e))
    kappa = to_dev(kappa0.copy())
    xo, ko = np.zeros(p.n_u + p.n_e), kappa0.copy()
    rhs_old = np.zeros(len(eqs))
    for step in range(1, 8):
        rhs_new = np.zeros(len(eqs))
        rhs_new[n_fixed:] = 5e-5 * 6 * step  # displacement of the pulled end: mean strain 5e-5 per step
        info = ctx.quasistatic_step(x, kappa, to_dev(rhs_old), to_dev(rhs_new), **LIN)
        xo, ko, it = O.quasistatic_step(p.orc, [p.n_u, p.n_e], [eqs, []], xo, ko, rhs_old, rhs_new)
        assert info["iterations"] == it, (step, info, it)
        assert rel(x.cpu().numpy(), xo) <= 1e-8, (step, rel(x.cpu().numpy(), xo))
        assert rel(kappa.cpu().numpy(), ko) <= 1e-8, (step, rel(kappa.cpu().numpy(), ko))
        rhs_old = rhs_new
    assert ko.max() > 3e-4  # the damage grows beyond the weak cell's initial history: it localises


def test_no_convergence_leaves_state_unchanged():
    p, eqs, n_fixed, kappa0 = bar(4)
    ctx = p.context()
    ctx.set_constraints(0, eqs)
    x = to_dev(np.zeros(p.n_u + p.n_e))
    kappa = to_dev(kappa0.copy())
    rhs = np.zeros(len(eqs))
    rhs[n_fixed:] = 0.05
    with pytest.raises(N.NutoError) as err:
        ctx.quasistatic_step(x, kappa, None, to_dev(rhs), tol=1e-30, max_iter=1, **LIN)
    assert err.value.code == N.NO_CONVERGENCE
    assert torch.count_nonzero(x).item() == 0
    assert np.array_equal(kappa.cpu().numpy(), kappa0)


def test_linear_elastic_step_without_newton_iterations():
    X, C = make_mesh(N.QUAD4, (5, 3), (2.0, 1.0, 1.0))
    p = Pair(N.QUAD4, N.QUAD4, X, C, N.linear_elastic_params(2, 30000.0, 0.2),
             O.Params(kind=0, dim=2, E=30000.0, nu=0.2, plane_state=0), False)
    eqs, rhs, field = boundary_setup(p, node_coords(p))
    ctx = p.context()
    ctx.set_constraints(0, eqs)
    x = to_dev(np.zeros(p.n_u))
    info = ctx.quasistatic_step(x, None, None, to_dev(rhs), **LIN)
    assert info["iterations"] == 0
    np.testing.assert_allclose(x.cpu().numpy(), field.ravel(), rtol=0, atol=1e-11 * np.abs(field).max())


def test_post_process_of_device_fields(tmp_path):
    """Damage and kappa straight from the device into .vtu cell data (Visualize::PostProcess with
    Cell::Eval), equal to the oracle's IP fields averaged per cell."""
    from test_visualize import read_vtu

    p, eqs, n_fixed, kappa0 = bar()
    ctx = p.context()
    ctx.set_constraints(0, eqs)
    x = to_dev(np.zeros(p.n_u + p.n_e))
    kappa = to_dev(kappa0.copy())
    rhs_old = np.zeros(len(eqs))
    for step in (1, 2):  # the load steps of the localising bar test
        rhs = np.zeros(len(eqs))
        rhs[n_fixed:] = 5e-5 * 6 * step
        ctx.quasistatic_step(x, kappa, to_dev(rhs_old), to_dev(rhs), **LIN)
        rhs_old = rhs
    u, e = x[:p.n_u], x[p.n_u:]
    pp = N.visualize.PostProcess(str(tmp_path))
    pp.define_visualizer("GDM", N.HEX8, p.X, p.C)
    pp.add_dof("GDM", "Displacements", u, N.cell_dofs(p.ranks, 3), 3)
    pp.add_cell_function("GDM", lambda: ctx.eval_ip(N.IP_DAMAGE, u, e, kappa), "Damage")
    pp.add_cell_function("GDM", lambda: ctx.eval_ip(N.IP_KAPPA, u, e, kappa), "Kappa")
    pp.plot(1.0)
    d, _, ncells = read_vtu(str(tmp_path / "GDM0.vtu"))
    un, en, kn = u.cpu().numpy(), e.cpu().numpy(), kappa.cpu().numpy()
    ref_d = p.orc.eval_ip(3, un, en, kn).mean(1)[:, 0]
    ref_k = p.orc.eval_ip(2, un, en, kn).mean(1)[:, 0]
    assert ncells == p.ncells and ref_d.max() > 0
    np.testing.assert_allclose(d["Damage"], ref_d, rtol=1e-6, atol=1e-9)
    np.testing.assert_allclose(d["Kappa"], ref_k, rtol=1e-6, atol=1e-12)
