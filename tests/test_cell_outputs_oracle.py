"""Pins of the oracle's lumped mass (SimpleAssembler::BuildDiagonallyLumpedMatrix with
DynamicMomentumBalance::Hessian2, SimpleAssembler.cpp:67-98, DynamicMomentumBalance.h:22-30) and of
its fields at the integration points (Cell::Eval, Cell.h:80-93), by closed forms that do not
restate the oracle's formulas:

* lumped mass: the reference scales each cell's consistent-mass diagonal by (sum of all entries /
  sum of the diagonal) / number of components, so every component of the lumped masses sums to
  rho V / dim (not rho V); on regular grids of hex8, quad4 and tet4 cells every node of a cell
  receives the same share, rho V_cell / (nodes per cell * dim).
* fields: a linear displacement u = A x is reproduced by every element, so the strain at every IP
  is the engineering strain of A ([xx, yy, zz, yz, zx, xy] order, shear doubled); the linear
  elastic stress is lambda tr(eps) + 2 mu eps (normal) and mu gamma (shear); kappa, damage and the
  equivalent strain compose the scalar laws pinned in test_oracle_pins.py.
"""
import numpy as np
import pytest

import nuto_amd as N
import oracle as O

from pairs import RULES, make_mesh

TYPES = [(N.QUAD4, N.QUAD4), (N.HEX8, N.HEX8), (N.TET4, N.TET4), (N.TET4, N.TET10), (N.HEX8, N.HEX20)]
NN = {(N.QUAD4, N.QUAD4): (3, 2), (N.HEX8, N.HEX8): (3, 2, 2), (N.TET4, N.TET4): (2, 2, 1),
      (N.TET4, N.TET10): (2, 1, 2), (N.HEX8, N.HEX20): (2, 1, 1)}
LENGTHS = (2.0, 1.5, 0.5)


def problem(ct, dt, params, jitter=0.0, two=False):
    X, C = make_mesh(ct, NN[(ct, dt)], LENGTHS, jitter=jitter)
    dim = N.DIM[ct]
    om = O.Mesh.from_arrays(ct, X, C)
    Ld = om.add_dof(dt, dim)
    Le = om.add_dof(dt, 1) if two else -1
    rule = RULES[(ct, dt)]
    return X, C, om, O.Problem(om, params, rule[0], rule[1], Ld, Le), Ld, Le


def dof_node_coords(X, C, ct, dt, om, layer):
    """Coordinates of the dof nodes of a layer: the coordinate element at the dof node's local
    coordinates (the layer's nodes are placed there, MeshFemDofConvert.cpp:164-210)."""
    lay = om.dof_layer(layer)
    conn, numbers = lay["conn"], lay["numbers"]
    loc = O.local_coords(dt)
    dim = N.DIM[ct]
    xs = np.zeros((numbers.shape[0], dim))
    for c in range(C.shape[0]):
        for a in range(conn.shape[1]):
            Nc, _ = O.shape(ct, loc[a])
            xs[conn[c, a]] = Nc @ X[C[c], :dim]
    return xs, numbers


@pytest.mark.parametrize("ct,dt", TYPES)
def test_lumped_mass_component_sums(ct, dt):
    dim = N.DIM[ct]
    _, _, _, P, _, _ = problem(ct, dt, O.Params(kind=0, dim=dim, E=1.0, nu=0.2))
    rho = 2.5
    M, Mb = P.lumped_mass(rho)
    V = np.prod(LENGTHS[:dim])
    sums = M.reshape(-1, dim).sum(axis=0)
    assert np.allclose(sums, rho * V / dim, rtol=1e-13, atol=0), (sums, rho * V / dim)
    assert np.all(M > 0) and np.allclose(Mb, M, rtol=1e-15)
    # every component of a node holds the same mass
    assert np.array_equal(M.reshape(-1, dim), np.repeat(M.reshape(-1, dim)[:, :1], dim, axis=1))


@pytest.mark.parametrize("ct,dt,npc", [(N.HEX8, N.HEX8, 8), (N.QUAD4, N.QUAD4, 4), (N.TET4, N.TET4, 4)])
def test_lumped_mass_regular_grid_shares(ct, dt, npc):
    dim = N.DIM[ct]
    X, C, om, P, Ld, _ = problem(ct, dt, O.Params(kind=0, dim=dim, E=1.0, nu=0.2))
    rho = 1.75
    M, _ = P.lumped_mass(rho)
    lay = om.dof_layer(Ld)
    nodes = lay["numbers"][:, 0] // dim  # node -> rank
    V_cell = np.prod(LENGTHS[:dim]) / C.shape[0]
    count = np.bincount(np.asarray(C).ravel(), minlength=X.shape[0])  # cells per coordinate node
    # coordinate nodes are the dof nodes here; map node id -> rank through the layer's conn
    rank_of = np.zeros(X.shape[0], dtype=np.int64)
    rank_of[np.asarray(C).ravel()] = nodes[lay["conn"].ravel()]
    expect = count * rho * V_cell / (npc * dim)
    got = M.reshape(-1, dim)[rank_of, 0]
    assert np.allclose(got, expect, rtol=1e-13), np.abs(got - expect).max()


@pytest.mark.parametrize("ct,dt", TYPES)
def test_eval_strain_and_stress_of_a_linear_field(ct, dt):
    dim = N.DIM[ct]
    E, nu = 30000.0, 0.2
    X, C, om, P, Ld, _ = problem(ct, dt, O.Params(kind=0, dim=dim, E=E, nu=nu, plane_state=1), jitter=0.1)
    A = np.array([[1.0, 2.0, -0.5], [0.3, -1.2, 0.7], [0.9, 0.4, 0.25]])[:dim, :dim] * 1e-3
    xs, numbers = dof_node_coords(X, C, ct, dt, om, Ld)
    u = np.zeros(P.ndof[0])
    for node in range(xs.shape[0]):
        for c in range(dim):
            u[numbers[node, c]] = A[c] @ xs[node]
    eps = P.eval_ip(0, u)
    if dim == 2:
        exact = np.array([A[0, 0], A[1, 1], A[0, 1] + A[1, 0]])
    else:
        exact = np.array([A[0, 0], A[1, 1], A[2, 2], A[1, 2] + A[2, 1], A[0, 2] + A[2, 0], A[0, 1] + A[1, 0]])
    assert np.allclose(eps, exact[None, None, :], rtol=0, atol=1e-14), np.abs(eps - exact).max()
    sig = P.eval_ip(1, u)
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    tr = exact[:dim].sum()
    ref = np.concatenate([lam * tr + 2 * mu * exact[:dim], mu * exact[dim:]])
    assert np.allclose(sig, ref[None, None, :], rtol=1e-12, atol=1e-12 * np.abs(ref).max())


def test_eval_damage_fields_compose_the_scalar_laws():
    E, nu, k, k0, beta, alpha = 30000.0, 0.2, 10.0, 1e-4, 100.0, 0.99
    p = O.Params(kind=1, dim=3, E=E, nu=nu, k=k, kappa0=k0, beta=beta, alpha=alpha, c=4.0)
    X, C, om, P, Ld, Le = problem(N.HEX8, N.HEX8, p, jitter=0.1, two=True)
    rng = np.random.default_rng(5)
    u = rng.uniform(-2e-4, 2e-4, P.ndof[0])
    c0 = 1.7e-4
    e = np.full(P.ndof[1], c0)  # constant nodal eeq: eeq = c0 at every IP (partition of unity)
    kap = rng.uniform(0.0, 3e-4, P.ncells * P.nip)
    eps = P.eval_ip(0, u, e, kap)
    kappa = P.eval_ip(2, u, e, kap)[:, :, 0]
    hist = kap.reshape(P.nip, P.ncells).T  # history column-major: cell + ip * n_cells
    assert np.allclose(kappa, np.maximum(hist, c0), rtol=1e-15, atol=0)
    omega = P.eval_ip(3, u, e, kap)[:, :, 0]
    sig = P.eval_ip(1, u, e, kap)
    eq = P.eval_ip(4, u, e, kap)[:, :, 0]
    for c in range(P.ncells):
        for ip in range(P.nip):
            w, _ = O.damage_exp(k0, beta, alpha, kappa[c, ip])
            assert omega[c, ip] == pytest.approx(w, rel=1e-14, abs=1e-300)
            s_ref = O.led(E, nu, 0, eps[c, ip], w)[0]
            assert np.allclose(sig[c, ip], s_ref, rtol=1e-13, atol=1e-13 * np.abs(s_ref).max())
            assert eq[c, ip] == pytest.approx(O.mises(nu, k, eps[c, ip])[0], rel=1e-14)


def test_eval_damage_fields_rejected_for_linear_elasticity():
    _, _, _, P, _, _ = problem(N.HEX8, N.HEX8, O.Params(kind=0, dim=3, E=1.0, nu=0.2))
    with pytest.raises(ValueError):
        P.eval_ip(2, np.zeros(P.ndof[0]))
