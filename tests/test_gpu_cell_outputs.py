"""Parity of the device's lumped mass (nuto_build_lumped_mass) and fields at the integration points
(nuto_eval_ip) with the oracle, on jittered meshes of every element pair, through the C ABI.

Bar: the lumped mass within 1e-12 of the oracle's magnitude scale (every term is positive, so that
is 1e-12 relative); fields within 1e-12 relative plus 1e-12 of the field's largest magnitude (the
fields are continuous in their inputs, so branch points such as kappa = kappa0 stay within it).
"""
import numpy as np
import pytest

import nuto_amd as N
import oracle as O
from pairs import Pair, check_values, make_mesh, to_dev
from test_gpu_parity import gdm_params, ld_params, le_params

pytestmark = pytest.mark.gpu

TYPES = [(N.QUAD4, N.QUAD4, (5, 4)), (N.HEX8, N.HEX8, (3, 3, 2)), (N.TET4, N.TET4, (3, 2, 2)),
         (N.TET4, N.TET10, (2, 2, 2)), (N.HEX8, N.HEX20, (2, 2, 2))]


def close(got, ref, what):
    scale = np.abs(ref).max()
    err = np.abs(got - ref)
    lim = 1e-12 * np.abs(ref) + 1e-12 * scale
    assert (err <= lim).all(), f"{what}: worst {err.max()!r} at {int(np.argmax(err - lim))} (scale {scale!r})"


@pytest.mark.parametrize("ct,dt,n", TYPES)
def test_lumped_mass_parity(ct, dt, n):
    dim = N.DIM[ct]
    X, C = make_mesh(ct, n, (1.2, 0.8, 1.0), jitter=0.15)
    p = Pair(ct, dt, X, C, *le_params(dim), two_fields=False)
    ctx = p.context()
    rho = 2.4e-9
    M = ctx.build_lumped_mass(rho)
    ctx.synchronize()
    ref, bound = p.orc.lumped_mass(rho)
    check_values(M.cpu().numpy(), ref, bound, "lumped mass")
    # a second call on a fresh output gives the same bits (row-owner sums, no atomics)
    assert np.array_equal(ctx.build_lumped_mass(rho).cpu().numpy(), M.cpu().numpy())


def test_lumped_mass_gradient_damage_context():
    X, C = make_mesh(N.HEX8, (3, 2, 2), (1.0, 1.0, 1.0), jitter=0.1)
    p = Pair(N.HEX8, N.HEX8, X, C, *gdm_params(), two_fields=True)
    ctx = p.context()
    M = ctx.build_lumped_mass(1.0)
    ref, bound = p.orc.lumped_mass(1.0)
    check_values(M.cpu().numpy(), ref, bound, "lumped mass (two dof types, displacements only)")


@pytest.mark.parametrize("ct,dt,n", TYPES)
def test_eval_ip_linear_elastic(ct, dt, n):
    dim = N.DIM[ct]
    X, C = make_mesh(ct, n, (1.0, 1.0, 1.0), jitter=0.1)
    p = Pair(ct, dt, X, C, *le_params(dim), two_fields=False)
    ctx = p.context()
    u, _, _ = p.inputs()
    for field in (N.IP_STRAIN, N.IP_STRESS):
        got = ctx.eval_ip(field, to_dev(u)).cpu().numpy()
        close(got, p.orc.eval_ip(field, u), f"field {field}")
    with pytest.raises(N.NutoError):
        ctx.eval_ip(N.IP_KAPPA, to_dev(u))


@pytest.mark.parametrize("app", [N.FULL, N.UNILATERAL])
def test_eval_ip_gradient_damage(app):
    X, C = make_mesh(N.HEX8, (3, 3, 3), (1.0, 1.0, 1.0), jitter=0.1)
    p = Pair(N.HEX8, N.HEX8, X, C, *gdm_params(app), two_fields=True)
    ctx = p.context()
    u, e, kap = p.inputs()
    for field in (N.IP_STRAIN, N.IP_STRESS, N.IP_KAPPA, N.IP_DAMAGE, N.IP_EQ_STRAIN):
        got = ctx.eval_ip(field, to_dev(u), to_dev(e), to_dev(kap)).cpu().numpy()
        close(got, p.orc.eval_ip(field, u, e, kap), f"field {field}")


@pytest.mark.parametrize("ct,dt,n", [(N.HEX8, N.HEX8, (3, 2, 2)), (N.TET4, N.TET10, (2, 2, 2)),
                                     (N.HEX8, N.HEX20, (2, 1, 2))])
def test_eval_ip_local_damage(ct, dt, n):
    X, C = make_mesh(ct, n, (1.0, 1.0, 1.0), jitter=0.1)
    p = Pair(ct, dt, X, C, *ld_params(), two_fields=False)
    ctx = p.context()
    u, _, kap = p.inputs()
    for field in (N.IP_STRAIN, N.IP_STRESS, N.IP_KAPPA, N.IP_DAMAGE, N.IP_EQ_STRAIN):
        got = ctx.eval_ip(field, to_dev(u), None, to_dev(kap)).cpu().numpy()
        close(got, p.orc.eval_ip(field, u, None, kap), f"field {field}")


def test_eval_ip_kappa_matches_update_history():
    """The kappa field is what nuto_update_history writes (GradientDamage::Update stores Kappa)."""
    X, C = make_mesh(N.HEX8, (4, 3, 2), (1.0, 1.0, 1.0))
    p = Pair(N.HEX8, N.HEX8, X, C, *gdm_params(), two_fields=True)
    ctx = p.context()
    u, e, kap = p.inputs()
    kd = to_dev(kap)
    field = ctx.eval_ip(N.IP_KAPPA, to_dev(u), to_dev(e), kd).cpu().numpy()[:, :, 0]
    ctx.update_history(to_dev(u), to_dev(e), kd)
    updated = kd.cpu().numpy().reshape(p.nip, p.ncells).T
    assert np.array_equal(field, updated)
