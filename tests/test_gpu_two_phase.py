"""The two forms of the generic 3D matrix assembly agree with the oracle and with each other:
the two-phase kernels (cell matrices by columns, in Gram form with one or two row nodes per lane or
on the matrix cores, NUTO_AMD_CELL_GRAM=0 / 1 / 2 / 3, then the row gather, with the wave-per-node LDS gather or the
lane-per-node gather that serves nodes whose rows exceed the LDS budget) and the row-owner kernel
(NUTO_AMD_TWO_PHASE=0), for linear elasticity and local damage on tet4, tet10, hex8 and hex20
(jittered meshes, per-cell material fields), at the parity bar of the other tests."""
import os

import numpy as np
import pytest

import nuto_amd as N
from pairs import Pair, check_values, make_mesh, to_dev
from test_gpu_parity import ld_params, le_params

pytestmark = pytest.mark.gpu
TYPES = [(N.TET4, N.TET4), (N.TET4, N.TET10), (N.HEX8, N.HEX8), (N.HEX8, N.HEX20)]


def assemble(p, u, kap, env):
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        ctx = p.context()
        R, K = ctx.build_vector_matrix(to_dev(u), None, to_dev(kap))
        return R.cpu().numpy(), K.cpu().numpy()
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


@pytest.mark.parametrize("ct,dt", TYPES)
@pytest.mark.parametrize("kind", ["le", "ld"])
def test_two_phase_forms(ct, dt, kind):
    X, C = make_mesh(ct, (3, 2, 2), (3.0, 2.0, 2.0), jitter=0.1, seed=11)
    p = Pair(ct, dt, X, C, *(le_params(3) if kind == "le" else ld_params()), False)
    u, _, kap = p.inputs(u_range=4e-4, k_range=1.5e-4)
    ref = p.orc.assemble(u, None, kap if kind == "ld" else None)
    forms = {"wave": {}, "lane": {"NUTO_AMD_GATHER_LANE": "1"}, "row-owner": {"NUTO_AMD_TWO_PHASE": "0"},
             "columns": {"NUTO_AMD_CELL_GRAM": "0"}, "gram1": {"NUTO_AMD_CELL_GRAM": "1"},
             "gram2": {"NUTO_AMD_CELL_GRAM": "2"}, "mfma": {"NUTO_AMD_CELL_GRAM": "3"}}
    out = {}
    for name, env in forms.items():
        R, K = assemble(p, u, kap, env)
        check_values(R, ref["R"], ref["R_bound"], f"R ({name})")
        check_values(K, ref["val"], ref["bound"], f"K ({name})")
        out[name] = (R, K)
    # the two gathers add the same staged values in the same order
    assert np.array_equal(out["wave"][0], out["lane"][0]) and np.array_equal(out["wave"][1], out["lane"][1])


def test_two_phase_with_material_fields():
    X, C = make_mesh(N.HEX8, (3, 2, 2), (3.0, 2.0, 2.0), jitter=0.1, seed=5)
    dev, orc = ld_params()
    p = Pair(N.HEX8, N.HEX20, X, C, dev, orc, False)
    u, _, kap = p.inputs(u_range=4e-4, k_range=1.5e-4)
    rng = np.random.default_rng(3)
    E = rng.uniform(25000.0, 35000.0, p.ncells)
    ft = rng.uniform(2.5, 3.5, p.ncells)
    gf = rng.uniform(0.05, 0.15, p.ncells)
    p.orc.set_theta(E, ft, gf)
    ref = p.orc.assemble(u, None, kap)
    for env in ({}, {"NUTO_AMD_TWO_PHASE": "0"}, {"NUTO_AMD_CELL_GRAM": "2"}, {"NUTO_AMD_CELL_GRAM": "3"},
                {"NUTO_AMD_CELL_GRAM": "0"}):
        old = {k: os.environ.get(k) for k in env}
        os.environ.update(env)
        try:
            ctx = p.context()
            ctx.set_material_fields(E, ft, gf)
            R, K = ctx.build_vector_matrix(to_dev(u), None, to_dev(kap))
        finally:
            for k, v in old.items():
                if v is None:
                    os.environ.pop(k, None)
                else:
                    os.environ[k] = v
        check_values(R.cpu().numpy(), ref["R"], ref["R_bound"], f"R {env}")
        check_values(K.cpu().numpy(), ref["val"], ref["bound"], f"K {env}")


@pytest.mark.parametrize("ct,dt", [(N.TET4, N.TET10), (N.HEX8, N.HEX20)])
def test_two_phase_unilateral_damage(ct, dt):
    """LinearElasticDamage with UNILATERAL application (H = 0 in compression) in the cell kernel."""
    X, C = make_mesh(ct, (2, 2, 2), (2.0, 2.0, 2.0), jitter=0.1, seed=9)
    p = Pair(ct, dt, X, C, *ld_params(N.UNILATERAL), False)
    u, _, kap = p.inputs(u_range=6e-4, k_range=2e-4)
    ref = p.orc.assemble(u, None, kap)
    for form in ("1", "2", "3", "0"):  # the second Gram runs forms 2 and 3 as form 1
        R, K = assemble(p, u, kap, {"NUTO_AMD_CELL_GRAM": form})
        check_values(R, ref["R"], ref["R_bound"], f"R (form {form})")
        check_values(K, ref["val"], ref["bound"], f"K (form {form})")
