"""The parity rule is not vacuous: plausible mistakes planted in the oracle (a dropped term, a
flipped sign, a wrong shear factor, a dropped history max, a transposed operand, the nonlocal length
squared or negated in R and K alike) change R or K by
more than 1e-12 of the contribution magnitudes on the committed 3x3x3 gradient-damage fixture, so
the same rule applied to the device path would reject each of them.
"""
import os

import numpy as np
import pytest

import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def assemble(fix, mutation, jitter=False):
    f = fix
    om = O.Mesh.from_arrays(O.HEX8, f["coords_j"] if jitter else f["coords"], f["conn"])
    Ld, Le = om.add_dof(O.HEX8, 3), om.add_dof(O.HEX8, 1)
    p = O.Params(kind=1, dim=3, E=30000.0, nu=0.2, k=10.0, kappa0=1e-4, beta=100.0, alpha=0.99, c=4.0)
    O.set_mutation(mutation)
    try:
        return O.Problem(om, p, 0, 2, Ld, Le).assemble(f["u"], f["e"], f["kappa"])
    finally:
        O.set_mutation(0)


@pytest.fixture(scope="module")
def fix():
    f = dict(np.load(os.path.join(GOLD, "hex8_gdm_3x3x3.npz")))
    # a jittered copy of the coordinates (J not diagonal, so J^-1 and J^-T differ)
    rng = np.random.default_rng(3)
    f["coords_j"] = f["coords"] + rng.uniform(-0.1, 0.1, f["coords"].shape) * (f["coords"].max() / 3)
    return f


def violates(ref, got):
    """Does `got` fail the parity rule against `ref` (bounds from ref's magnitude shadow)?"""
    r = np.abs(got["R"] - ref["R"]) > 1e-12 * ref["R_bound"]
    k = np.abs(got["val"] - ref["val"]) > 1e-12 * ref["bound"]
    return bool(r.any() or k.any())


def test_unmutated_oracle_passes_its_own_rule(fix):
    a, b = assemble(fix, 0), assemble(fix, 0)
    assert not violates(a, b)


@pytest.mark.parametrize("mutation", [1, 2, 3, 4, 5, 7, 8])
def test_planted_mistake_is_rejected(fix, mutation):
    ref = assemble(fix, 0)
    assert violates(ref, assemble(fix, mutation)), f"mutation {mutation} passed the parity rule"


def test_transposed_jacobian_inverse_is_rejected(fix):
    ref = assemble(fix, 0, jitter=True)
    assert violates(ref, assemble(fix, 6, jitter=True))
