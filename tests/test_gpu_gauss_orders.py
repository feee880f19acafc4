"""Tensor Gauss rules above order 3 (Gauss1D by Newton iteration on the Legendre polynomial, any
order the reference's IntegrationTypeTensorProduct takes up to 10; kernels for quad4 with 3, 4, 5
points and hex8 with 3, 4 points per direction, the hex8 cell matrices in each of their four forms)
against the oracle, which restates the reference's
own Newton iteration (Legendre.cpp:23-55): pattern bit-equal, R and K within 1e-12 of the
contribution magnitudes, for linear elasticity, local damage and gradient damage. Gradient damage
with other than 2x2x2 points runs the generic row-owner kernel (the node kernel and the adjoint
take 8 IPs)."""
import os

import numpy as np
import pytest

import nuto_amd as N
from pairs import Pair, check_values, make_mesh, to_dev
from test_gpu_parity import gdm_params, ld_params, le_params

pytestmark = pytest.mark.gpu


def run(p, two=False, env=None):
    u, e, kap = p.inputs(u_range=4e-4, k_range=1.5e-4)
    old = {k: os.environ.get(k) for k in (env or {})}
    os.environ.update(env or {})
    try:
        ctx = p.context()
        R, K = ctx.build_vector_matrix(to_dev(u), to_dev(e), to_dev(kap))
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    ref = p.orc.assemble(u, e, kap)
    rp, ci = ctx.pattern()
    assert np.array_equal(rp, ref["row_ptr"]) and np.array_equal(ci, ref["col"])
    check_values(R.cpu().numpy(), ref["R"], ref["R_bound"], f"R (rule {p.rule})")
    check_values(K.cpu().numpy(), ref["val"], ref["bound"], f"K (rule {p.rule})")


@pytest.mark.parametrize("n", [3, 4, 5])
def test_quad4_linear_elastic(n):
    X, C = make_mesh(N.QUAD4, (4, 3), (1.0, 0.75, 1.0))
    p = Pair(N.QUAD4, N.QUAD4, X, C, *le_params(2), False, rule=(0, n))
    assert p.nip == n * n
    run(p)


@pytest.mark.parametrize("form", ["0", "1", "2", "3"])  # cell matrices: columns, Gram (2 forms), MFMA
@pytest.mark.parametrize("n", [3, 4])
@pytest.mark.parametrize("kind", ["le", "ld"])
def test_hex8_momentum(n, kind, form):
    X, C = make_mesh(N.HEX8, (3, 2, 2), (3.0, 2.0, 2.0), jitter=0.1, seed=5)
    p = Pair(N.HEX8, N.HEX8, X, C, *(le_params(3) if kind == "le" else ld_params()), False, rule=(0, n))
    assert p.nip == n ** 3
    run(p, env={"NUTO_AMD_CELL_GRAM": form})


@pytest.mark.parametrize("n", [3, 4])
def test_hex8_gradient_damage(n):
    X, C = make_mesh(N.HEX8, (3, 2, 2), (3.0, 2.0, 2.0), jitter=0.1, seed=6)
    p = Pair(N.HEX8, N.HEX8, X, C, *gdm_params(), True, rule=(0, n))
    run(p, two=True)


def test_orders_without_kernels_are_refused():
    X, C = make_mesh(N.HEX8, (2, 2, 2), (2.0, 2.0, 2.0))
    p = Pair(N.HEX8, N.HEX8, X, C, *le_params(3), False, rule=(0, 2))
    ctx = p.context()
    with pytest.raises(N.NutoError):
        ctx.set_rule(0, 11)  # beyond the rules Gauss1D provides
    ctx.set_rule(0, 5)  # 125 points: a valid rule, no kernel instantiated for it
    ctx.build_pattern()
    u = to_dev(np.zeros(p.n_u))
    with pytest.raises(N.NutoError) as ei:
        ctx.build_vector_matrix(u, None, None)
    assert ei.value.code == 5  # NUTO_ERR_UNSUPPORTED
