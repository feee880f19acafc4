"""The allocation-free assembly entry point (nuto_build_vector_matrix_noalloc) against the allocating
one (nuto_build_vector_matrix) through the C ABI: bit-identical R and K on two calls in a row, the
second on a smaller mesh; sizes beyond the scratch maximum and contexts without prepared node
tables are refused with an error."""
import ctypes

import numpy as np
import pytest
import torch

import nuto_amd as N
from nuto_amd import binding as B
from nuto_amd import synthetic as S

pytestmark = pytest.mark.gpu
ERR_ARGUMENT, ERR_STATE, ERR_UNSUPPORTED = 1, 2, 5  # NUTO_ERR_* (include/nuto_amd.h)


def _context(n, jitter=0.0):
    X, C = N.structured_mesh(N.HEX8, n, n, n, (float(n), float(n), float(n)), jitter, 3)
    ranks, nn = N.dof_layer(N.HEX8, C, N.HEX8)
    ctx = N.Context()
    ctx.set_cells(N.HEX8, X, C)
    ctx.set_dof_type(0, N.HEX8, 3, 3 * nn, N.cell_dofs(ranks, 3))
    ctx.set_dof_type(1, N.HEX8, 1, nn, N.cell_dofs(ranks, 1))
    ctx.set_integrand(N.gradient_damage_params())
    ctx.set_rule(0, 2)
    ctx.build_pattern()  # with the integrand and rule set: also the node kernel's tables
    nc = C.shape[0]
    u = torch.from_numpy(S.uniform(S.SEED_U, 3 * nn, -5e-3, 5e-3)).cuda()
    e = torch.from_numpy(S.uniform(S.SEED_E, nn, 0.0, 5e-4)).cuda()
    kap = torch.from_numpy(S.uniform(S.SEED_KAPPA, nc * 8, 0.0, 3e-4)).cuda()
    return ctx, u, e, kap


def _noalloc(ctx, u, e, kap, R, K):
    """The C ABI call itself (status code), with device pointers of caller-owned tensors."""
    return B.lib().nuto_build_vector_matrix_noalloc(ctx.h, *B._ptrs(u, e, kap, R, K))


@pytest.mark.parametrize("jitter", [0.0, 0.15])
def test_noalloc_matches_the_allocating_entry_point(jitter):
    """Affine cells (records without J^-1) and jittered cells (records with it); the second call
    runs on a smaller mesh, so the scratch holds the first call's larger contents."""
    for n in (30, 12):
        ctx, u, e, kap = _context(n, jitter)
        R = torch.full((ctx.n_rows,), float("nan"), dtype=torch.float64, device="cuda")
        K = torch.full((ctx.nnz,), float("nan"), dtype=torch.float64, device="cuda")
        assert _noalloc(ctx, u, e, kap, R, K) == 0, B.lib().nuto_last_error(ctx.h).decode()
        Rr, Kr = ctx.build_vector_matrix(u, e, kap)
        torch.cuda.synchronize()
        assert torch.equal(R, Rr) and torch.equal(K, Kr), n


def test_noalloc_refusals():
    ctx, u, e, kap = _context(40)  # 64000 cells: records beyond the scratch maximum
    R = torch.empty(ctx.n_rows, dtype=torch.float64, device="cuda")
    K = torch.empty(ctx.nnz, dtype=torch.float64, device="cuda")
    assert _noalloc(ctx, u, e, kap, R, K) == ERR_ARGUMENT
    assert "scratch" in B.lib().nuto_last_error(ctx.h).decode()
    # a context whose node tables were not built with the pattern (integrand set afterwards)
    ctx2, u2, e2, kap2 = _context(6)
    ctx2.set_integrand(N.gradient_damage_params())
    R2 = torch.empty(ctx2.n_rows, dtype=torch.float64, device="cuda")
    K2 = torch.empty(ctx2.nnz, dtype=torch.float64, device="cuda")
    assert _noalloc(ctx2, u2, e2, kap2, R2, K2) == ERR_STATE
    # after one allocating call the tables exist and the call goes through
    ctx2.build_vector_matrix(u2, e2, kap2)
    assert _noalloc(ctx2, u2, e2, kap2, R2, K2) == 0
    Rr, Kr = ctx2.build_vector_matrix(u2, e2, kap2)
    torch.cuda.synchronize()
    assert torch.equal(R2, Rr) and torch.equal(K2, Kr)
    # another integrand: not provided
    ctx3 = N.Context()
    X, C = N.structured_mesh(N.HEX8, 3, 3, 3, (3.0, 3.0, 3.0), 0.0, 3)
    ranks, nn = N.dof_layer(N.HEX8, C, N.HEX8)
    ctx3.set_cells(N.HEX8, X, C)
    ctx3.set_dof_type(0, N.HEX8, 3, 3 * nn, N.cell_dofs(ranks, 3))
    ctx3.set_integrand(N.linear_elastic_params(3, 30000.0, 0.2))
    ctx3.set_rule(0, 2)
    ctx3.build_pattern()
    u3 = torch.zeros(3 * nn, dtype=torch.float64, device="cuda")
    R3 = torch.empty(ctx3.n_rows, dtype=torch.float64, device="cuda")
    K3 = torch.empty(ctx3.nnz, dtype=torch.float64, device="cuda")
    assert B.lib().nuto_build_vector_matrix_noalloc(ctx3.h, *B._ptrs(u3, None, None, R3, K3)) == ERR_UNSUPPORTED
