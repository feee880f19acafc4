"""Builds the same problem on the device path (nuto_amd, through the C ABI) and in the oracle,
from the same mesh arrays and seeded inputs, and compares them with the parity rule:
pattern / row_ptr / col / dof numbering bit-equal; values |x - ref| <= 1e-12 * sum|contributions|.
"""
import numpy as np

import nuto_amd as N
import oracle as O
from nuto_amd import synthetic as S

RTOL = 1e-12

RULES = {  # (coord type, dof type) -> (rule kind, n)
    (N.QUAD4, N.QUAD4): (0, 2),
    (N.HEX8, N.HEX8): (0, 2),
    (N.TET4, N.TET4): (1, 1),
    (N.TET4, N.TET10): (1, 2),
    (N.TET10, N.TET10): (1, 2),
    (N.HEX8, N.HEX20): (0, 3),
}


def make_mesh(coord_type, n, lengths=None, jitter=0.0, seed=7):
    if coord_type == N.QUAD4:
        X, C = N.structured_mesh(N.QUAD4, n[0], n[1], 0, lengths or (1.0, 1.0, 1.0))
    else:
        X, C = N.structured_mesh(coord_type, n[0], n[1], n[2], lengths or (1.0, 1.0, 1.0), jitter, seed)
    return X, C


class Pair:
    def __init__(self, coord_type, dof_type, X, C, params_dev, params_orc, two_fields, rule=None):
        self.coord_type, self.dof_type = coord_type, dof_type
        self.X, self.C = X, C
        dim = N.DIM[coord_type]
        self.ncomp = dim
        self.ranks, self.nn = N.dof_layer(coord_type, C, dof_type)
        self.two = two_fields
        self.rule = rule or RULES[(coord_type, dof_type)]
        # oracle side: NuTo's coordinate-matching AddDofInterpolation on the same arrays
        self.om = O.Mesh.from_arrays(coord_type, X, C)
        Ld = self.om.add_dof(dof_type, dim)
        Le = self.om.add_dof(dof_type, 1) if two_fields else -1
        self.orc = O.Problem(self.om, params_orc, self.rule[0], self.rule[1], Ld, Le)
        self.layer_d = self.om.dof_layer(Ld)
        self.params_dev = params_dev
        self.ncells = C.shape[0]
        self.nip = self.orc.nip
        self.n_u = self.nn * dim
        self.n_e = self.nn if two_fields else 0

    def numbering_matches(self):
        lay = self.layer_d
        orc_rank = lay["numbers"][lay["conn"], 0] // self.ncomp
        return np.array_equal(orc_rank, self.ranks)

    def context(self):
        ctx = N.Context()
        ctx.set_cells(self.coord_type, self.X, self.C)
        ctx.set_dof_type(0, self.dof_type, self.ncomp, self.n_u, N.cell_dofs(self.ranks, self.ncomp))
        if self.two:
            ctx.set_dof_type(1, self.dof_type, 1, self.n_e, N.cell_dofs(self.ranks, 1))
        ctx.set_integrand(self.params_dev)
        ctx.set_rule(*self.rule)
        ctx.build_pattern()
        return ctx

    def inputs(self, u_range=5e-3, e_range=5e-4, k_range=3e-4, seed_shift=0):
        u = S.uniform(S.SEED_U + seed_shift, self.n_u, -u_range, u_range)
        e = S.uniform(S.SEED_E + seed_shift, self.n_e, 0.0, e_range) if self.two else None
        kap = S.uniform(S.SEED_KAPPA + seed_shift, self.ncells * self.nip, 0.0, k_range)
        return u, e, kap


def to_dev(a):
    import torch

    return None if a is None else torch.from_numpy(np.ascontiguousarray(a)).cuda()


def check_values(got, ref, bound, what, floor=0.0):
    """Parity rule; `floor` is an absolute allowance for entries that are pure rounding noise
    (zero in exact arithmetic), used only where a test states why its inputs produce them."""
    err = np.abs(got - ref)
    lim = RTOL * bound + floor
    bad = err > lim
    if bad.any():
        i = int(np.argmax(err - lim))
        raise AssertionError(f"{what}: {int(bad.sum())} entries beyond 1e-12*sum|contrib|; worst idx {i}: got "
                             f"{got[i]!r} ref {ref[i]!r} bound {bound[i]!r}")
    big = np.abs(ref) > 1e-8 * max(np.abs(ref).max(), 1e-300)
    return float((err[big] / np.abs(ref[big])).max()) if big.any() else 0.0
