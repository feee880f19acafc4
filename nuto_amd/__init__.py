"""nuto_amd: MI355X-native element assembly for NuTo's SimpleAssembler path.

The compute path is libnuto_amd.so (HIP kernels for gfx950 behind the C ABI in
include/nuto_amd.h); this package is its Python binding.
"""
from .binding import (DIM, NUM_NODES, FULL, CONSTANT, DECREASING, GRADIENT_DAMAGE, HEX8, HEX20, IP_DAMAGE, IP_EQ_STRAIN, IP_KAPPA,
                      IP_STRAIN, IP_STRESS, MOMENTUM_LINEAR_ELASTIC, MOMENTUM_LOCAL_DAMAGE, SOLVER_BICGSTAB, SOLVER_CG,
                      PLANE_STRAIN, PLANE_STRESS, QUAD4, TET4, TET10, UNILATERAL, Batch, Context, GmshMesh, NutoError,
                      Params, cell_dofs, dof_layer, gradient_damage_params, lib, linear_elastic_params,
                      local_damage_params, morton_decode2d, morton_decode3d, morton_encode2d, morton_encode3d,
                      morton_neighbor3d, morton_order, structured_mesh, NO_CONVERGENCE, adaptive_solve)

from . import autograd  # noqa: E402  (torch.autograd binding of the residual)
from . import visualize  # noqa: E402  (Visualize::PostProcess: .vtu / .pvd output)

__all__ = [
    "DIM", "NUM_NODES", "Batch", "Context", "NutoError", "Params", "cell_dofs", "dof_layer", "gradient_damage_params", "lib",
    "linear_elastic_params", "local_damage_params", "structured_mesh", "QUAD4", "HEX8", "TET4", "TET10", "HEX20",
    "MOMENTUM_LINEAR_ELASTIC", "GRADIENT_DAMAGE", "MOMENTUM_LOCAL_DAMAGE", "PLANE_STRESS", "PLANE_STRAIN", "FULL",
    "UNILATERAL", "CONSTANT", "DECREASING", "GmshMesh", "morton_order", "morton_encode3d", "morton_encode2d", "morton_decode3d",
    "morton_decode2d", "morton_neighbor3d", "NO_CONVERGENCE", "adaptive_solve", "IP_STRAIN", "IP_STRESS", "IP_KAPPA", "IP_DAMAGE", "IP_EQ_STRAIN", "SOLVER_CG", "SOLVER_BICGSTAB", "autograd", "visualize",
]
