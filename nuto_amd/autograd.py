"""torch.autograd binding of the gradient-damage residual R(u, eeq; theta).

`residual(ctx, u, e, kappa, theta)` returns R = BuildVector (SimpleAssembler.cpp:38-65 with
GradientDamage<3>::Gradient, GradientDamage.h:70-90) as a differentiable function of the
displacements u, the nonlocal equivalent strains e and the per-cell material fields
theta = (E, f_t, G_f) (shape (3, n_cells)); the history kappa is a constant. The backward pass is
the hand-written adjoint kernel (nuto_vjp): K^T lambda for (u, e) and dR/dtheta^T lambda for theta,
with the integration-point state recomputed, not stored.
"""
import torch


class GradientDamageResidual(torch.autograd.Function):
    @staticmethod
    def forward(fctx, ctx, u, e, kappa, theta):
        ctx.set_material_fields_device(theta)
        R = ctx.build_vector(u.contiguous(), e.contiguous(), kappa.contiguous())
        fctx.ctx = ctx
        fctx.save_for_backward(u, e, kappa, theta)
        return R

    @staticmethod
    def backward(fctx, gR):
        u, e, kappa, theta = fctx.saved_tensors
        ctx = fctx.ctx
        ctx.set_material_fields_device(theta)
        d_state, d_theta = ctx.vjp(u.contiguous(), e.contiguous(), kappa.contiguous(), gR.contiguous(),
                                   want_theta=fctx.needs_input_grad[4])
        n_u = u.numel()
        du = d_state[:n_u] if fctx.needs_input_grad[1] else None
        de = d_state[n_u:] if fctx.needs_input_grad[2] else None
        return None, du, de, None, d_theta


def residual(ctx, u, e, kappa, theta):
    """Differentiable R(u, e; theta) on an assembled-pattern gradient-damage Context."""
    return GradientDamageResidual.apply(ctx, u, e, kappa, theta)
