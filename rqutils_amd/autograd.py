r"""Differentiable subspace action ``y = H(alpha) x`` and Hellmann-Feynman gradients.

No counterpart in the reference: JAX reverse-mode autodiff cannot go through
``get_diagonal``'s ``while_loop`` (``rqutils/sqd.py:611-628``).  Here

* ``grad_x = H^dagger grad_y = H grad_y`` (``H`` is Hermitian for real alpha)
  reuses the forward kernel;
* ``grad_alpha_k = Re sum_b conj(gy_b) (P_k x)_b`` with
  ``(P_k x)_b = (-i)^{x_k.z_k} (-1)^{popcount(b & z_k)} x[b ^ x_k]`` is one
  pass of the gradient kernel (a per-term wave reduction over the hits);
* ``hellmann_feynman(plan, psi)`` gives ``dE/dalpha_k = <psi|P_k|psi>`` for a
  normalised ground vector.

``alpha`` is indexed like ``plan.tables`` (X-grouped order); use
``TermTables.from_masks(..., return_index=True)`` to map input terms to it.
"""
import torch

from rqutils_amd.plan import Plan


class PauliMatvec(torch.autograd.Function):
    """``y = H(alpha) x`` on a ``Plan``; differentiable in ``alpha`` and ``x``."""

    @staticmethod
    def forward(ctx, alpha, x, plan):
        plan.set_alphas(alpha.detach())
        y = plan.matvec(x.detach().contiguous())
        ctx.save_for_backward(x)
        ctx.plan = plan
        return y

    @staticmethod
    def backward(ctx, gy):
        (x,) = ctx.saved_tensors
        plan = ctx.plan
        gy = gy.contiguous()
        grad_alpha = grad_x = None
        if ctx.needs_input_grad[1]:
            grad_x = plan.matvec(gy)
        if ctx.needs_input_grad[0]:
            grad_alpha = plan.grad_coef(x.contiguous(), gy.to(x.dtype))
        return grad_alpha, grad_x, None


def pauli_matvec(plan: Plan, alpha: torch.Tensor, x: torch.Tensor) -> torch.Tensor:
    """Differentiable ``H(alpha) x`` (alpha: float64, one per term of ``plan.tables``)."""
    return PauliMatvec.apply(alpha, x, plan)


def hellmann_feynman(plan: Plan, psi: torch.Tensor) -> torch.Tensor:
    """``dE/dalpha_k = <psi|P_k|psi> / <psi|psi>`` for the ground vector ``psi``."""
    return plan.expect_terms(psi)
