"""torch.autograd wrappers over the discrete-adjoint entry points of the C ABI.

Tensors are CUDA float64 of shape (nElems, N+1, N+1, N+1, 5), C-contiguous: the same memory as the reference's
U(5,0:N,0:N,0:N,nElems) in Fortran order (axes reversed), so their device pointers go straight through the
ABI.  The operator's context holds one state and one recorded trajectory at a time: RHS.backward resets the
state to the saved input, LSERK.backward needs the trajectory of its own forward call (no forward of the same
operator in between).
"""
from __future__ import annotations

import torch

from .dg import DGOperator


def _sync():
    # the ABI works on its own stream and synchronises it before returning; torch's producers must be done
    torch.cuda.current_stream().synchronize()


def _check(U: torch.Tensor, op: DGOperator):
    if not (U.is_cuda and U.dtype == torch.float64):
        raise TypeError("U must be a CUDA float64 tensor")
    want = (op.nElems, op.n, op.n, op.n, 5)
    if tuple(U.shape) != want:
        raise ValueError(f"U has shape {tuple(U.shape)}, expected {want}")


class RHS(torch.autograd.Function):
    """Ut = DGTimeDerivative_weakForm(U); backward: Ubar = (dR/dU)^T Utbar."""

    @staticmethod
    def forward(ctx, U, op):
        _check(U, op)
        U = U.contiguous()
        Ut = torch.empty_like(U)
        _sync()
        op.set_state_device(U.data_ptr())
        op.rhs_only()
        op.get_ut_device(Ut.data_ptr())
        ctx.op = op
        ctx.save_for_backward(U)
        return Ut

    @staticmethod
    def backward(ctx, lam):
        (U,) = ctx.saved_tensors
        lam = lam.contiguous()
        Ubar = torch.empty_like(U)
        _sync()
        ctx.op.set_state_device(U.data_ptr())
        ctx.op.rhs_vjp_device(lam.data_ptr(), Ubar.data_ptr())
        return Ubar, None


class LSERK(torch.autograd.Function):
    """U(t+nSteps*dt) = TimeStepByLSERKW2^nSteps(U); backward: checkpointed reverse sweep."""

    @staticmethod
    def forward(ctx, U0, op, dt, nSteps, ckptEvery):
        _check(U0, op)
        U0 = U0.contiguous()
        out = torch.empty_like(U0)
        _sync()
        op.set_state_device(U0.data_ptr())
        op.forward_ckpt(0.0, dt, nSteps, ckptEvery)
        op.get_state_device(out.data_ptr())
        ctx.op = op
        return out

    @staticmethod
    def backward(ctx, Ubar):
        g = Ubar.contiguous().clone()
        _sync()
        ctx.op.lserk_adjoint_device(g.data_ptr())
        return g, None, None, None, None


def rhs(U, op):
    return RHS.apply(U, op)


def lserk(U0, op, dt, nSteps, ckptEvery):
    return LSERK.apply(U0, op, dt, nSteps, ckptEvery)
