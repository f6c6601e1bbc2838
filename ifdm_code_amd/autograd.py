"""torch.autograd binding of the KDK evolution and its hand-written adjoint.

``evolve(psi0, solver, schedule, K)`` returns psi after len(schedule) steps
(frozen time steps) computed by libifdm.so; its backward is the library's
reverse sweep (ifdm_adjoint_device) with checkpoints every K steps.  psi0 is
a complex128 CUDA tensor on the solver's device, shape (N, N, N).  Gradients
follow torch's convention for real losses of complex tensors,
dL/dRe + i dL/dIm, which is the convention of ifdm_adjoint.

Import torch (and initialise its GPU context) before the library is first
used in a process: torch ships its own HIP runtime with the same soname as
the system one libifdm.so links, and whichever is loaded first serves both.
"""
from __future__ import annotations

import torch


class KDKEvolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, psi0, solver, schedule, K):
        if psi0.dtype != torch.complex128 or not psi0.is_cuda:
            raise TypeError("psi0 must be a complex128 CUDA tensor")
        src = psi0.detach().contiguous()
        torch.cuda.synchronize(src.device)
        solver.set_psi_device(src.data_ptr())
        # run_checkpointed solves Poisson for psi0 first, so the result does
        # not depend on what the solver held before
        ctx.serial = solver.run_checkpointed(schedule, int(K))
        out = torch.empty_like(src)
        solver.get_psi_device(out.data_ptr())
        ctx.solver = solver
        ctx.schedule = list(schedule)
        return out

    @staticmethod
    def backward(ctx, g):
        # the solver keeps the checkpoints of its last forward run only, and
        # they serve one reverse sweep
        if getattr(ctx.solver, "ck_serial", None) != ctx.serial:
            raise RuntimeError("evolve(): the solver ran another forward pass after this one; "
                               "its checkpoints are gone (use one solver per graph)")
        g = g.detach().contiguous().clone()
        torch.cuda.synchronize(g.device)
        ctx.solver.adjoint_device(ctx.schedule, g.data_ptr())
        ctx.solver.ck_serial = None
        return g, None, None, None


def evolve(psi0: torch.Tensor, solver, schedule, K: int = 20) -> torch.Tensor:
    return KDKEvolve.apply(psi0, solver, schedule, K)
