"""torch.autograd binding of the discrete adjoint (include/dns.h,
dns_adjoint_*): the final enstrophy after `nsteps` RK4 steps as a
differentiable function of the initial velocity.

    J = final_enstrophy(u0, nu, dt, nsteps, every=20)   # u0: [3,n,n,n] float64
    J.backward()                                        # u0.grad = dJ/du0

The forward pass runs the GPU steps keeping a checkpoint every `every`
steps; the backward pass replays each segment from its checkpoint and runs
the exact transpose of the discrete step.  CPU and GPU tensors are both
accepted (GPU tensors are passed to the library by device pointer).
"""
import numpy as np
import torch

from . import Solver, _check, lib
import ctypes


class FinalEnstrophy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u0, nu, dt, nsteps, every):
        n = u0.shape[-1]
        if u0.dtype != torch.float64 or tuple(u0.shape) != (3, n, n, n):
            raise ValueError("u0 must be a float64 tensor of shape [3, n, n, n]")
        solver = Solver(n, nu, dt, "fp64")
        x = u0.detach().contiguous()
        if x.is_cuda:
            torch.cuda.synchronize(x.device)
            z = ctypes.c_double()
            _check(lib().dns_adjoint_forward(solver._ctx, ctypes.c_void_p(x.data_ptr()), int(nsteps), int(every),
                                             ctypes.byref(z)), "dns_adjoint_forward")
            J = z.value
        else:
            J = solver.adjoint_forward(x.numpy(), nsteps, every)
        ctx.solver = solver
        ctx.like = (u0.dtype, u0.device, tuple(u0.shape))
        return u0.new_tensor(J)

    @staticmethod
    def backward(ctx, grad_out):
        dtype, device, shape = ctx.like
        solver = ctx.solver
        seed = float(grad_out.detach().cpu())
        if device.type == "cuda":
            g = torch.empty(shape, dtype=dtype, device=device)
            torch.cuda.synchronize(device)
            _check(lib().dns_adjoint_backward(solver._ctx, seed, ctypes.c_void_p(g.data_ptr())),
                   "dns_adjoint_backward")
        else:
            g = torch.from_numpy(np.ascontiguousarray(solver.adjoint_backward(seed)))
        solver.close()
        ctx.solver = None
        return g, None, None, None, None


def final_enstrophy(u0, nu, dt, nsteps, every=20):
    """Final enstrophy sum (as printed by dns) after nsteps steps from u0."""
    return FinalEnstrophy.apply(u0, nu, dt, nsteps, every)
