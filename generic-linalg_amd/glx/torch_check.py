"""torch.autograd wrapper of the device gradient -- for checking only.

S(theta, m) = phi^dag (D^dag D)^-1 phi for the U(1) staggered operator
(operators.cpp:444-453) with links U = exp(i theta).  forward() evaluates the
estimator 2 Re phi^dag chi - |D chi|^2 on the device (exact S up to second
order in the solve error), backward() returns the device gradient
(glx_dev_staggered_grad): dS/dtheta for every link and dS/dm.  The solver
library is the product; this wrapper only lets tests compare it with
autograd-style directional derivatives.
"""
from __future__ import annotations

import numpy as np
import torch

from . import DeviceArray, Lattice, staggered_grad


class StaggeredAction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, mass, phi, lx: int, ly: int, precision: int = 0, eps: float = 1e-12):
        links = np.exp(1j * theta.detach().cpu().numpy().astype(np.float64))
        lat = Lattice(links, lx, ly)
        dphi = DeviceArray(np.ascontiguousarray(phi.detach().cpu().numpy(), dtype=np.complex128))
        g = staggered_grad(lat, float(mass), dphi, precision=precision, eps=eps)
        ctx.save_for_backward(torch.from_numpy(g.grad).to(device=theta.device, dtype=theta.dtype),
                              torch.tensor(g.dS_dm, dtype=theta.dtype, device=theta.device))
        return torch.tensor(g.S_robust, dtype=theta.dtype, device=theta.device)

    @staticmethod
    def backward(ctx, gout):
        g, gm = ctx.saved_tensors
        return gout * g, gout * gm, None, None, None, None, None
