"""torch.autograd wrapper of the adjoint, for the tests only.

BiquadIR.apply(coeffs, sim, nsteps, ncheckpoints) runs the recording forward
through the C ABI (virs_forward_record) with the given material rows and
returns the receiver IRs as a tensor; its backward runs virs_adjoint with
the incoming gradient and returns dL/dcoeffs.  The simulation itself runs on
the GPU through libvirs_hip.so; torch only carries the values and the chain
rule around it.
"""
import numpy as np
import torch


class BiquadIR(torch.autograd.Function):
    @staticmethod
    def forward(ctx, coeffs, sim, nsteps, ncheckpoints):
        sim.set_materials(coeffs.detach().cpu().numpy().astype(np.float64))
        sim.start(nsteps)
        sim.forward_record(nsteps, ncheckpoints)
        ctx.sim = sim
        return torch.from_numpy(sim.receivers(True)).to(coeffs.dtype)

    @staticmethod
    def backward(ctx, grad_ir):
        g = ctx.sim.adjoint(grad_ir.detach().cpu().numpy().astype(np.float32))
        return torch.from_numpy(g).to(grad_ir.dtype), None, None, None
