"""torch.autograd.Function over the device chain and its hand-written adjoint.

    rho = cldos(tre, tim, mu, Vre, Vim, engine, hops, omegas, eta)

tre / tim are the real and imaginary parts of the hopping amplitudes in the
order of `hops` (a synth.HOP_DTYPE array giving R and the orbital pair of each
term); mu is the chemical potential (the `fermi` of the input file); Vre / Vim
the impurity potential.  The forward runs the whole FP64 chain on the GPU
(keeping the per-energy G0(R) windows, never G0(k)); backward calls the
library's adjoint, which rebuilds G0(k) per energy from the cached eigen data.
Maps and dL/drho stay in device memory.  The returned maps have shape
(E, W*W) in the library's x + y*W order.
"""
from __future__ import annotations

import numpy as np
import torch

from . import synth


class CldosFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, tre, tim, mu, Vre, Vim, engine, hops, omegas, eta):
        h = np.array(hops, dtype=synth.HOP_DTYPE, copy=True)
        h["tre"] = tre.detach().cpu().double().numpy()
        h["tim"] = tim.detach().cpu().double().numpy()
        engine.set_model(h, float(mu.detach().cpu()), False)
        V = Vre.detach().cpu().double().numpy() + 1j * Vim.detach().cpu().double().numpy()
        engine.set_impurity(V)
        engine.set_keep_windows(True)
        engine.run(np.asarray(omegas, np.float64), float(eta), host=False)
        rho = engine.ldos_tensor().clone()  # the library buffer is reused by the next run
        ctx.engine = engine
        ctx.device = tre.device
        return rho.to(tre.device)

    @staticmethod
    def backward(ctx, grad_rho):
        g = grad_rho.detach().to(device="cuda", dtype=torch.float64).contiguous()
        torch.cuda.synchronize()
        gh, gmu, gV = ctx.engine.gradient_device(g.data_ptr())
        dev = ctx.device
        return (torch.from_numpy(np.ascontiguousarray(gh[:, 0])).to(dev),
                torch.from_numpy(np.ascontiguousarray(gh[:, 1])).to(dev),
                torch.tensor(gmu, dtype=torch.float64, device=dev),
                torch.from_numpy(np.ascontiguousarray(gV.real)).to(dev),
                torch.from_numpy(np.ascontiguousarray(gV.imag)).to(dev),
                None, None, None, None)


def cldos(tre, tim, mu, Vre, Vim, engine, hops, omegas, eta):
    return CldosFunction.apply(tre, tim, mu, Vre, Vim, engine, hops, omegas, eta)
