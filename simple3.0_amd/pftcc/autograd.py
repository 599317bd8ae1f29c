"""torch.autograd.Function around the batched adjoint (gencorr_grad_for_rot_8 for many pairs).

For test harnesses: inputs and outputs are CPU float64 tensors; the objective and its gradient
come from a backend with the signature of PolarFTCorrCalc.grad_batch
(irefs, iptcls, shvecs, irots, with_grad) -> f or (f, grad[:, 3]).  The forward pass evaluates
the objective only; the backward pass calls the backend again for the gradient (the phase ramps
are recomputed on the device, never stored).  Reference: simple_polarft_corrcalc.f90:1849-1925
(gencorr_grad_for_rot_8), whose objective is k-weighted unless match_filt.
"""
import numpy as np
import torch


class CorrForRot(torch.autograd.Function):
    """f[i] = corr(iref[i], iptcl[i], shift[i], irot[i] + dpsi[i]) for CPU float64 tensors.

    shifts: [N, 2] (pixels); dpsi: [N] in-plane angle offsets in radians (only the derivative at
    dpsi = 0 is available: the objective is evaluated at the grid rotation irot)."""

    @staticmethod
    def forward(ctx, shifts, dpsi, backend, irefs, iptcls, irots):
        if torch.any(dpsi != 0):
            raise ValueError("CorrForRot evaluates at grid rotations only (dpsi must be 0)")
        sh = shifts.detach().cpu().numpy()
        f = backend(irefs, iptcls, sh, irots, False)
        ctx.backend = backend
        ctx.args = (irefs, iptcls, irots)
        ctx.save_for_backward(shifts)
        return torch.from_numpy(np.asarray(f, np.float64))

    @staticmethod
    def backward(ctx, gout):
        (shifts,) = ctx.saved_tensors
        irefs, iptcls, irots = ctx.args
        _, g = ctx.backend(irefs, iptcls, shifts.detach().cpu().numpy(), irots, True)
        g = torch.from_numpy(np.asarray(g, np.float64))
        gs = gout[:, None] * g[:, :2] if ctx.needs_input_grad[0] else None
        gp = gout * g[:, 2] if ctx.needs_input_grad[1] else None
        return gs, gp, None, None, None, None


def corr_for_rot(pftcc, irefs, iptcls, shifts, irots, dpsi=None):
    """Differentiable objective of pairs through a PolarFTCorrCalc (or any grad_batch backend)."""
    backend = pftcc.grad_batch if hasattr(pftcc, "grad_batch") else pftcc
    if dpsi is None:
        dpsi = torch.zeros(shifts.shape[0], dtype=torch.float64)
    return CorrForRot.apply(shifts, dpsi, backend, np.asarray(irefs, np.int32), np.asarray(iptcls, np.int32),
                            np.asarray(irots, np.int32))
