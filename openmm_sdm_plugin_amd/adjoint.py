"""Parameter gradients of the SDM alchemical potential for force-field fitting.

PerturbationEnergy is a torch.autograd.Function mapping the selected atoms' nonbonded
parameters (q, sigma, epsilon) to u_f = U(x_f + d) - U(x_f) for every frame f.  Its
forward and backward are hand-written HIP kernels (sdm_du_dparams_frames): the backward
recomputes every pair term from the frame coordinates instead of storing them.  The
soft core u -> u' and the bias W(u') (LangevinIntegratorSDM.cpp:125-149,
ReferenceSDMKernels.cpp:264-282) are differentiable torch expressions on top, so
dW/d(alpha, u0, w0) and the chain W'(u') du'/du du/dtheta come from autograd.  The
reference computes no parameter gradients (its dW/dalpha is hard-coded 0,
ReferenceSDMKernels.cpp:294); this module has no counterpart there.
"""
from __future__ import annotations

import numpy as np
import torch
import torch.nn.functional as Fn

from . import _native
from .systems import SDMParams, SDMSystem


def softcore(u, method, umax, acore, ubcore):
    """u' = SoftCoreF(u) as a torch expression (LangevinIntegratorSDM.cpp:125-149)."""
    if method == 0:
        return u
    if method == 1:
        up = umax * torch.tanh((u - ubcore) / umax) + ubcore
    elif method == 2:
        gu = (u - ubcore) / (acore * (umax - ubcore))
        zeta = 1.0 + 2.0 * gu * (gu + 1.0)
        zp = zeta ** acore
        up = (umax - ubcore) * (zp - 1.0) / (zp + 1.0) + ubcore
    else:
        raise ValueError("Unknown soft core method")
    return torch.where(u > ubcore, up, u)


def bias(up, p: SDMParams, alpha=None, u0=None, w0=None):
    """W(u') for the three bias methods (ReferenceSDMKernels.cpp:269-282), with a stable
    softplus for ln(1 + e^z)."""
    alpha = p.alpha if alpha is None else alpha
    u0 = p.u0 if u0 is None else u0
    w0 = p.w0coeff if w0 is None else w0
    if p.bias_method == 1:
        return 0.5 * p.gamma * up * up + p.wbcoeff * up + w0
    if p.bias_method == 2:
        z = -alpha * (up - u0)
        w = p.lambda2 * up + w0
        a_pos = alpha > 0 if torch.is_tensor(alpha) else alpha > 0
        if torch.is_tensor(a_pos):
            return w + torch.where(a_pos, (p.lambda2 - p.lambda1) / alpha * Fn.softplus(z),
                                   torch.zeros_like(w))
        return w + ((p.lambda2 - p.lambda1) / alpha * Fn.softplus(z) if a_pos else 0.0)
    return p.lambda_ * up


class AdjointModel:
    """The SDM system and the atoms whose parameters are fitted."""

    def __init__(self, system: SDMSystem, sel, params: SDMParams):
        self.system = system
        self.sel = np.asarray(sel, dtype=np.int32)
        self.params = params
        self.native = _native.NativeContext(system, 1)
        self.q = np.asarray(system.charge, dtype=np.float64).copy()
        self.sig = np.asarray(system.sigma, dtype=np.float64).copy()
        self.eps = np.asarray(system.epsilon, dtype=np.float64).copy()

    def upload(self, q_sel, sig_sel, eps_sel):
        self.q[self.sel] = q_sel.detach().double().cpu().numpy()
        self.sig[self.sel] = sig_sel.detach().double().cpu().numpy()
        self.eps[self.sel] = eps_sel.detach().double().cpu().numpy()
        self.native.set_nonbonded_params(self.q, self.sig, self.eps)

    def mean_W_and_u(self, q_sel, sig_sel, eps_sel, alpha, u0, w0, frames):
        """(mean_f W(u'(u_f)), mean_f u_f) for frames [F, N, 3] (float64, on the GPU)."""
        u = PerturbationEnergy.apply(q_sel, sig_sel, eps_sel, frames, self)
        p = self.params
        up = softcore(u, p.softcore_method, p.umax, p.acore, p.ubcore)
        W = bias(up, p, alpha, u0, w0)
        return W.mean(), u.mean()


class PerturbationEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, q_sel, sig_sel, eps_sel, frames, model: AdjointModel):
        if frames.dtype != torch.float64 or not frames.is_cuda or not frames.is_contiguous():
            raise ValueError("frames must be a contiguous float64 CUDA tensor [F, N, 3]")
        F = frames.shape[0]
        model.upload(q_sel, sig_sel, eps_sel)
        u = torch.empty(F, dtype=torch.float64, device=frames.device)
        torch.cuda.synchronize()
        model.native.du_dparams_frames(frames.data_ptr(), F, model.sel, 0, u.data_ptr())
        ctx.model = model
        ctx.save_for_backward(q_sel, sig_sel, eps_sel, frames)
        return u

    @staticmethod
    def backward(ctx, gu):
        q_sel, sig_sel, eps_sel, frames = ctx.saved_tensors
        model = ctx.model
        F = frames.shape[0]
        nsel = len(model.sel)
        model.upload(q_sel, sig_sel, eps_sel)
        grad = torch.empty(F, nsel, 3, dtype=torch.float64, device=frames.device)
        torch.cuda.synchronize()
        model.native.du_dparams_frames(frames.data_ptr(), F, model.sel, grad.data_ptr(), 0)
        g = (gu.to(torch.float64)[:, None, None] * grad).sum(0)
        return (g[:, 0].to(q_sel.dtype), g[:, 1].to(sig_sel.dtype), g[:, 2].to(eps_sel.dtype),
                None, None)
