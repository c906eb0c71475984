"""torch.autograd.Function wrappers with the hand-written adjoint kernels.

MassApply:        y = M_s(eps) x with gradients in x (M_s is self-adjoint) and in the
                  permittivity samples, geps = NF * Re(conj(ifftn gy) .* ifftn x)
                  (derivative of exp1/applyMij_fft.m:11 with respect to eps).
LowestEigenvalues: lambda_1..k of the Method-C eigensolve as a function of eps; the
                  backward is the Hellmann-Feynman gradient
                  d lambda/d eps_j = -lambda * NF * sum_c |ifftn(x_c)_j|^2 of the
                  M-normalised recovered eigenvector x (exp2/recover_original_eigenvectors_exp2.m:18-51).
                  The forward keeps only the reduced Ritz vectors (2NF x k, the
                  checkpoint); the backward recovers x from them.
Neither has a reference counterpart: the reference computes no gradients.
"""
import numpy as np
import torch

from .eigs import QSystem, recover_original_eigenvectors, run_method_C
from .mass import MassBackend


class MassApply(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, eps, mass: MassBackend):
        # the backend computes with this eps and gets its previous permittivity back afterwards
        with mass.using_eps(eps.detach()):
            y = mass.MscalarOp(x.detach())
        ctx.mass = mass
        ctx.save_for_backward(x, eps)
        return y

    @staticmethod
    def backward(ctx, gy):
        x, eps = ctx.saved_tensors
        mass = ctx.mass
        gx = None
        if ctx.needs_input_grad[0]:
            with mass.using_eps(eps.detach()):  # M_s of the forward's eps, whatever was set since
                gx = mass.MscalarOp(gy.contiguous())
        geps = None
        if ctx.needs_input_grad[1]:
            geps = mass.adjoint_eps(x.detach().reshape(-1, mass.NF), gy.contiguous().reshape(-1, mass.NF))
        return gx, geps, None


def mass_apply(x, eps, mass):
    return MassApply.apply(x, eps, mass)


class LowestEigenvalues(torch.autograd.Function):
    @staticmethod
    def forward(ctx, eps, mass: MassBackend, q, k, opts):
        with mass.using_eps(eps.detach()):
            sysq = QSystem(mass, q)
            out = run_method_C(sysq, k=k, want_vectors=True, **opts)
        ctx.mass = mass
        ctx.sysq = sysq
        ctx.lam = out["lambda_"].copy()
        ctx.Y = out["Y"]  # checkpoint: reduced Ritz vectors
        ctx.save_for_backward(eps)
        return torch.as_tensor(out["lambda_"], dtype=torch.float64, device=eps.device)

    @staticmethod
    def backward(ctx, glam):
        from . import _lib
        import ctypes as C
        (eps,) = ctx.saved_tensors
        mass, sysq = ctx.mass, ctx.sysq
        g = torch.zeros(mass.NF, dtype=torch.float64, device=mass.device)
        gl = np.ascontiguousarray(glam.detach().cpu().numpy().astype(np.float64))
        lam = np.ascontiguousarray(ctx.lam.astype(np.float64))
        # the recovery X = M^{-1} V Y and the gradient use the forward's eps
        with mass.using_eps(eps.detach()):
            X, MX, _ = recover_original_eigenvectors(ctx.Y, sysq)
            _lib.call("fame6d_eig_grad_eps", mass.handle, _lib.ptr(X), X.shape[0], lam.ctypes.data_as(C.c_void_p),
                      gl.ctypes.data_as(C.c_void_p), _lib.ptr(g))
        sysq.close()
        return g.to(glam.device) if glam.device != g.device else g, None, None, None, None


def lowest_eigenvalues(eps, mass, q, k=5, **opts):
    """lambda_1..k (ascending) differentiable with respect to eps (NF float64 tensor)."""
    defaults = dict(p=40, tol=1e-10, max_restart=50, tol_cg=1e-10, maxit_cg=3000, seed=20260905)
    defaults.update(opts)
    return LowestEigenvalues.apply(eps, mass, q, k, defaults)
