"""torch.autograd interface of the direct-RPA energy.

    E = brpa.autograd.ecrpa(R, eocc, evir, freqs, weights)   # 0-d float64 tensor
    E.backward()                                              # R.grad, eocc.grad, evir.grad

The forward pass is the fused GPU E_RPA (FP64 or split-BF16); the backward
pass is the hand-written FP64 adjoint of libbrpa (brpa_ecrpa_grad), which
recomputes Pi(u) per frequency instead of keeping it from the forward pass.
The frequency grid (freqs, weights) is a constant: in the reference it comes
from a histogram scan that is piecewise constant in the orbital energies.
"""
import numpy as np
import torch

from . import Context

_default_ctx = None


def _context():
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(torch.cuda.current_device())
    return _default_ctx


def _np(x):
    if isinstance(x, torch.Tensor):
        return x.detach().to("cpu", torch.float64).numpy()
    return np.asarray(x, np.float64)


class _ERPA(torch.autograd.Function):
    @staticmethod
    def forward(fctx, R, eocc, evir, freqs, weights, bctx, precision):
        Rc = R.detach().contiguous()
        e, _ = bctx.ecrpa(Rc, _np(eocc), _np(evir), _np(freqs), _np(weights), precision)
        fctx.save_for_backward(R, eocc, evir)
        fctx.bctx = bctx
        fctx.grid = (_np(freqs), _np(weights))
        return torch.tensor(e, dtype=torch.float64, device=R.device)

    @staticmethod
    def backward(fctx, gE):
        R, eocc, evir = fctx.saved_tensors
        dR, deo, dev = fctx.bctx.ecrpa_grad(R.detach().contiguous(), _np(eocc), _np(evir),
                                            *fctx.grid)
        g = gE.to(torch.float64)
        gR = g * dR if fctx.needs_input_grad[0] else None
        go = (g.to(eocc.device) * torch.from_numpy(deo).to(eocc.device)) if fctx.needs_input_grad[1] else None
        gv = (g.to(evir.device) * torch.from_numpy(dev).to(evir.device)) if fctx.needs_input_grad[2] else None
        return gR, go, gv, None, None, None, None


def ecrpa(R, eocc, evir, freqs, weights, context=None, precision="fp64"):
    """Differentiable E_RPA of one fragment (R: CUDA float64 tensor (ov, N))."""
    if not (isinstance(eocc, torch.Tensor) and isinstance(evir, torch.Tensor)):
        raise TypeError("eocc and evir must be torch tensors")
    return _ERPA.apply(R, eocc, evir, freqs, weights, context or _context(), precision)
