"""Reverse-mode rules (torch.autograd) for analyze and synthesize.

The backward of each transform is its hand-written adjoint in libsurfhip:
  NSSurfacelet analyze    -> surfhip_ns_adjoint (synthesize for a tight bank)
  NSSurfacelet synthesize -> surfhip_ns_synthesize_adjoint (analyze if tight)
  Surfacelet analyze      -> surfhip_s_adjoint (the decimated transform is
                             not tight, so this is not synthesize)
These are ChainRulesCore-style rrules (the reverse-mode rule a Julia caller
would register), expressed for torch tensors.  The transforms are real-linear and act on complex data as two real
channels, so the same adjoint serves complex inputs.
"""
import torch

from . import decimated as _dec
from . import ns as _ns
from ._arrays import as_colmajor


class _NSAnalyze(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, ws):
        ctx.ws = ws
        return _ns.analyze(ws, x.detach())

    @staticmethod
    def backward(ctx, gc):
        return _ns.adjoint(ctx.ws, as_colmajor(gc)), None


class _NSSynthesize(torch.autograd.Function):
    @staticmethod
    def forward(ctx, c, ws):
        ctx.ws = ws
        return _ns.synthesize(ws, as_colmajor(c.detach()))

    @staticmethod
    def backward(ctx, gx):
        return _ns.synthesize_adjoint(ctx.ws, as_colmajor(gx)), None


class _DecAnalyze(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, ws):
        ctx.ws = ws
        return _dec.analyze(ws, x.detach()).data

    @staticmethod
    def backward(ctx, gflat):
        c = _dec.SubbandVector(gflat.contiguous(), ctx.ws.layout)
        return _dec.adjoint(ctx.ws.transform, c, workspace=ctx.ws), None


def analyze(ws, x):
    """Differentiable analyze with a prebuilt workspace.  NS: returns the
    (n1, n2, n3, S) coefficients; decimated: the flat coefficient vector."""
    if isinstance(ws, _ns.NSSurfaceletWorkspace):
        return _NSAnalyze.apply(x, ws)
    if isinstance(ws, _dec.SurfaceletWorkspace):
        return _DecAnalyze.apply(x, ws)
    raise TypeError("analyze needs a workspace of NSSurfacelet or Surfacelet")


def synthesize(ws, c):
    """Differentiable NS synthesize with a prebuilt workspace."""
    if isinstance(ws, _ns.NSSurfaceletWorkspace):
        return _NSSynthesize.apply(c, ws)
    raise TypeError("a differentiable synthesize is provided for NSSurfacelet")
