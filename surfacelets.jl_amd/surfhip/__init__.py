"""surfhip -- the MI355X (gfx950) analyze/synthesize path of the 3-D
surfacelet transforms, with the reference's API (hakkelt/Surfacelets.jl and
AbstractMultiscaleTransforms) over the C ABI of libsurfhip.so.

Julia's mutating `analyze!`/`synthesize!` are spelled `analyze_`/`synthesize_`.
Arrays are torch tensors on a HIP device, column-major (axis 1 contiguous) as
in the reference; see `surfhip._arrays`.
"""
from ._lib import ArgumentError, DeviceError, DimensionMismatch, lib
from ._arrays import as_colmajor, empty_colmajor, from_numpy, to_numpy
from . import ns as _ns
from .ns import NSSurfacelet, NSSurfaceletWorkspace
from .directional import (hourglass_decompose, hourglass_decompose_, hourglass_mask, hourglass_reconstruct,
                          hourglass_reconstruct_, hourglass_weight, make_hourglass, make_ndfb, ndfb_decompose,
                          ndfb_decompose_, ndfb_mask, ndfb_nsubbands, ndfb_reconstruct, ndfb_reconstruct_,
                          ndfb_sharing, ndfb_subband_index, ndfb_subband_triple, ndfb_weight)


def _is_dec(t):
    from . import decimated as _dec
    return isinstance(t, (_dec.Surfacelet, _dec.SurfaceletWorkspace))


def _mod(t):
    from . import decimated as _dec
    return _dec if _is_dec(t) else _ns


def make_workspace(t, *args, **kw):
    return _mod(t).make_workspace(t, *args, **kw)


def analyze(*args, **kw):
    t = args[1] if (len(args) == 3 and callable(args[0])) else args[0]
    return _mod(t).analyze(*args, **kw)


def analyze_(c, t, x, **kw):
    return _mod(t).analyze_(c, t, x, **kw)


def synthesize(t, c, **kw):
    return _mod(t).synthesize(t, c, **kw)


def synthesize_(*args, **kw):
    t = args[2] if callable(args[0]) else args[1]
    return _mod(t).synthesize_(*args, **kw)


def subband_layout(t, dims):
    return _mod(t).subband_layout(t, dims)


def nsubbands(t, dims):
    return _mod(t).nsubbands(t, dims)


def similar_coefficients(t, x):
    return _mod(t).similar_coefficients(t, x)


def subband(t, c, i):
    return _mod(t).subband(t, c, i)


def signal_size(t, c):
    return _mod(t).signal_size(t, c)


def supports_streaming(t):
    return not _is_dec(t)


def transform(ws):
    return ws.transform


def __getattr__(name):
    if name in ("Surfacelet", "SurfaceletWorkspace", "SubbandVector", "surfacelet_redundancy",
                "checkerboard_orders"):
        from . import decimated as _dec
        return getattr(_dec, name)
    raise AttributeError(name)
