"""The directional filter banks on their own: the hourglass bank and the
directional (NDFB) bank, on the NS kernels.

The reference builds them as public blocks (src/directional/hourglass.jl:
149-496, src/directional/ndfb.jl:91-468).  Here a bank is a libsurfhip plan
of one NSSurfacelet scale without the pyramid (surfhip_ndfb_plan_create):
the group factor is the hourglass weight W_i alone and the lowpass is zero,
so plan subband 1 is zero and plan subbands 2..S are the bank's 3 * 2^levels
directional subbands in ndfb_subband_index order.  Decomposition is the
plan's analysis and reconstruction its synthesis, which is a plain sum when
the bank is not tight (the masks are then a partition of unity and the
synthesis masks are 1), as in the reference.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from ._arrays import as_colmajor, check_device, ctypes_void, empty_colmajor, is_colmajor, ptr
from ._lib import ArgumentError, DimensionMismatch
from .ns import _eltype


class NDFBParams(ctypes.Structure):
    _fields_ = [
        ("levels", ctypes.c_int32),
        ("bernstein_order", ctypes.c_int32),
        ("halfband_order", ctypes.c_int32),
        ("transition", ctypes.c_double),
        ("tight", ctypes.c_int32),
    ]


def _wedge_levels(L):
    """src/directional/ndfb_wedges.jl:19: (cld(L, 2), fld(L, 2))."""
    return (L + 1) // 2, L // 2


class NDFB:
    """make_ndfb(T, dims; levels, bernstein_order, halfband_order, tight,
    transition) (src/directional/ndfb.jl:91-104) on the device of `device`.
    One call at a time."""

    def __init__(self, dtype, dims, *, levels=2, bernstein_order=12, halfband_order=1, tight=True,
                 transition=1.0, device=None):
        if levels < 0:
            raise ArgumentError("levels must be ≥ 0")
        if not transition >= 0:
            raise ArgumentError("transition must be ≥ 0")
        if bernstein_order < 1:
            raise ArgumentError("Bernstein order must be ≥ 1")
        dims = tuple(int(d) for d in dims)
        if len(dims) != 3 or any(d < 1 for d in dims):
            raise ArgumentError("volume dimensions must be positive")
        self.dims = dims
        self.levels = int(levels)
        self.bernstein_order = int(bernstein_order)
        self.halfband_order = int(halfband_order)
        self.tight = bool(tight)
        self.transition = float(transition)
        self.dtype = _eltype(dtype)
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.nsub = 3 * (1 << self.levels)
        self._plan = ctypes.c_void_p()
        code = _lib.F32 if self.dtype == torch.float32 else _lib.C64
        with torch.cuda.device(self.device):
            _lib.call("surfhip_ndfb_plan_create", ctypes.byref(self._plan), ctypes.byref(self.c_params()),
                      dims[0], dims[1], dims[2], code, ctypes_void(torch.cuda.current_stream().cuda_stream))
        self._band = None

    def c_params(self, tight=None):
        p = NDFBParams()
        p.levels = self.levels
        p.bernstein_order = self.bernstein_order
        p.halfband_order = self.halfband_order
        p.transition = self.transition
        p.tight = int(self.tight if tight is None else tight)
        return p

    def __del__(self):
        plan = getattr(self, "_plan", None)
        if plan is not None and plan.value:
            try:
                _lib.lib().surfhip_ns_plan_destroy(plan)
            except Exception:
                pass
            self._plan = None

    def __repr__(self):
        return (f"NDFB{{{self.dtype}}}({'×'.join(map(str, self.dims))}; levels = {self.levels}, "
                f"bernstein_order = {self.bernstein_order}, tight = {str(self.tight).lower()})")

    def _bind(self):
        _lib.call("surfhip_ns_set_stream", self._plan,
                  ctypes_void(torch.cuda.current_stream(self.device).cuda_stream))

    def _input(self, volume):
        check_device(volume, "volume")
        if tuple(volume.shape) != self.dims:
            raise DimensionMismatch("volume size does not match the plan")
        if volume.is_complex() and self.dtype != torch.complex64:
            raise ArgumentError("a complex volume needs a bank built for complex data")
        return as_colmajor(volume if volume.dtype == self.dtype else volume.to(self.dtype))

    def _check_band(self, band, k):
        if tuple(band.shape) != self.dims:
            raise DimensionMismatch(f"subband {k} has size {tuple(band.shape)}, but the plan is {self.dims}")
        return band


def make_ndfb(dtype, dims, *, levels=2, bernstein_order=12, halfband_order=1, tight=True, transition=1.0,
              device=None):
    """src/directional/ndfb.jl:91-104,129-131."""
    return NDFB(dtype, dims, levels=levels, bernstein_order=bernstein_order, halfband_order=halfband_order,
                tight=tight, transition=transition, device=device)


def ndfb_sharing(fb, *, levels=2, halfband_order=None, transition=1.0):
    """src/directional/ndfb.jl:118-127: a directional bank on an existing
    hourglass bank's size and hourglass parameters."""
    return NDFB(fb.dtype, fb.dims, levels=levels, bernstein_order=fb.bernstein_order,
                halfband_order=fb.halfband_order if halfband_order is None else halfband_order,
                tight=fb.tight, transition=transition, device=fb.device)


def ndfb_nsubbands(fb):
    """src/directional/ndfb.jl:159: 3 * 2^levels."""
    return fb.nsub


def _counts(fb):
    la, lb = _wedge_levels(fb.levels)
    return 1 << la, 1 << lb


def ndfb_subband_index(fb, hourglass, p, q):
    """src/directional/ndfb.jl:184-190."""
    na, nb = _counts(fb)
    if not 1 <= hourglass <= 3:
        raise ArgumentError("hourglass index must be 1, 2 or 3")
    if not 1 <= p <= na:
        raise ArgumentError(f"wedge index p out of range 1:{na}")
    if not 1 <= q <= nb:
        raise ArgumentError(f"wedge index q out of range 1:{nb}")
    return ((hourglass - 1) * na + (p - 1)) * nb + q


def ndfb_subband_triple(fb, idx):
    """src/directional/ndfb.jl:199-203."""
    na, nb = _counts(fb)
    if not 1 <= idx <= 3 * na * nb:
        raise ArgumentError("subband index out of range")
    k = idx - 1
    return k // (na * nb) + 1, (k // nb) % na + 1, k % nb + 1


def _mask_host(fb, idx, tight):
    n1, n2, n3 = fb.dims
    h = (n1 // 2 + 1, n2, n3)
    out = np.empty(int(np.prod(h)), dtype=np.float32)
    _lib.call("surfhip_ndfb_mask_host", ctypes.byref(fb.c_params(tight)), n1, n2, n3, idx,
              out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    return out.reshape(h, order="F")


def ndfb_mask(fb, idx):
    """src/directional/ndfb.jl:215-220: the analysis mask of subband idx on the
    rfft grid, as the device evaluates it (numpy, (n1 // 2 + 1, n2, n3))."""
    ndfb_subband_triple(fb, idx)
    return _mask_host(fb, idx, fb.tight)


def ndfb_weight(fb, idx):
    """src/directional/ndfb.jl:230-241: the weight before the tight
    normalisation; the weights sum to 1."""
    ndfb_subband_triple(fb, idx)
    return _mask_host(fb, idx, False)


def ndfb_decompose(*args):
    """ndfb_decompose(volume, fb) -> list of 3 * 2^levels subbands, or the
    streaming form ndfb_decompose(f, volume, fb): f(idx, band) for idx in
    order, `band` valid until f returns (src/directional/ndfb.jl:263-346)."""
    if len(args) == 3 and callable(args[0]):
        f, volume, fb = args
        x = fb._input(volume)
        fb._bind()
        if fb._band is None:
            fb._band = empty_colmajor(fb.dims, fb.dtype, fb.device)
        _lib.call("surfhip_ns_stream_begin", fb._plan, ptr(x))
        for idx in range(1, fb.nsub + 1):
            _lib.call("surfhip_ns_stream_subband", fb._plan, idx + 1, ptr(fb._band))
            f(idx, fb._band)
        return None
    volume, fb = args
    x = fb._input(volume)
    c = empty_colmajor(fb.dims + (fb.nsub + 1,), fb.dtype, fb.device)
    fb._bind()
    _lib.call("surfhip_ns_analyze", fb._plan, ptr(x), ptr(c))
    return [c[:, :, :, k] for k in range(1, fb.nsub + 1)]


def ndfb_decompose_(dst, volume, fb):
    """src/directional/ndfb.jl:276-290."""
    if len(dst) != fb.nsub:
        raise DimensionMismatch(f"dst must hold {fb.nsub} subbands")
    for k in range(fb.nsub):
        fb._check_band(dst[k], k + 1)

    def put(idx, band):
        dst[idx - 1].copy_(band)

    ndfb_decompose(put, volume, fb)
    return dst


def ndfb_reconstruct_(*args):
    """ndfb_reconstruct!(volume, subbands, fb) or the streaming form
    ndfb_reconstruct!(f, volume, fb) with f(k) returning subband k, consumed
    in index order (src/directional/ndfb.jl:394-425)."""
    if callable(args[0]) and not isinstance(args[0], torch.Tensor):
        f, volume, fb = args
    else:
        volume, subbands, fb = args
        if len(subbands) != fb.nsub:
            raise DimensionMismatch(f"expected {fb.nsub} subbands")
        f = lambda k: subbands[k - 1]  # noqa: E731
    check_device(volume, "volume")
    if tuple(volume.shape) != fb.dims:
        raise DimensionMismatch("volume size does not match the plan")
    fb._bind()
    _lib.call("surfhip_ns_stream_inverse_begin", fb._plan)
    for k in range(1, fb.nsub + 1):
        b = fb._input(fb._check_band(f(k), k))
        _lib.call("surfhip_ns_stream_accumulate", fb._plan, k + 1, ptr(b))
    if is_colmajor(volume) and volume.dtype == fb.dtype:
        _lib.call("surfhip_ns_stream_finish", fb._plan, ptr(volume))
    else:
        y = empty_colmajor(fb.dims, fb.dtype, fb.device)
        _lib.call("surfhip_ns_stream_finish", fb._plan, ptr(y))
        volume.copy_(y)
    return volume


def ndfb_reconstruct(subbands, fb):
    """src/directional/ndfb.jl:384-387."""
    if len(subbands) != fb.nsub:
        raise DimensionMismatch(f"expected {fb.nsub} subbands")
    out = empty_colmajor(fb.dims, fb.dtype, fb.device)
    return ndfb_reconstruct_(out, list(subbands), fb)


# ------------------------------------------------------------ hourglass bank

def make_hourglass(dtype, dims, *, bernstein_order=12, halfband_order=1, tight=True, device=None):
    """src/directional/hourglass.jl:149-170,191-198: the three-subband bank
    (the directional bank with levels = 0)."""
    return NDFB(dtype, dims, levels=0, bernstein_order=bernstein_order, halfband_order=halfband_order,
                tight=tight, device=device)


def _hourglass_index(fb, i):
    if fb.levels != 0:
        raise ArgumentError("an hourglass bank has levels = 0")
    if not 1 <= i <= 3:
        raise ArgumentError("hourglass subband index must be 1, 2 or 3")
    return i


def hourglass_mask(fb, i):
    """src/directional/hourglass.jl:264-270."""
    return _mask_host(fb, _hourglass_index(fb, i), fb.tight)


def hourglass_weight(fb, i):
    """src/directional/hourglass.jl:279-285."""
    return _mask_host(fb, _hourglass_index(fb, i), False)


def hourglass_decompose(volume, fb):
    """src/directional/hourglass.jl:330-333: (s1, s2, s3)."""
    _hourglass_index(fb, 1)
    return tuple(ndfb_decompose(volume, fb))


def hourglass_decompose_(dst, volume, fb):
    """src/directional/hourglass.jl:343-398."""
    _hourglass_index(fb, 1)
    return tuple(ndfb_decompose_(list(dst), volume, fb))


def hourglass_reconstruct(subbands, fb):
    """src/directional/hourglass.jl:435-438."""
    _hourglass_index(fb, 1)
    return ndfb_reconstruct(subbands, fb)


def hourglass_reconstruct_(volume, subbands, fb):
    """src/directional/hourglass.jl:446-459."""
    _hourglass_index(fb, 1)
    return ndfb_reconstruct_(volume, list(subbands), fb)
