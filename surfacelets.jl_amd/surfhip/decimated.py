"""Surfacelet (decimated): the reference's transform object and its AMT methods,
running on libsurfhip.

Mirrors src/types/transforms.jl:159-315 (the transform),
src/transforms/surfacelet.jl:126-464 (workspace, size rules, layout, analyze!,
synthesize!) and the SubbandVector container of AbstractMultiscaleTransforms
(one flat vector, subbands at cumulative ranges in subband_layout order).
Integer rules (sizes, layout, nearest fitting length) run in the native
library so that every binding shares them.
"""
import ctypes
from collections import namedtuple

import numpy as np
import torch

from . import _lib
from ._arrays import as_colmajor, check_device, ctypes_void, empty_colmajor, is_colmajor, ptr
from ._lib import ArgumentError, DimensionMismatch
from .ns import _check_volume, _jl_dims, _scales_and_levels, _check_scales

SLayoutEntry = namedtuple("SLayoutEntry", ["scale", "hourglass", "wedge"])

CHECKERBOARD_ORDERS = [4, 6, 8, 10, 12, 14, 16, 18]


def checkerboard_orders():
    """src/filters/checkerboard.jl:61."""
    return list(CHECKERBOARD_ORDERS)


def _transverse(K):
    return (1, 2) if K == 0 else ((0, 2) if K == 1 else (0, 1))


def _surf_levels(L):
    """src/types/transforms.jl:263-286 -- levels[K][m] (0-based) of one scale."""
    if isinstance(L, (bool, np.bool_)):
        raise ArgumentError(f"a directional level must be an integer or a 3 × 3 integer matrix, got {L!r}")
    if isinstance(L, (int, np.integer)):
        L = int(L)
        la, lb = (L + 1) // 2, L // 2
        out = []
        for K in range(3):
            a, b = _transverse(K)
            out.append(tuple(la if m == a else (lb if m == b else 0) for m in range(3)))
        return tuple(out), L
    M = np.asarray(L)
    if M.ndim == 2 and np.issubdtype(M.dtype, np.integer):
        if M.shape != (3, 3):
            raise ArgumentError(f"a level matrix must be 3 × 3, got size {_jl_dims(M.shape)}")
        return tuple(tuple(0 if m == K else int(M[K, m]) for m in range(3)) for K in range(3)), M.astype(int)
    raise ArgumentError(f"a directional level must be an integer or a 3 × 3 integer matrix, got {_jl_repr(L)}")


def _jl_repr(v):
    if isinstance(v, np.ndarray):
        rows = ["".join(" ".join(str(int(x)) for x in r)) for r in v]
        return "[" + "; ".join(rows) + "]"
    return repr(v)


def _jl_float(v):
    v = float(v)
    return repr(v) if v != int(v) else f"{v:.1f}"


class Surfacelet:
    """Lu and Do's decimated 3-D surfacelet transform (src/types/transforms.jl:222-261)."""

    def __init__(self, scales=None, direction_levels=None, pyramid_ratio=1.5, checkerboard_order=12,
                 diamond_length=15, kaiser_beta=2.5, diamond_power=4):
        J, levels = _scales_and_levels(scales, direction_levels)
        J = int(J)
        _check_scales(J, levels)
        parsed = [_surf_levels(L) for L in levels]
        lv = [p[0] for p in parsed]
        if not all(all(v >= 0 for row in l for v in row) for l in lv):
            raise ArgumentError("directional levels must be ≥ 0")
        if float(pyramid_ratio) not in (1.0, 1.5, 2.0):
            raise ArgumentError(f"pyramid_ratio must be 1, 1.5 or 2, got {pyramid_ratio}")
        if checkerboard_order not in CHECKERBOARD_ORDERS:
            raise ArgumentError(
                f"checkerboard_order must be one of [{', '.join(map(str, CHECKERBOARD_ORDERS))}], got {checkerboard_order}")
        if not (diamond_length % 2 == 1 and diamond_length >= 3):
            raise ArgumentError(f"diamond_length must be an odd integer ≥ 3, got {diamond_length}")
        if not kaiser_beta > 0:
            raise ArgumentError(f"kaiser_beta must be > 0, got {kaiser_beta}")
        if not (diamond_power % 2 == 0 and diamond_power > 0):
            raise ArgumentError(f"diamond_power must be a positive even integer, got {diamond_power}")
        self.scales = J
        self.direction_levels = [p[1] for p in parsed]
        self.levels = lv
        self.pyramid_ratio = float(pyramid_ratio)
        self.checkerboard_order = int(checkerboard_order)
        self.diamond_length = int(diamond_length)
        self.kaiser_beta = float(kaiser_beta)
        self.diamond_power = int(diamond_power)

    def _key(self):
        dl = tuple(("m", tuple(map(tuple, d.tolist()))) if isinstance(d, np.ndarray) else ("i", d)
                   for d in self.direction_levels)
        return (self.scales, dl, self.pyramid_ratio, self.checkerboard_order, self.diamond_length,
                self.kaiser_beta, self.diamond_power)

    def __eq__(self, o):
        return isinstance(o, Surfacelet) and self._key() == o._key()

    def __hash__(self):
        return hash(("Surfacelet",) + self._key())

    def __repr__(self):
        L = "[" + ", ".join(_jl_repr(d) if isinstance(d, np.ndarray) else str(d) for d in self.direction_levels) + "]"
        return (f"Surfacelet(scales = {self.scales}, direction_levels = {L}, pyramid_ratio = {_jl_float(self.pyramid_ratio)}, "
                f"checkerboard_order = {self.checkerboard_order}, diamond_length = {self.diamond_length}, "
                f"kaiser_beta = {_jl_float(self.kaiser_beta)}, diamond_power = {self.diamond_power})")

    def c_params(self):
        if self.scales > _lib.MAX_SCALES:
            raise ArgumentError(f"at most {_lib.MAX_SCALES} scales are supported, got {self.scales}")
        p = _lib.SParams()
        p.scales = self.scales
        for j, lv in enumerate(self.levels):
            for K in range(3):
                for m in range(3):
                    p.levels[j][K][m] = lv[K][m]
        p.pyramid_ratio = self.pyramid_ratio
        p.checkerboard_order = self.checkerboard_order
        p.diamond_length = self.diamond_length
        p.kaiser_beta = self.kaiser_beta
        p.diamond_power = self.diamond_power
        return p


class SubbandLayout:
    """Sizes of the subbands (layout order), their ranges in the flat vector and
    the signal size (AbstractMultiscaleTransforms.SubbandLayout)."""

    def __init__(self, sizes, signal_size):
        self.sizes = [tuple(int(v) for v in s) for s in sizes]
        self.signal_size = tuple(signal_size)
        self.offsets = []
        off = 0
        for s in self.sizes:
            self.offsets.append(off)
            off += int(np.prod(s))
        self.length = off
        self.ranges = [range(o, o + int(np.prod(s))) for o, s in zip(self.offsets, self.sizes)]

    def __eq__(self, o):
        return isinstance(o, SubbandLayout) and self.sizes == o.sizes and self.signal_size == o.signal_size

    def __hash__(self):
        return hash((tuple(self.sizes), self.signal_size))


class SubbandVector:
    """Coefficients of a decimated Surfacelet: one flat device vector holding the
    subbands at the ranges of `layout`, each column-major."""

    def __init__(self, data, layout):
        if data.dim() != 1 or data.numel() != layout.length:
            raise DimensionMismatch(f"a vector of length {data.numel()} does not hold a layout of length {layout.length}")
        self.data = data
        self.layout = layout

    @property
    def dtype(self):
        return self.data.dtype

    @property
    def device(self):
        return self.data.device

    def subband(self, i):
        """Subband i (1-based) as a column-major view."""
        s = self.layout.sizes[i - 1]
        o = self.layout.offsets[i - 1]
        flat = self.data[o:o + int(np.prod(s))]
        return flat.view(tuple(reversed(s))).permute(2, 1, 0)

    def __len__(self):
        return self.layout.length


def parent(c):
    return c.data


def _sizes(t, dims):
    dims = _check_volume(dims)
    P = t.c_params()
    nsub = ctypes.c_int32()
    ncoef = ctypes.c_int64()
    _lib.call("surfhip_s_counts", ctypes.byref(P), dims[0], dims[1], dims[2], ctypes.byref(nsub), ctypes.byref(ncoef))
    sizes = (ctypes.c_int64 * (3 * nsub.value))()
    offs = (ctypes.c_int64 * nsub.value)()
    _lib.call("surfhip_s_layout", ctypes.byref(P), dims[0], dims[1], dims[2], sizes, offs, nsub.value)
    return [tuple(sizes[3 * i:3 * i + 3]) for i in range(nsub.value)]


def surfacelet_redundancy(t, dims):
    """src/transforms/surfacelet.jl:220-222."""
    dims = _check_volume(dims)
    return sum(int(np.prod(s)) for s in _sizes(t, dims)) / int(np.prod(dims))


def subband_layout(t, dims):
    """src/transforms/surfacelet.jl:301-308."""
    _sizes(t, dims)
    lay = [SLayoutEntry(0, 0, 0)]
    for j in range(1, t.scales + 1):
        for K in range(1, 4):
            for w in range(1, (1 << sum(t.levels[j - 1][K - 1])) + 1):
                lay.append(SLayoutEntry(j, K, w))
    return lay


def nsubbands(t, dims):
    """src/transforms/surfacelet.jl:310."""
    return len(_sizes(t, dims))


def supports_streaming(t):
    return False


def _eltype(x):
    dt = x if isinstance(x, torch.dtype) else x.dtype
    if dt in (torch.float32, torch.complex64):
        return dt
    if dt in (torch.int8, torch.int16, torch.int32, torch.int64, torch.uint8, torch.bool):
        return torch.float32
    raise ArgumentError(f"libsurfhip computes in Float32 and ComplexF32; got element type {dt}")


class SurfaceletWorkspace:
    """A libsurfhip plan of a decimated Surfacelet for one size and element type
    (src/transforms/surfacelet.jl:46-61).  It serves one call at a time."""

    def __init__(self, t, dtype, dims, device, tasks=1):
        self.transform = t
        self.dims = dims
        self.dtype = dtype
        self.device = torch.device(device)
        self.tasks = tasks
        self.layout = SubbandLayout(_sizes(t, dims), dims)
        self._plan = ctypes.c_void_p()
        code = _lib.F32 if dtype == torch.float32 else _lib.C64
        with torch.cuda.device(self.device):
            _lib.call("surfhip_s_plan_create", ctypes.byref(self._plan), ctypes.byref(t.c_params()),
                      dims[0], dims[1], dims[2], code, ctypes_void(torch.cuda.current_stream().cuda_stream))

    def __del__(self):
        plan = getattr(self, "_plan", None)
        if plan is not None and plan.value:
            try:
                _lib.lib().surfhip_s_plan_destroy(plan)
            except Exception:
                pass
            self._plan = None

    def __repr__(self):
        return f"SurfaceletWorkspace{{{self.dtype}}}({'×'.join(map(str, self.dims))}, {self.transform}, tasks = {self.tasks})"

    def bind_stream(self):
        _lib.call("surfhip_s_set_stream", self._plan, ctypes_void(torch.cuda.current_stream(self.device).cuda_stream))


def make_workspace(t, x_or_type, dims=None, *, flags=None, tasks=None, tune=False, prewarm=True, device=None):
    """make_workspace(t, T, dims) or make_workspace(t, x): src/transforms/
    surfacelet.jl:126-199.  `flags` and `tune` are accepted and ignored."""
    if tasks is not None and tasks < 1:
        raise ArgumentError(f"tasks must be ≥ 1, got {tasks}")
    if isinstance(x_or_type, torch.Tensor):
        check_device(x_or_type, "volume")
        dims = _check_volume(x_or_type.shape)
        dt = _eltype(x_or_type)
        device = x_or_type.device
    else:
        dims = _check_volume(dims)
        dt = _eltype(x_or_type)
        device = device or torch.device("cuda", torch.cuda.current_device())
    ws = SurfaceletWorkspace(t, dt, dims, device, tasks or 1)
    if prewarm:
        z = as_colmajor(torch.zeros(dims, dtype=dt, device=device))
        synthesize(ws, analyze(ws, z))
    return ws


def check_workspace(ws, t, eltype, dims):
    if not isinstance(ws, SurfaceletWorkspace):
        raise ArgumentError(f"a Surfacelet needs a SurfaceletWorkspace, got {type(ws).__name__}")
    if ws.transform != t:
        raise ArgumentError(f"the workspace was built for {ws.transform}, not {t}")
    if eltype != ws.dtype:
        raise ArgumentError(f"the workspace is for element type {ws.dtype}, got {eltype}")
    if tuple(dims) != tuple(ws.dims):
        raise DimensionMismatch(f"the workspace is for size {_jl_dims(ws.dims)}, got {_jl_dims(dims)}")
    return ws


def _split(t_or_ws, workspace):
    if isinstance(t_or_ws, SurfaceletWorkspace):
        return t_or_ws.transform, t_or_ws
    return t_or_ws, workspace


def _workspace_for(workspace, t, eltype, dims, device):
    """src/transforms/surfacelet.jl:327-334."""
    if workspace is None:
        return make_workspace(t, eltype, dims, prewarm=False, device=device)
    return check_workspace(workspace, t, eltype, dims)


def similar_coefficients(t, x):
    """src/transforms/surfacelet.jl:439-442."""
    dims = _check_volume(x.shape)
    lay = SubbandLayout(_sizes(t, dims), dims)
    return SubbandVector(torch.empty(lay.length, dtype=_eltype(x), device=x.device), lay)


def subband(t, c, i):
    return c.subband(i)


def signal_size(t, c):
    return c.layout.signal_size


def _check_coefficients(c, ws):
    """src/transforms/surfacelet.jl:421-427, plus what the device needs: the
    vector lives on the GPU (a host pointer must never reach a kernel)."""
    check_device(c.data, "coefficient vector")
    if c.layout != ws.layout:
        raise DimensionMismatch(f"the coefficients' subband layout is not that of {ws.transform} on size {_jl_dims(ws.dims)}")
    if c.dtype != ws.dtype:
        raise ArgumentError(f"coefficients of element type {c.dtype}, the workspace is for {ws.dtype}")
    return c


def analyze_(c, t, x, *, workspace=None, threaded=None):
    """analyze!(c, t, x): src/transforms/surfacelet.jl:446-454."""
    t, workspace = _split(t, workspace)
    check_device(x, "volume")
    ws = _workspace_for(workspace, t, _eltype(x), _check_volume(x.shape), x.device)
    _check_coefficients(c, ws)
    xin = as_colmajor(x if x.dtype == ws.dtype else x.to(ws.dtype))
    ws.bind_stream()
    if c.data.is_contiguous():
        _lib.call("surfhip_s_analyze", ws._plan, ptr(xin), ptr(c.data))
    else:
        # A strided view (e.g. big[::2]) is written through a packed buffer,
        # never as if it were packed.
        tmp = torch.empty(c.data.numel(), dtype=ws.dtype, device=ws.device)
        _lib.call("surfhip_s_analyze", ws._plan, ptr(xin), ptr(tmp))
        c.data.copy_(tmp)
    return c


def analyze(t, x, *, workspace=None, threaded=None):
    t, workspace = _split(t, workspace)
    check_device(x, "volume")
    ws = _workspace_for(workspace, t, _eltype(x), _check_volume(x.shape), x.device)
    c = SubbandVector(torch.empty(ws.layout.length, dtype=ws.dtype, device=x.device), ws.layout)
    return analyze_(c, t, x, workspace=ws, threaded=threaded)


def synthesize_(x, t, c, *, workspace=None, threaded=None):
    """synthesize!(x, t, c): src/transforms/surfacelet.jl:456-464."""
    t, workspace = _split(t, workspace)
    check_device(x, "volume")
    ws = _workspace_for(workspace, t, _eltype(x), _check_volume(x.shape), x.device)
    _check_coefficients(c, ws)
    cd = c.data.contiguous()
    ws.bind_stream()
    if is_colmajor(x) and x.dtype == ws.dtype:
        _lib.call("surfhip_s_synthesize", ws._plan, ptr(cd), ptr(x))
    else:
        y = empty_colmajor(ws.dims, ws.dtype, ws.device)
        _lib.call("surfhip_s_synthesize", ws._plan, ptr(cd), ptr(y))
        x.copy_(y)
    return x


def synthesize(t, c, *, workspace=None, threaded=None):
    t, workspace = _split(t, workspace)
    dims = c.layout.signal_size
    if workspace is not None:
        dims = workspace.dims
    x = empty_colmajor(dims, c.dtype, c.device)
    return synthesize_(x, t, c, workspace=workspace, threaded=threaded)


def adjoint(t, c, *, workspace=None):
    """x = A^T c, the adjoint of analyze (not its inverse: the decimated
    transform is a frame but not a tight one, src/transforms/surfacelet.jl:24-26)."""
    t, workspace = _split(t, workspace)
    dims = workspace.dims if workspace is not None else c.layout.signal_size
    ws = _workspace_for(workspace, t, c.dtype, dims, c.device)
    _check_coefficients(c, ws)
    cd = c.data.contiguous()
    x = empty_colmajor(ws.dims, ws.dtype, ws.device)
    ws.bind_stream()
    _lib.call("surfhip_s_adjoint", ws._plan, ptr(cd), ptr(x))
    return x


INDEX_MAP_KINDS = {"box": 0, "box_low": 1, "split_gather": 2, "mirror": 3, "block_offsets": 4}


def index_map(ws, kind, level, *, axis=0, hourglass=0):
    """An integer index map of the workspace's plan as the device holds or
    computes it (surfhip_s_index_maps; level, axis and hourglass 0-based,
    finest level first): 'box' / 'box_low' the pyramid box bins and their
    lowpass bins along `axis`; 'split_gather' the box coordinates each
    lowpass bin gathers from, shape (n_low, 2), -1 when absent; 'mirror' the
    Hermitian mirror index along `axis` of `hourglass` at scale `level`, -1
    on the direct half; 'block_offsets' the flat offset of each block of
    `hourglass` at scale `level`."""
    k = INDEX_MAP_KINDS[kind]
    n = ctypes.c_int64()
    _lib.call("surfhip_s_index_maps", ws._plan, k, level, hourglass, axis, None, 0, ctypes.byref(n))
    buf = (ctypes.c_int64 * max(1, n.value))()
    _lib.call("surfhip_s_index_maps", ws._plan, k, level, hourglass, axis, buf, n.value, ctypes.byref(n))
    out = np.frombuffer(buf, dtype=np.int64, count=n.value).copy()
    return out.reshape(-1, 2) if k == 2 else out


def check_size(t, dims):
    """Raise the reference's ArgumentError (with the nearest fitting length)
    when `dims` does not fit `t` (src/transforms/surfacelet.jl:253-269)."""
    dims = _check_volume(dims)
    _lib.call("surfhip_s_check_size", ctypes.byref(t.c_params()), dims[0], dims[1], dims[2])
