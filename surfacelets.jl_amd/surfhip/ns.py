"""NSSurfacelet: the reference's transform object and its AMT methods, running
on libsurfhip.

Mirrors src/types/transforms.jl:86-157 (the transform) and
src/transforms/ns_surfacelet.jl:88-173,504-641 (workspace, layout, analyze!,
synthesize!, streaming) with the same argument meaning and error texts.
Julia's `analyze!`/`synthesize!` are `analyze_`/`synthesize_` here.
"""
import ctypes
from collections import namedtuple

import torch

from . import _lib
from ._arrays import (as_colmajor, check_device, ctypes_void, data_dtype, empty_colmajor, is_colmajor, ptr)
from ._lib import ArgumentError, DimensionMismatch

NSLayoutEntry = namedtuple("NSLayoutEntry", ["scale", "direction"])


def _parabolic_levels(scales):
    """src/types/transforms.jl:147."""
    return [1 + (j - 1) // 2 for j in range(1, scales + 1)]


def _scales_and_levels(scales, direction_levels):
    """src/types/transforms.jl:136-142."""
    if direction_levels is None:
        if scales is None:
            raise ArgumentError("give `scales`, `direction_levels` or both")
        return scales, _parabolic_levels(scales)
    return (len(direction_levels) if scales is None else scales), list(direction_levels)


def _check_scales(scales, direction_levels):
    """src/types/transforms.jl:149-154."""
    if scales < 1:
        raise ArgumentError(f"scales must be ≥ 1, got {scales}")
    if len(direction_levels) != scales:
        raise ArgumentError(f"length(direction_levels) must equal scales = {scales}, got {len(direction_levels)}")


class NSSurfacelet:
    """The nonsubsampled surfacelet transform (src/types/transforms.jl:86-113)."""

    def __init__(self, scales=None, direction_levels=None, bernstein_order=12, halfband_order=1,
                 transition=1.0, tight=True):
        J, levels = _scales_and_levels(scales, direction_levels)
        J = int(J)
        _check_scales(J, levels)
        levels = [int(v) for v in levels]
        if not all(v >= 0 for v in levels):
            raise ArgumentError(f"direction_levels entries must be ≥ 0, got [{', '.join(map(str, levels))}]")
        if bernstein_order < 1:
            raise ArgumentError(f"bernstein_order must be ≥ 1, got {bernstein_order}")
        if halfband_order < 0:
            raise ArgumentError(f"halfband_order must be ≥ 0, got {halfband_order}")
        if not transition >= 0:
            raise ArgumentError(f"transition must be ≥ 0, got {transition}")
        self.scales = J
        self.direction_levels = levels
        self.bernstein_order = int(bernstein_order)
        self.halfband_order = int(halfband_order)
        self.transition = float(transition) + 0.0
        self.tight = bool(tight)

    def _key(self):
        return (self.scales, tuple(self.direction_levels), self.bernstein_order, self.halfband_order,
                self.transition, self.tight)

    def __eq__(self, o):
        return isinstance(o, NSSurfacelet) and self._key() == o._key()

    def __hash__(self):
        return hash(("NSSurfacelet",) + self._key())

    def __repr__(self):
        return (f"NSSurfacelet(scales = {self.scales}, direction_levels = [{', '.join(map(str, self.direction_levels))}], "
                f"bernstein_order = {self.bernstein_order}, halfband_order = {self.halfband_order}, "
                f"transition = {self.transition}, tight = {str(self.tight).lower()})")

    def c_params(self):
        p = _lib.NSParams()
        p.scales = self.scales
        if self.scales > _lib.MAX_SCALES:
            raise ArgumentError(f"at most {_lib.MAX_SCALES} scales are supported, got {self.scales}")
        for j, v in enumerate(self.direction_levels):
            p.direction_levels[j] = v
        p.bernstein_order = self.bernstein_order
        p.halfband_order = self.halfband_order
        p.transition = self.transition
        p.tight = 1 if self.tight else 0
        return p


def _check_volume(dims):
    """src/transforms/ns_surfacelet.jl:506-507."""
    dims = tuple(int(d) for d in dims)
    if len(dims) != 3:
        raise ArgumentError(f"the surfacelet transforms take volumes (three axes), got size {_jl_dims(dims)}")
    return dims


def _jl_dims(dims):
    return "(" + ", ".join(map(str, dims)) + ("," if len(dims) == 1 else "") + ")"


def _nbands(t, j):
    return 3 * (1 << t.direction_levels[j - 1])


def nsubbands(t, dims):
    """src/transforms/ns_surfacelet.jl:527."""
    _check_volume(dims)
    return 1 + sum(_nbands(t, j) for j in range(1, t.scales + 1))


def subband_layout(t, dims):
    """src/transforms/ns_surfacelet.jl:518-525."""
    _check_volume(dims)
    lay = [NSLayoutEntry(0, 0)]
    for j in range(1, t.scales + 1):
        lay += [NSLayoutEntry(j, k) for k in range(1, _nbands(t, j) + 1)]
    return lay


def supports_streaming(t):
    return True


class NSSurfaceletWorkspace:
    """A libsurfhip plan: the tables on the device and the transform's buffers
    for one size and element type (src/transforms/ns_surfacelet.jl:55-76).
    It serves one call at a time."""

    def __init__(self, t, dtype, dims, device):
        self.transform = t
        self.dims = dims
        self.dtype = dtype
        self.device = torch.device(device)
        self._plan = ctypes.c_void_p()
        code = _lib.F32 if dtype == torch.float32 else _lib.C64
        with torch.cuda.device(self.device):
            _lib.call("surfhip_ns_plan_create", ctypes.byref(self._plan), ctypes.byref(t.c_params()),
                      dims[0], dims[1], dims[2], code, ctypes_void(torch.cuda.current_stream().cuda_stream))
        self.nsub = nsubbands(t, dims)
        self._band = None

    def __del__(self):
        plan = getattr(self, "_plan", None)
        if plan is not None and plan.value:
            try:
                _lib.lib().surfhip_ns_plan_destroy(plan)
            except Exception:
                pass
            self._plan = None

    def __repr__(self):
        return f"NSSurfaceletWorkspace{{{self.dtype}}}({'×'.join(map(str, self.dims))}, {self.transform}, tasks = 1)"

    @property
    def handle(self):
        return self._plan

    def bind_stream(self):
        _lib.call("surfhip_ns_set_stream", self._plan, ctypes_void(torch.cuda.current_stream(self.device).cuda_stream))

    def tune(self):
        """Run the plan's autotuner (block shapes and radix orders of the
        general-size kernels; the factored kernels have nothing to tune) and
        return the configuration it keeps."""
        cfg = (ctypes.c_int32 * 6)()
        self.bind_stream()
        _lib.call("surfhip_ns_plan_tune", self._plan, cfg)
        def radices(code):
            out = []
            while code:
                out.append(code % 100)
                code //= 100
            return out[::-1]
        return {"factored": bool(cfg[0]), "rows_per_block": int(cfg[1]), "pencils_per_block": int(cfg[2]),
                "radix_orders": [radices(int(cfg[3 + a])) for a in range(3)]}

    def set_timing(self, enable):
        """Bracket every kernel launch with HIP events (a benchmarking hook)."""
        _lib.call("surfhip_ns_set_timing", self._plan, 1 if enable else 0)

    def timing(self):
        """{kernel class: (launches, total ms)} since timing was switched on."""
        n = 5
        launches = (ctypes.c_int64 * n)()
        ms = (ctypes.c_double * n)()
        names = ctypes.create_string_buffer(256)
        _lib.call("surfhip_ns_timing", self._plan, launches, ms, names, 256)
        keys = names.value.decode().split(",")
        return {k: (int(launches[i]), float(ms[i])) for i, k in enumerate(keys)}

    def band_buffer(self):
        if self._band is None:
            self._band = empty_colmajor(self.dims, self.dtype, self.device)
        return self._band


def transform(ws):
    return ws.transform


def _eltype(x):
    if isinstance(x, torch.dtype):
        dt = x
    else:
        dt = x.dtype
    if dt in (torch.float32, torch.complex64):
        return dt
    if dt in (torch.int8, torch.int16, torch.int32, torch.int64, torch.uint8, torch.bool):
        return torch.float32
    raise ArgumentError(f"libsurfhip computes in Float32 and ComplexF32; got element type {dt}")


def make_workspace(t, x_or_type, dims=None, *, flags=None, tasks=None, tune=False, prewarm=True,
                   cache_bandpass=True, device=None):
    """make_workspace(t, T, dims) or make_workspace(t, x): src/transforms/
    ns_surfacelet.jl:88-173.  `flags` and `cache_bandpass` are accepted (the
    device evaluates the bandpass on the fly); `tasks` must be ≥ 1.  `tune`
    runs the plan's autotuner (the reference accepts and ignores it)."""
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
    ws = NSSurfaceletWorkspace(t, dt, dims, device)
    if tune:
        ws.tune()
    if prewarm:
        z = torch.zeros(dims, dtype=dt, device=device)
        synthesize(ws, analyze(ws, as_colmajor(z)))
    return ws


def check_workspace(ws, t, eltype, dims):
    """The checks of AbstractMultiscaleTransforms.check_workspace as the
    reference's tests pin them (test/items/test_ns_surfacelet.jl:181-191)."""
    if not isinstance(ws, NSSurfaceletWorkspace):
        raise ArgumentError(f"an NSSurfacelet needs an NSSurfaceletWorkspace, got {type(ws).__name__}")
    if ws.transform != t:
        raise ArgumentError(f"the workspace was built for {ws.transform}, not {t}")
    if eltype != ws.dtype:
        raise ArgumentError(f"the workspace is for element type {ws.dtype}, got {eltype}")
    if tuple(dims) != tuple(ws.dims):
        raise DimensionMismatch(f"the workspace is for size {_jl_dims(ws.dims)}, got {_jl_dims(dims)}")
    return ws


def _workspace_for(workspace, t, x):
    """src/transforms/ns_surfacelet.jl:247-252."""
    if workspace is None:
        return make_workspace(t, x, prewarm=False)
    return check_workspace(workspace, t, _eltype(x), _check_volume(x.shape))


def _split(t_or_ws, workspace):
    if isinstance(t_or_ws, NSSurfaceletWorkspace):
        return t_or_ws.transform, t_or_ws
    return t_or_ws, workspace


def similar_coefficients(t, x):
    """src/transforms/ns_surfacelet.jl:537-540 -- (size(x)..., S), column-major."""
    dims = _check_volume(x.shape)
    return empty_colmajor(dims + (nsubbands(t, dims),), _eltype(x), x.device)


def subband(t, c, i):
    """src/transforms/ns_surfacelet.jl:542 -- view of subband i (1-based)."""
    return c[:, :, :, i - 1]


def signal_size(t, c):
    return tuple(c.shape[:3])


def _check_coefficients(c, t, dims):
    """src/transforms/ns_surfacelet.jl:545-549."""
    expect = tuple(dims) + (nsubbands(t, dims),)
    if tuple(c.shape) != expect:
        raise DimensionMismatch(f"coefficients of size {_jl_dims(tuple(c.shape))}, expected {_jl_dims(expect)}")
    return c


def _input(x, ws):
    x = x if x.dtype == ws.dtype else x.to(ws.dtype)
    return as_colmajor(x)


def analyze_(c, t, x, *, workspace=None, threaded=None):
    """analyze!(c, t, x): src/transforms/ns_surfacelet.jl:551-566."""
    t, workspace = _split(t, workspace)
    check_device(x, "volume")
    check_device(c, "coefficient array")
    ws = _workspace_for(workspace, t, x)
    _check_coefficients(c, t, ws.dims)
    if ws.dtype == torch.complex64 and c.dtype != torch.complex64:
        raise ArgumentError(f"coefficients of element type {c.dtype}, the workspace is for {ws.dtype}")
    xin = _input(x, ws)
    ws.bind_stream()
    if is_colmajor(c) and c.dtype == ws.dtype:
        _lib.call("surfhip_ns_analyze", ws.handle, ptr(xin), ptr(c))
    else:
        # Coefficients the library cannot write directly go through the
        # streaming path (as the reference routes non-stack arrays).
        _lib.call("surfhip_ns_stream_begin", ws.handle, ptr(xin))
        band = ws.band_buffer()
        for i in range(1, ws.nsub + 1):
            _lib.call("surfhip_ns_stream_subband", ws.handle, i, ptr(band))
            c[:, :, :, i - 1].copy_(band)
    return c


def analyze(*args, workspace=None, threaded=None):
    """analyze(t, x) -> coefficients, or the streaming form analyze(f, t, x)
    (src/transforms/ns_surfacelet.jl:620-624): f(i, band) for the directional
    subbands in increasing i, then the lowpass (i = 1); `band` is valid only
    until f returns."""
    if len(args) == 3 and callable(args[0]):
        f, t, x = args
        t, workspace = _split(t, workspace)
        check_device(x, "volume")
        ws = _workspace_for(workspace, t, x)
        xin = _input(x, ws)
        ws.bind_stream()
        _lib.call("surfhip_ns_stream_begin", ws.handle, ptr(xin))
        band = ws.band_buffer()
        for i in list(range(2, ws.nsub + 1)) + [1]:
            _lib.call("surfhip_ns_stream_subband", ws.handle, i, ptr(band))
            f(i, band)
        return None
    t, x = args
    t, workspace = _split(t, workspace)
    check_device(x, "volume")
    ws = _workspace_for(workspace, t, x)
    c = similar_coefficients(t, x if x.dtype == ws.dtype else x.to(ws.dtype))
    return analyze_(c, t, x, workspace=ws, threaded=threaded)


def synthesize_(*args, workspace=None, threaded=None):
    """synthesize!(x, t, c) (src/transforms/ns_surfacelet.jl:568-581), or the
    streaming form synthesize!(f, x, t) (:637-641) where f(i) returns subband i."""
    if callable(args[0]) and not isinstance(args[0], torch.Tensor):
        f, x, t = args
        t, workspace = _split(t, workspace)
        check_device(x, "volume")
        ws = _workspace_for(workspace, t, x)
        ws.bind_stream()
        _lib.call("surfhip_ns_stream_inverse_begin", ws.handle)
        for i in list(range(2, ws.nsub + 1)) + [1]:
            b = f(i)
            if tuple(b.shape) != tuple(ws.dims):
                raise DimensionMismatch(f"subband {i} has size {_jl_dims(tuple(b.shape))}, expected {_jl_dims(ws.dims)}")
            b = _input(b, ws)
            _lib.call("surfhip_ns_stream_accumulate", ws.handle, i, ptr(b))
        _finish_into(x, ws)
        return x
    x, t, c = args
    t, workspace = _split(t, workspace)
    check_device(x, "volume")
    check_device(c, "coefficient array")
    ws = _workspace_for(workspace, t, x)
    _check_coefficients(c, t, ws.dims)
    ws.bind_stream()
    if is_colmajor(c) and c.dtype == ws.dtype:
        if is_colmajor(x) and x.dtype == ws.dtype:
            _lib.call("surfhip_ns_synthesize", ws.handle, ptr(c), ptr(x))
        else:
            y = empty_colmajor(ws.dims, ws.dtype, ws.device)
            _lib.call("surfhip_ns_synthesize", ws.handle, ptr(c), ptr(y))
            x.copy_(y)
        return x
    return synthesize_(lambda i: subband(t, c, i), x, t, workspace=ws)


def _finish_into(x, ws):
    if is_colmajor(x) and x.dtype == ws.dtype:
        _lib.call("surfhip_ns_stream_finish", ws.handle, ptr(x))
    else:
        y = empty_colmajor(ws.dims, ws.dtype, ws.device)
        _lib.call("surfhip_ns_stream_finish", ws.handle, ptr(y))
        x.copy_(y)


def synthesize(t, c, *, workspace=None, threaded=None):
    """synthesize(t, c) -> volume."""
    t, workspace = _split(t, workspace)
    check_device(c, "coefficient array")
    dims = _check_volume(c.shape[:3])
    x = empty_colmajor(dims, _eltype(c), c.device)
    return synthesize_(x, t, c, workspace=workspace, threaded=threaded)


def ns_mask_host(t, dims, i):
    """Host-only: the analysis mask of subband i on the rfft grid, as the
    device evaluates it; numpy array of shape (n1 // 2 + 1, n2, n3)."""
    import numpy as np
    dims = _check_volume(dims)
    h = (dims[0] // 2 + 1, dims[1], dims[2])
    out = np.empty(int(np.prod(h)), dtype=np.float32)
    _lib.call("surfhip_ns_mask_host", ctypes.byref(t.c_params()), dims[0], dims[1], dims[2], i,
              out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    return out.reshape(h, order="F")


def adjoint(t, c, *, workspace=None):
    """x = A^T c, the adjoint of analyze (synthesize with the analysis masks;
    equal to synthesize for a tight bank)."""
    t, workspace = _split(t, workspace)
    check_device(c, "coefficient array")
    dims = _check_volume(c.shape[:3])
    ws = workspace if workspace is not None else make_workspace(t, _eltype(c), dims, prewarm=False, device=c.device)
    check_workspace(ws, t, _eltype(c), dims)
    _check_coefficients(c, t, ws.dims)
    cin = as_colmajor(c if c.dtype == ws.dtype else c.to(ws.dtype))
    x = empty_colmajor(ws.dims, ws.dtype, ws.device)
    ws.bind_stream()
    _lib.call("surfhip_ns_adjoint", ws.handle, ptr(cin), ptr(x))
    return x


def synthesize_adjoint(t, x, *, workspace=None):
    """c = S^T x, the adjoint of synthesize (analyze with the synthesis masks;
    equal to analyze for a tight bank)."""
    t, workspace = _split(t, workspace)
    check_device(x, "volume")
    ws = _workspace_for(workspace, t, x)
    xin = _input(x, ws)
    c = empty_colmajor(ws.dims + (ws.nsub,), ws.dtype, ws.device)
    ws.bind_stream()
    _lib.call("surfhip_ns_synthesize_adjoint", ws.handle, ptr(xin), ptr(c))
    return c
