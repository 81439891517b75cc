"""Batched NSSurfacelet analyze / synthesize of many volumes of uneven sizes,
on libsurfhip's batch (surfhip_ns_batch_*).

The library holds one plan per (lane, volume size) and spreads the volumes
over `lanes` HIP streams, longest first onto the least-loaded lane, so
transforms of different volumes overlap.  A run is recorded once as a
hipGraph and replayed with a single launch while the mode and the arrays
repeat (graph=True); results equal each volume's lone transform on a plan of
the same configuration, bit for bit.  The reference transforms one volume per
call and has no device batching (src/transforms/ns_surfacelet.jl:147,551-581).
"""
import ctypes

import torch

from . import _lib
from . import ns as _ns
from ._arrays import ctypes_void, empty_colmajor, is_colmajor
from ._lib import ArgumentError

TUNE, NO_GRAPH = 1, 2
MODES = {"analyze": 0, "synthesize": 1, "both": 2}


class NSBatch:
    def __init__(self, t, shapes, dtype=torch.float32, device="cuda", lanes=4, tune=False, graph=True):
        if lanes < 1:
            raise ArgumentError(f"lanes must be ≥ 1, got {lanes}")
        if len(shapes) == 0:
            raise ArgumentError("a batch needs at least one volume")
        self.transform = t
        self.shapes = [_ns._check_volume(s) for s in shapes]
        self.dtype = _ns._eltype(dtype)
        self.device = torch.device(device)
        if self.device.index is None:
            self.device = torch.device(self.device.type, torch.cuda.current_device())
        self.lanes = min(lanes, len(self.shapes))
        self.nsub = [_ns.nsubbands(t, s) for s in self.shapes]
        dims = (ctypes.c_int32 * (3 * len(self.shapes)))(*[d for s in self.shapes for d in s])
        flags = (TUNE if tune else 0) | (0 if graph else NO_GRAPH)
        self._b = ctypes.c_void_p()
        code = _lib.F32 if self.dtype == torch.float32 else _lib.C64
        with torch.cuda.device(self.device):
            _lib.call("surfhip_ns_batch_create", ctypes.byref(self._b), ctypes.byref(t.c_params()),
                      len(self.shapes), dims, code, lanes, flags,
                      ctypes_void(torch.cuda.current_stream(self.device).cuda_stream))
        lane_of = (ctypes.c_int32 * len(self.shapes))()
        _lib.call("surfhip_ns_batch_info", self._b, lane_of, None)
        self.lane_of = list(lane_of)

    def __del__(self):
        b = getattr(self, "_b", None)
        if b is not None and b.value:
            try:
                _lib.lib().surfhip_ns_batch_destroy(b)
            except Exception:
                pass
            self._b = None

    def graph_recorded(self):
        rec = ctypes.c_int32()
        _lib.call("surfhip_ns_batch_info", self._b, None, ctypes.byref(rec))
        return bool(rec.value)

    def coefficients(self):
        """Column-major coefficient arrays for every volume of the batch."""
        return [empty_colmajor(s + (n,), self.dtype, self.device) for s, n in zip(self.shapes, self.nsub)]

    def _check(self, name, arrs, coef):
        if arrs is None:
            raise ArgumentError(f"{name} are needed for this mode")
        if len(arrs) != len(self.shapes):
            raise ArgumentError(f"{len(arrs)} {name} for a batch of {len(self.shapes)}")
        for k, (a, s, n) in enumerate(zip(arrs, self.shapes, self.nsub)):
            shape = s + (n,) if coef else s
            if (not isinstance(a, torch.Tensor) or a.device != self.device or tuple(a.shape) != shape
                    or not is_colmajor(a) or a.dtype != self.dtype):
                raise ArgumentError(f"{name} {k} must be a column-major {self.dtype} array of size {shape} "
                                    f"on {self.device}")

    def run(self, xs=None, cs=None, ys=None, mode="both"):
        """mode 'analyze': cs[k] = analyze(xs[k]); 'synthesize': ys[k] =
        synthesize(cs[k]); 'both': the two in turn, for every volume.
        Ordered after the current stream's earlier work and before its later
        work."""
        m = MODES[mode]
        if m != 1:
            self._check("volumes", xs, False)
        self._check("coefficients", cs, True)
        if m != 0:
            self._check("outputs", ys, False)
        n = len(self.shapes)
        P = ctypes.c_void_p * n
        xa = P(*[x.data_ptr() for x in xs]) if m != 1 else None
        ca = P(*[c.data_ptr() for c in cs])
        ya = P(*[y.data_ptr() for y in ys]) if m != 0 else None
        _lib.call("surfhip_ns_batch_set_stream", self._b,
                  ctypes_void(torch.cuda.current_stream(self.device).cuda_stream))
        _lib.call("surfhip_ns_batch_run", self._b, m, xa, ca, ya)
        return ys if m != 0 else cs
