"""Batched NSSurfacelet analyze+synthesize of many volumes of uneven sizes.

The volumes are spread over `lanes` HIP streams (longest-first onto the
least-loaded lane), each lane with its own workspace per distinct size, so
transforms of different volumes run concurrently and fill the GPU where one
volume's kernels would not.  `capture` records the whole batch, every lane
forked from and joined back to the capturing stream, as one graph (hipGraph,
through torch.cuda.CUDAGraph) that `replay` launches with a single call.
Workspaces are autotuned by default (make_workspace(...; tune=true)).

Graph capture records kernel launches only: the transforms' launch sequence
holds no host synchronisation and no memset (buffers are cleared by
kernels), so a replay recomputes exactly what the eager run computes.
"""
import torch

from . import ns as _ns
from ._arrays import empty_colmajor, is_colmajor
from ._lib import ArgumentError


class NSBatch:
    def __init__(self, t, shapes, dtype=torch.float32, device="cuda", lanes=4, tune=True):
        if lanes < 1:
            raise ArgumentError(f"lanes must be ≥ 1, got {lanes}")
        if len(shapes) == 0:
            raise ArgumentError("a batch needs at least one volume")
        self.transform = t
        self.shapes = [tuple(int(d) for d in s) for s in shapes]
        self.dtype = dtype
        self.device = torch.device(device)
        self.lanes = min(lanes, len(self.shapes))
        # Longest-processing-time assignment: biggest volumes first, each onto
        # the lane with the least voxels so far.
        order = sorted(range(len(self.shapes)), key=lambda k: -_voxels(self.shapes[k]))
        load = [0] * self.lanes
        self.lane_of = [0] * len(self.shapes)
        for k in order:
            lane = min(range(self.lanes), key=lambda q: load[q])
            self.lane_of[k] = lane
            load[lane] += _voxels(self.shapes[k])
        self.streams = [torch.cuda.Stream(self.device) for _ in range(self.lanes)]
        self.workspaces = {}
        self.tuning = {}
        for k, s in enumerate(self.shapes):
            key = (self.lane_of[k], s)
            if key not in self.workspaces:
                ws = _ns.make_workspace(t, dtype, s, device=self.device, prewarm=False)
                if tune:
                    self.tuning[key] = ws.tune()
                self.workspaces[key] = ws

    def workspace(self, k):
        """The workspace that transforms volume k."""
        return self.workspaces[(self.lane_of[k], self.shapes[k])]

    def coefficients(self):
        """Column-major coefficient arrays for every volume of the batch."""
        return [_ns.similar_coefficients(self.transform, empty_colmajor(s, self.dtype, self.device))
                for s in self.shapes]

    def _check(self, xs, cs, ys):
        for name, arrs in (("volumes", xs), ("coefficients", cs), ("outputs", ys)):
            if len(arrs) != len(self.shapes):
                raise ArgumentError(f"{len(arrs)} {name} for a batch of {len(self.shapes)}")
        for k, s in enumerate(self.shapes):
            for name, a, shape in (("volume", xs[k], s), ("output", ys[k], s)):
                if tuple(a.shape) != s or not is_colmajor(a) or a.dtype != self.dtype:
                    raise ArgumentError(f"{name} {k} must be a column-major {self.dtype} array of size {s}")
            if not is_colmajor(cs[k]) or cs[k].dtype != self.dtype:
                raise ArgumentError(f"coefficients {k} must be a column-major {self.dtype} array")

    def run(self, xs, cs, ys):
        """cs[k] = analyze(xs[k]); ys[k] = synthesize(cs[k]) for every volume,
        the lanes running concurrently; ordered after the current stream's
        earlier work and before its later work."""
        self._check(xs, cs, ys)
        main = torch.cuda.current_stream(self.device)
        for s in self.streams:
            s.wait_stream(main)
        for k in range(len(self.shapes)):
            with torch.cuda.stream(self.streams[self.lane_of[k]]):
                ws = self.workspace(k)
                _ns.analyze_(cs[k], ws, xs[k])
                _ns.synthesize_(ys[k], ws, cs[k])
        for s in self.streams:
            main.wait_stream(s)
        for ws in self.workspaces.values():
            ws.bind_stream()
        return ys

    def capture(self, xs, cs, ys):
        """Record run(xs, cs, ys) as one graph; replay it with graph.replay()."""
        self._check(xs, cs, ys)
        side = torch.cuda.Stream(self.device)
        side.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(side):
            self.run(xs, cs, ys)  # eager pass: first-use work stays out of the graph
        torch.cuda.current_stream(self.device).wait_stream(side)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=side):
            self.run(xs, cs, ys)
        return g


def _voxels(s):
    return s[0] * s[1] * s[2]
