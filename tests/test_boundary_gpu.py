"""Failure handling at the C boundary, forced from the host (SURVEY's error
model; the reference's DimensionMismatch / ArgumentError texts are at
test/items/test_ns_surfacelet.jl:181-193).

- A device buffer that ends before the extent a plan reads or writes is
  refused with SURFHIP_ERR_DIM before anything is launched; a host pointer
  with SURFHIP_ERR_ARG.  The undersized buffers come from hipMalloc directly
  (an exact allocation), so the check sees their true extent.
- A device allocation the library refuses at plan creation returns
  SURFHIP_ERR_OOM: above the device's total memory, or above the
  SURFHIP_MAX_ALLOCATION cap, which the test sets.
No kernel runs on an undersized buffer: the check precedes every launch.
"""
import ctypes
import os

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

surfhip = pytest.importorskip("surfhip")
from surfhip import _lib  # noqa: E402
from surfhip import decimated as SD  # noqa: E402
from surfhip import ns as NS  # noqa: E402


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return "cuda"


def _hip():
    for name in ("libamdhip64.so", os.path.join("/opt/rocm/lib", "libamdhip64.so")):
        try:
            h = ctypes.CDLL(name)
        except OSError:
            continue
        h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        h.hipMalloc.restype = ctypes.c_int
        h.hipFree.argtypes = [ctypes.c_void_p]
        h.hipFree.restype = ctypes.c_int
        return h
    pytest.skip("libamdhip64 not found")


class _Raw:
    """One exact hipMalloc allocation, freed on exit."""

    def __init__(self, nbytes):
        self.h = _hip()
        self.p = ctypes.c_void_p()
        assert self.h.hipMalloc(ctypes.byref(self.p), nbytes) == 0
        self.nbytes = nbytes

    def __enter__(self):
        return self.p

    def __exit__(self, *exc):
        torch.cuda.synchronize()
        assert self.h.hipFree(self.p) == 0


def _status(name, *args):
    return getattr(_lib.lib(), name)(*args)


def test_ns_undersized_and_host_buffers_are_refused():
    dev = _dev()
    dims = (32, 32, 32)
    t = NS.NSSurfacelet(scales=3, direction_levels=[2, 2, 3])
    ws = NS.make_workspace(t, torch.float32, dims, prewarm=False)
    N, S = 32 ** 3, NS.nsubbands(t, dims)
    x = torch.zeros(dims, device=dev)
    c = torch.zeros((N * S,), device=dev)
    torch.cuda.synchronize()
    xp, cp = ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(c.data_ptr())
    with _Raw(N * S) as small_c:            # a quarter of the 4 N S bytes needed
        for name, args in (("surfhip_ns_analyze", (xp, small_c)), ("surfhip_ns_synthesize", (small_c, xp)),
                           ("surfhip_ns_adjoint", (small_c, xp)), ("surfhip_ns_synthesize_adjoint", (xp, small_c))):
            assert _status(name, ws.handle, *args) == _lib.ERR_DIM, name
            assert "coefficient array" in _lib.last_error()
    with _Raw(N) as small_x:                 # a quarter of one volume
        assert _status("surfhip_ns_analyze", ws.handle, small_x, cp) == _lib.ERR_DIM
        assert "volume" in _lib.last_error()
        assert _status("surfhip_ns_stream_begin", ws.handle, small_x) == _lib.ERR_DIM
        assert _status("surfhip_ns_stream_finish", ws.handle, small_x) == _lib.ERR_DIM
    host = np.zeros(N * S, dtype=np.float32)
    assert _status("surfhip_ns_analyze", ws.handle, xp, ctypes.c_void_p(host.ctypes.data)) == _lib.ERR_ARG
    assert "not memory the device can address" in _lib.last_error()
    # The plan still works after the refusals.
    y = NS.synthesize(ws, NS.analyze(ws, x + 1.0))
    torch.cuda.synchronize()
    assert float((y - 1.0).abs().max()) < 1e-5


def test_decimated_undersized_buffers_are_refused():
    dev = _dev()
    t = SD.Surfacelet(scales=2, direction_levels=[1, 2])
    dims = (48, 24, 36)
    ws = SD.make_workspace(t, torch.float32, dims)
    x = torch.zeros(dims, device=dev)
    n = ws.layout.length
    torch.cuda.synchronize()
    with _Raw(n) as small:                   # a quarter of the 4 ncoef bytes needed
        xp = ctypes.c_void_p(x.data_ptr())
        for name, args in (("surfhip_s_analyze", (xp, small)), ("surfhip_s_synthesize", (small, xp)),
                           ("surfhip_s_adjoint", (small, xp))):
            assert _status(name, ws._plan, *args) == _lib.ERR_DIM, name
            assert "coefficient vector" in _lib.last_error()


def test_plan_creation_reports_allocation_failure_as_oom(monkeypatch):
    # Holding all but 16 MiB of the free device memory with one hipMalloc
    # did not make the 128^3 plan's allocations fail on an MI355X (the plan
    # built; device memory is evidently oversubscribed rather than refused),
    # so the failure is forced with the
    # library's per-allocation cap: 8 MiB is below the 128^3 plans' first
    # large tables (NS: 43 MB, decimated: 12.8 MB), allocated before any
    # kernel runs.
    _dev()
    monkeypatch.setenv("SURFHIP_MAX_ALLOCATION", str(8 << 20))
    t = NS.NSSurfacelet(scales=3, direction_levels=[2, 2, 3])
    with pytest.raises(_lib.DeviceOutOfMemory, match="SURFHIP_MAX_ALLOCATION"):
        NS.NSSurfaceletWorkspace(t, torch.float32, (128, 128, 128), "cuda")
    assert _status("surfhip_ns_plan_create", ctypes.byref(ctypes.c_void_p()), ctypes.byref(t.c_params()),
                   128, 128, 128, _lib.F32, None) == _lib.ERR_OOM
    with pytest.raises(_lib.DeviceOutOfMemory):
        SD.SurfaceletWorkspace(SD.Surfacelet(scales=3, direction_levels=[2, 2, 3], pyramid_ratio=2),
                               torch.float32, (128, 128, 128), "cuda")
    monkeypatch.delenv("SURFHIP_MAX_ALLOCATION")
    ws = NS.NSSurfaceletWorkspace(t, torch.float32, (128, 128, 128), "cuda")
    del ws
