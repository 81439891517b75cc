"""surfhip_ns_analyze_noalloc against surfhip_ns_analyze, through the C ABI.

The allocation-free entry point runs a plan's analysis with its working
spectra and slabs in the library's static device scratch (64 MiB per GPU).
Its coefficients must equal the plan's own analysis bit for bit, on two
calls in a row where the second has a smaller input (the scratch then holds
the first call's data), on the factored and the general kernels and for
complex data; a plan whose working set exceeds the scratch is refused with
an ArgumentError before anything runs; and after a first call (which loads
the scratch) further calls take no device memory.
"""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

surfhip = pytest.importorskip("surfhip")
from surfhip import _lib  # noqa: E402
from surfhip import ns as NS  # noqa: E402
from surfhip._arrays import empty_colmajor, from_numpy  # noqa: E402


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return "cuda"


def _pair(dims, dtype, seed):
    t = NS.NSSurfacelet(scales=3, direction_levels=[2, 2, 3])
    ws = NS.make_workspace(t, dtype, dims, prewarm=False)
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(dims).astype(np.float32)
    if dtype == torch.complex64:
        x = (x + 1j * rng.standard_normal(dims)).astype(np.complex64)
    xd = from_numpy(x, "cuda")
    ref = NS.analyze(ws, xd)
    return ws, xd, ref


def _noalloc(ws, xd):
    c = empty_colmajor(tuple(ws.dims) + (ws.nsub,), ws.dtype, ws.device)
    c.fill_(float("nan"))
    ws.bind_stream()
    _lib.call("surfhip_ns_analyze_noalloc", ws.handle, ctypes.c_void_p(xd.data_ptr()),
              ctypes.c_void_p(c.data_ptr()))
    return c


@pytest.mark.parametrize("first,second", [
    (((64, 64, 64), torch.float32), ((32, 32, 32), torch.float32)),      # factored kernels
    (((48, 40, 36), torch.float32), ((24, 20, 18), torch.float32)),      # general kernels
    (((64, 32, 64), torch.complex64), ((32, 32, 32), torch.complex64)),  # two channels
    (((40, 48, 40), torch.float32), ((32, 64, 32), torch.float32)),      # general, then factored
])
def test_noalloc_analysis_equals_plan_analysis(first, second):
    _dev()
    print(f"case noalloc {first} then {second}", flush=True)
    runs = [_pair(first[0], first[1], 300), _pair(second[0], second[1], 301)]
    for ws, xd, ref in runs:                      # two calls in a row, the second smaller
        got = _noalloc(ws, xd)
        torch.cuda.synchronize()
        assert torch.equal(got, ref), ws.dims
    for ws, xd, ref in runs:                      # and again, reusing what the scratch holds
        got = _noalloc(ws, xd)
        torch.cuda.synchronize()
        assert torch.equal(got, ref), ws.dims


def test_noalloc_refuses_plans_beyond_the_scratch():
    _dev()
    t = NS.NSSurfacelet(scales=3, direction_levels=[2, 2, 3])
    ws = NS.make_workspace(t, torch.float32, (128, 128, 128), prewarm=False)
    x = torch.zeros((128, 128, 128), device="cuda")
    c = torch.zeros((128 ** 3 * ws.nsub,), device="cuda")
    torch.cuda.synchronize()
    status = _lib.lib().surfhip_ns_analyze_noalloc(ws.handle, ctypes.c_void_p(x.data_ptr()),
                                                   ctypes.c_void_p(c.data_ptr()))
    assert status == _lib.ERR_ARG
    assert "static scratch holds 67108864" in _lib.last_error()


def test_noalloc_takes_no_device_memory_after_the_first_call():
    _dev()
    ws, xd, ref = _pair((64, 64, 64), torch.float32, 302)
    c = empty_colmajor(tuple(ws.dims) + (ws.nsub,), ws.dtype, ws.device)
    args = (ws.handle, ctypes.c_void_p(xd.data_ptr()), ctypes.c_void_p(c.data_ptr()))
    ws.bind_stream()
    _lib.call("surfhip_ns_analyze_noalloc", *args)
    torch.cuda.synchronize()
    free0, _ = torch.cuda.mem_get_info()
    for _ in range(3):
        _lib.call("surfhip_ns_analyze_noalloc", *args)
    torch.cuda.synchronize()
    free1, _ = torch.cuda.mem_get_info()
    assert free1 == free0
    assert torch.equal(c, ref)
