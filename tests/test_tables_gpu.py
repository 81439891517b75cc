"""Group-factor tables built on the device (FP64, rounded once) are the host
build's bit for bit: both run the same per-voxel code with contraction off."""
import ctypes

import numpy as np
import pytest
import torch

surfhip = pytest.importorskip("surfhip")
from surfhip import _lib  # noqa: E402
from surfhip import ns as N  # noqa: E402


@pytest.mark.gpu
@pytest.mark.parametrize("dims,scales,levels,tight", [((32, 32, 32), 3, (2, 2, 3), True),
                                                      ((40, 48, 36), 2, (2, 3), False),
                                                      ((64, 32, 48), 3, (1, 2, 2), True)])
def test_device_tables_match_host_build(dims, scales, levels, tight):
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    t = N.NSSurfacelet(scales=scales, direction_levels=list(levels), tight=tight)
    ws = N.make_workspace(t, torch.float32, dims, prewarm=False)
    H = (dims[0] // 2 + 1) * dims[1] * dims[2]
    for g in range(3 * scales + 1):
        for syn in (0, 1):
            dev = np.empty(H, dtype=np.float32)
            host = np.empty(H, dtype=np.float32)
            _lib.call("surfhip_ns_group_table", ws.handle, g, syn, dev.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
            _lib.call("surfhip_ns_group_table_host", ctypes.byref(t.c_params()), dims[0], dims[1], dims[2], g, syn,
                      host.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
            assert np.array_equal(dev, host), (g, syn, int(np.sum(dev != host)))
