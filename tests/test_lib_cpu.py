"""CPU-side checks of libsurfhip: it loads, exports every entry point the
header declares, and its host-only functions (table construction, the mask
expression the kernels evaluate, parameter validation) agree with the oracle.
No device compute runs here."""
import ctypes
import os
import re

import numpy as np
import pytest

from oracle import ns as O

surfhip = pytest.importorskip("surfhip")
from surfhip import ns as N  # noqa: E402

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    src = open(os.path.join(ROOT, "include", "surfhip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(surfhip_\w+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    L = surfhip.lib()
    names = header_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert b"gfx950" in L.surfhip_version()


@pytest.mark.parametrize("tight", [True, False])
@pytest.mark.parametrize("dims", [(16, 16, 16), (15, 20, 8), (9, 7, 5), (26, 26, 26)])
def test_device_mask_expression_matches_oracle(dims, tight):
    t = N.NSSurfacelet(scales=2, direction_levels=[1, 2], bernstein_order=6, tight=tight)
    masks = O.subband_masks(O.NSParams(2, (1, 2), bernstein_order=6, tight=tight), dims)
    for i, ref in enumerate(masks, start=1):
        got = N.ns_mask_host(t, dims, i)
        assert got.shape == ref.shape
        assert np.max(np.abs(got - ref)) < 5e-7, i


def test_headline_masks_match_oracle():
    dims = (40, 36, 44)
    t = N.NSSurfacelet(scales=3, direction_levels=[2, 2, 3])
    masks = O.subband_masks(O.NSParams(3, (2, 2, 3)), dims)
    for i in (1, 2, 14, 30, 49):
        assert np.max(np.abs(N.ns_mask_host(t, dims, i) - masks[i - 1])) < 5e-7


def test_native_validation_messages():
    p = N.NSSurfacelet(scales=4, direction_levels=[0, 0, 0, 0], bernstein_order=4)
    with pytest.raises(surfhip.ArgumentError, match=r"size \(8, 8, 8\) is too small for 4 scales: scale 1 would be empty"):
        N.ns_mask_host(p, (8, 8, 8), 1)
    S = ctypes.c_int32()
    surfhip._lib.call("surfhip_ns_nsubbands", ctypes.byref(N.NSSurfacelet(scales=3, direction_levels=[2, 2, 3]).c_params()),
                      ctypes.byref(S))
    assert S.value == 49


def test_host_transform_object():
    # test/items/test_ns_surfacelet.jl:1-38 and :316-324
    t = N.NSSurfacelet(scales=3, direction_levels=[1, 2, 2])
    assert t == N.NSSurfacelet(scales=3, direction_levels=[1, 2, 2])
    assert hash(N.NSSurfacelet(scales=1, transition=-0.0)) == hash(N.NSSurfacelet(scales=1, transition=0.0))
    assert t != N.NSSurfacelet(scales=3, direction_levels=[1, 2, 2], tight=False)
    assert repr(t) == ("NSSurfacelet(scales = 3, direction_levels = [1, 2, 2], bernstein_order = 12, "
                       "halfband_order = 1, transition = 1.0, tight = true)")
    lay = N.subband_layout(t, (16, 16, 16))
    assert N.nsubbands(t, (16, 16, 16)) == len(lay) == 31
    assert lay[0] == (0, 0) and lay[1] == (1, 1) and lay[7] == (2, 1)
    with pytest.raises(surfhip.ArgumentError, match="three axes"):
        N.subband_layout(t, (16, 16))
    for kw, msg in [(dict(scales=0, direction_levels=[]), "scales must be ≥ 1"),
                    (dict(scales=2, direction_levels=[1]), "length\\(direction_levels\\) must equal scales"),
                    (dict(scales=1, direction_levels=[-1]), "direction_levels entries must be ≥ 0"),
                    (dict(scales=1, direction_levels=[1], bernstein_order=0), "bernstein_order must be ≥ 1"),
                    (dict(scales=1, direction_levels=[1], halfband_order=-1), "halfband_order must be ≥ 0"),
                    (dict(scales=1, direction_levels=[1], transition=-1), "transition must be ≥ 0"),
                    (dict(), "give `scales`, `direction_levels` or both")]:
        with pytest.raises(surfhip.ArgumentError, match=msg):
            N.NSSurfacelet(**kw)
    assert N.NSSurfacelet(scales=4).direction_levels == [1, 1, 2, 2]
    assert N.NSSurfacelet(direction_levels=[1, 2]).scales == 2
