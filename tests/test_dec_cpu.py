"""Host-side checks of the decimated Surfacelet boundary: transform object,
size rules with the reference's messages, subband layout and redundancy, all
through the native library's integer code, against the oracle and the
reference's own expectations (test/items/test_surfacelet.jl,
test/items/test_irc.jl:186-193)."""
import numpy as np
import pytest

from oracle import decimated as D
from oracle.golden import load_golden

surfhip = pytest.importorskip("surfhip")
from surfhip import decimated as S  # noqa: E402

L111 = np.array([[-1, 1, 1], [1, -1, 1], [1, 1, -1]])
L212 = np.array([[-1, 2, 1], [1, -1, 2], [2, 1, -1]])


def test_construction_validation_and_equality():
    # test/items/test_surfacelet.jl:1-58
    t = S.Surfacelet(scales=2, direction_levels=[1, 2])
    assert t.scales == 2
    assert t == S.Surfacelet(scales=2, direction_levels=[np.int8(1), np.int8(2)])
    assert hash(t) == hash(S.Surfacelet(scales=2, direction_levels=[1, 2]))
    assert t != S.Surfacelet(scales=2, direction_levels=[1, 2], checkerboard_order=8)
    assert repr(t) == ("Surfacelet(scales = 2, direction_levels = [1, 2], pyramid_ratio = 1.5, checkerboard_order = 12, "
                       "diamond_length = 15, kaiser_beta = 2.5, diamond_power = 4)")
    lay = S.subband_layout(t, (48, 48, 48))
    assert lay[0] == (0, 0, 0) and lay[1] == (1, 1, 1)
    assert sum(d.scale == 1 for d in lay) == 6 and sum(d.scale == 2 for d in lay) == 12
    assert S.nsubbands(t, (48, 48, 48)) == 19
    assert S.supports_streaming(t) is False
    q = S.Surfacelet(scales=1, direction_levels=[L212])
    assert q.levels[0] == ((0, 2, 1), (1, 0, 2), (2, 1, 0))
    assert S.nsubbands(q, (96, 96, 96)) == 25
    assert S.checkerboard_orders() == list(range(4, 19, 2))
    for kw, msg in [
        (dict(scales=0, direction_levels=[]), "scales must be ≥ 1"),
        (dict(scales=2, direction_levels=[1]), r"length\(direction_levels\) must equal scales"),
        (dict(scales=1, direction_levels=[-1]), "directional levels must be ≥ 0"),
        (dict(scales=1, direction_levels=[np.ones((2, 2), dtype=int)]), "3 × 3"),
        (dict(scales=1, direction_levels=[1.5]), "integer or a 3 × 3"),
        (dict(scales=1, direction_levels=[1], pyramid_ratio=3), "pyramid_ratio"),
        (dict(scales=1, direction_levels=[1], checkerboard_order=5), "checkerboard_order must be one of"),
        (dict(scales=1, direction_levels=[1], diamond_length=14), "diamond_length"),
        (dict(scales=1, direction_levels=[1], kaiser_beta=0), "kaiser_beta"),
        (dict(scales=1, direction_levels=[1], diamond_power=3), "diamond_power"),
    ]:
        with pytest.raises(surfhip.ArgumentError, match=msg):
            S.Surfacelet(**kw)


def test_size_rules_and_messages():
    # test/items/test_surfacelet.jl:175-207
    t = S.Surfacelet(scales=1, direction_levels=[1])
    with pytest.raises(surfhip.ArgumentError):
        S.check_size(t, (20, 24, 24))
    with pytest.raises(surfhip.ArgumentError):
        S.check_size(S.Surfacelet(scales=1, direction_levels=[7]), (24, 24, 24))
    with pytest.raises(surfhip.ArgumentError, match="three axes"):
        S.subband_layout(t, (24, 24))
    t2 = S.Surfacelet(scales=1, direction_levels=[0], pyramid_ratio=2)
    with pytest.raises(surfhip.ArgumentError, match=(
            r"axis 1: at scale 1 of 1 \(1 the coarsest\) the length is 10, too small for the pyramid's lowpass; "
            r"the nearest length that fits is 8")):
        S.check_size(t2, (10, 10, 10))
    with pytest.raises(surfhip.ArgumentError, match=r"axis 1: at scale 3 of 3 \(1 the coarsest\) the length is 20"):
        S.check_size(S.Surfacelet(scales=3), (20, 24, 24))
    with pytest.raises(surfhip.ArgumentError, match="resampling by 3/2 needs to be a multiple of 3"):
        S.check_size(S.Surfacelet(scales=1, direction_levels=[0]), (20, 24, 24))
    for f in (S.subband_layout, S.surfacelet_redundancy, S.nsubbands):
        with pytest.raises(surfhip.ArgumentError, match="too small for the pyramid's lowpass"):
            f(t2, (10, 10, 10))
    # The headline decimated workload does not fit the default ratio at 128.
    with pytest.raises(surfhip.ArgumentError, match="the nearest length that fits is 132"):
        S.check_size(S.Surfacelet(scales=3, direction_levels=[2, 2, 3]), (128, 128, 128))
    S.check_size(S.Surfacelet(scales=3, direction_levels=[2, 2, 3], pyramid_ratio=2), (128, 128, 128))


@pytest.mark.parametrize("dims,mode,J,L", [
    ((48, 48, 48), 1.5, 2, [1, 2]),
    ((48, 24, 36), 1.5, 1, [2]),
    ((32, 32, 32), 1, 2, [1, 1]),
    ((64, 32, 32), 2, 2, [1, 3]),
    ((128, 128, 128), 2, 3, [2, 2, 3]),
    ((96, 96, 96), 1.5, 2, [L212, L111]),
])
def test_layout_matches_oracle(dims, mode, J, L):
    t = S.Surfacelet(scales=J, direction_levels=L, pyramid_ratio=mode)
    to = D.SParams(scales=J, direction_levels=tuple(L), pyramid_ratio=mode)
    assert S._sizes(t, dims) == [tuple(s) for s in D.layout_sizes(to, dims)]
    assert [tuple(e) for e in S.subband_layout(t, dims)] == D.subband_layout(to, dims)
    assert S.surfacelet_redundancy(t, dims) == D.redundancy(to, dims)


def test_redundancy_matches_surfbox():
    # test/items/test_irc.jl:186-193 (golden values from SurfBox)
    G = load_golden()
    for nm, Lm, J in (("d24", [[-1, 1, 2], [2, -1, 1], [1, 2, -1]], 3), ("d96", [[-1, 2, 3], [3, -1, 2], [2, 3, -1]], 2)):
        t = S.Surfacelet(scales=J, direction_levels=[np.array(Lm)] * J, checkerboard_order=4)
        assert abs(S.surfacelet_redundancy(t, (96, 96, 96)) - G[f"red_96_{nm}"][0, 0]) < 1e-12
    # headline decimated workload: 3.424 coefficients per voxel
    t = S.Surfacelet(scales=3, direction_levels=[2, 2, 3], pyramid_ratio=2)
    assert sum(int(np.prod(s)) for s in S._sizes(t, (128, 128, 128))) == 7180288


def test_critical_sampling_per_hourglass():
    # test/items/test_surfacelet.jl:113-132
    red = lambda L: S.surfacelet_redundancy(S.Surfacelet(scales=2, direction_levels=L, checkerboard_order=4), (96, 96, 96))
    assert red([1, 1]) == red([3, 2]) == red([4, 4])
    assert abs(red([1, 1]) - (3 + 3 * (64 / 96) ** 3 + (32 / 96) ** 3)) < 1e-12


def test_coefficients_on_the_host_are_refused_before_any_launch():
    # A host-resident coefficient vector must never reach the kernels as a
    # pointer: the check raises before the workspace or the library is used.
    import types
    import torch
    t = S.Surfacelet(scales=1, direction_levels=[1])
    lay = S.SubbandLayout(S._sizes(t, (24, 24, 24)), (24, 24, 24))
    ws = types.SimpleNamespace(layout=lay, dtype=torch.float32, transform=t, dims=(24, 24, 24))
    c = S.SubbandVector(torch.zeros(lay.length), lay)
    with pytest.raises(surfhip.ArgumentError, match="device"):
        S._check_coefficients(c, ws)
