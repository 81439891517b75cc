"""NSSurfacelet on the GPU through libsurfhip vs the FP64 oracle.

Tolerances: FP32 results against the FP64 oracle, relative L2 over all
subbands <= 1e-5 (the FP32 accuracy bar) and per subband <= 1e-4 of that
subband's norm or 1e-6 of the input norm, whichever is larger (subbands with
almost no energy carry only rounding).  Structural properties the reference
asserts bit for bit (streaming == materialised, complex == re + i im,
workspace reuse) are asserted bit for bit here.
"""
import numpy as np
import pytest
import torch

from oracle import ns as O

pytestmark = pytest.mark.gpu

surfhip = pytest.importorskip("surfhip")
from surfhip import ns as N  # noqa: E402
from surfhip._arrays import from_numpy, to_numpy  # noqa: E402

RTOL = 1e-5


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return "cuda"


def _rel(a, b):
    return float(np.linalg.norm((np.asarray(a) - np.asarray(b)).ravel()) / max(np.linalg.norm(np.asarray(b).ravel()), 1e-300))


def _check_coeffs(c_gpu, c_ref, xnorm):
    S = len(c_ref)
    got = to_numpy(c_gpu)
    num = sum(np.sum(np.abs(got[..., i] - c_ref[i]) ** 2) for i in range(S))
    den = sum(np.sum(np.abs(c_ref[i]) ** 2) for i in range(S))
    assert np.sqrt(num / den) <= RTOL, np.sqrt(num / den)
    for i in range(S):
        err = np.linalg.norm((got[..., i] - c_ref[i]).ravel())
        ref = np.linalg.norm(c_ref[i].ravel())
        assert err <= max(1e-4 * ref, 1e-6 * xnorm), (i, err, ref)


CASES = [
    ((16, 16, 16), 2, (1, 2), 6, True),
    ((12, 20, 8), 3, (1, 1, 2), 6, True),
    ((15, 20, 8), 3, (1, 1, 2), 6, True),
    ((9, 7, 5), 1, (1,), 6, True),
    ((16, 12, 20), 2, (1, 2), 6, False),
    ((26, 26, 26), 2, (2, 2), 12, True),
    ((32, 24, 40), 3, (2, 2, 3), 12, True),
    # power-of-two grids take the fused fast path
    ((32, 32, 32), 2, (1, 2), 6, True),
    ((64, 32, 128), 3, (1, 2, 3), 12, True),
    ((128, 64, 32), 2, (2, 3), 12, True),
    ((32, 128, 64), 3, (2, 2, 3), 12, False),
    ((64, 64, 256), 2, (1, 2), 6, True),
    # composite radices 14, 9, 6 and 12 (28 = 14 x 2, 54 = 9 x 6, 36 = 12 x 3)
    ((28, 54, 36), 2, (1, 2), 6, True),
    # four scales, direction_levels [2, 2, 3, 3] (73 subbands), the ISTA
    # workload's transform: factored and general kernels
    ((64, 64, 64), 4, (2, 2, 3, 3), 12, True),
    ((72, 80, 64), 4, (2, 2, 3, 3), 12, True),
    # the factored kernels at lengths 80, 96, 112 (8 threads x 10, 12, 14
    # points per pencil, uneven phase-2 rows), mixed with the power-of-two
    # lengths and with each other
    ((96, 80, 112), 3, (2, 2, 3), 12, True),
    ((112, 96, 32), 2, (1, 2), 6, False),
    ((80, 128, 96), 2, (2, 3), 12, True),
]


@pytest.mark.parametrize("dims,J,L,bo,tight", CASES)
def test_ns_analyze_synthesize_match_oracle(dims, J, L, bo, tight):
    dev = _dev()
    print(f"case dims={dims} J={J} L={L} bo={bo} tight={tight}", flush=True)
    rng = np.random.default_rng(hash((dims, J)) % 2 ** 32)
    x = rng.standard_normal(dims).astype(np.float32)
    t = N.NSSurfacelet(scales=J, direction_levels=list(L), bernstein_order=bo, tight=tight)
    to = O.NSParams(J, tuple(L), bernstein_order=bo, tight=tight)
    tb = O.tables(to, dims)
    ref = O.analyze(to, x.astype(np.float64), tb)
    xd = from_numpy(x, dev)
    ws = N.make_workspace(t, xd)
    c = N.analyze(t, xd, workspace=ws)
    torch.cuda.synchronize()
    _check_coeffs(c, ref, np.linalg.norm(x))
    # synthesis of random coefficients against the oracle's synthesis
    r = [rng.standard_normal(dims).astype(np.float32) for _ in ref]
    rd = from_numpy(np.stack(r, axis=-1), dev)
    y = N.synthesize(t, rd, workspace=ws)
    yref = O.synthesize(to, [v.astype(np.float64) for v in r], tb)
    assert _rel(to_numpy(y), yref) <= RTOL
    # perfect reconstruction
    xr = N.synthesize(t, c, workspace=ws)
    assert _rel(to_numpy(xr), x) <= RTOL
    # streaming at full size equals the materialised transform bit for bit
    y2 = torch.empty_like(xr)
    N.synthesize_(lambda i: N.subband(t, c, i), y2, t, workspace=ws)
    assert torch.equal(xr, y2)


@pytest.mark.parametrize("route", ["strip", "axis"])
@pytest.mark.parametrize("dims,J,L,bo,tight", [CASES[1], CASES[2], CASES[4], CASES[6], CASES[11], CASES[12]])
def test_ns_per_axis_passes_match_oracle(dims, J, L, bo, tight, route, monkeypatch):
    """General sizes run axes 2 and 3 as one plane kernel where a (k2, k3)
    plane fits one block's LDS, else as one FFT pass per axis: a strip kernel
    (in-place passes over up to 32 pencils) where its radices allow, else the
    Stockham pencil kernel.  Both per-axis routes on the same cases: oracle
    parity, agreement with the plane route, and streaming synthesis equal to
    the materialised one bit for bit."""
    dev = _dev()
    print(f"case dims={dims} J={J} L={L} bo={bo} tight={tight} route={route}", flush=True)
    monkeypatch.setenv("SURFHIP_NS_NO_PLANE", "1")
    if route == "axis":
        monkeypatch.setenv("SURFHIP_NS_NO_STRIP", "1")
    test_ns_analyze_synthesize_match_oracle(dims, J, L, bo, tight)
    rng = np.random.default_rng(17)
    x = from_numpy(rng.standard_normal(dims).astype(np.float32), dev)
    t = N.NSSurfacelet(scales=J, direction_levels=list(L), bernstein_order=bo, tight=tight)
    ws_axis = N.make_workspace(t, x)
    monkeypatch.delenv("SURFHIP_NS_NO_PLANE")
    monkeypatch.delenv("SURFHIP_NS_NO_STRIP", raising=False)
    ws_plane = N.make_workspace(t, x)
    c_axis = N.analyze(t, x, workspace=ws_axis)
    c_plane = N.analyze(t, x, workspace=ws_plane)
    assert _rel(to_numpy(c_axis), to_numpy(c_plane)) <= RTOL
    y1 = N.synthesize(t, c_axis, workspace=ws_axis)
    y2 = torch.empty_like(y1)
    N.synthesize_(lambda i: N.subband(t, c_axis, i), y2, t, workspace=ws_axis)
    assert torch.equal(y1, y2)
    assert _rel(to_numpy(N.synthesize(t, c_axis, workspace=ws_plane)), to_numpy(y1)) <= RTOL


@pytest.mark.parametrize("nwork", [5, 1])
@pytest.mark.parametrize("dims,J,L,bo,tight", [CASES[5], CASES[6], CASES[4]])
def test_ns_general_path_several_batches(dims, J, L, bo, tight, nwork, monkeypatch):
    """The general path runs the subbands in batches of at most nwork (whole
    calls up to about 170^3; 15 at 256^3).  With the batch capped
    (SURFHIP_NS_NWORK), small sizes take the several-batch schedule: the
    first batch's synthesis sum starts from zero, later ones add to it, and
    the lowpass wraps into the last batch.  Oracle parity, perfect
    reconstruction and streaming == materialised (in the parity test), and
    the capped plan's results equal the uncapped plan's bit for bit (the same
    per-subband arithmetic, the sums in the same order)."""
    dev = _dev()
    print(f"case dims={dims} J={J} L={L} nwork={nwork}", flush=True)
    monkeypatch.setenv("SURFHIP_NS_NWORK", str(nwork))
    test_ns_analyze_synthesize_match_oracle(dims, J, L, bo, tight)
    rng = np.random.default_rng(19)
    x = from_numpy(rng.standard_normal(dims).astype(np.float32), dev)
    t = N.NSSurfacelet(scales=J, direction_levels=list(L), bernstein_order=bo, tight=tight)
    ws_cap = N.make_workspace(t, x)
    monkeypatch.delenv("SURFHIP_NS_NWORK")
    ws_all = N.make_workspace(t, x)
    c_cap = N.analyze(t, x, workspace=ws_cap)
    c_all = N.analyze(t, x, workspace=ws_all)
    assert torch.equal(c_cap, c_all)
    r = from_numpy(rng.standard_normal(dims + (N.nsubbands(t, dims),)).astype(np.float32), dev)
    assert torch.equal(N.synthesize(t, r, workspace=ws_cap), N.synthesize(t, r, workspace=ws_all))


@pytest.mark.parametrize("dims", [(16, 16, 16), (32, 64, 32)])
def test_ns_streaming_matches_materialised_bitwise(dims):
    dev = _dev()
    t = N.NSSurfacelet(scales=2, direction_levels=[1, 2], bernstein_order=6)
    x = from_numpy(np.random.default_rng(92).standard_normal(dims).astype(np.float32), dev)
    ws = N.make_workspace(t, x)
    ref = N.analyze(t, x, workspace=ws)
    seen = []

    def f(i, band):
        seen.append(i)
        assert torch.equal(band, N.subband(t, ref, i))

    assert N.analyze(f, t, x, workspace=ws) is None
    assert seen == list(range(2, N.nsubbands(t, dims) + 1)) + [1]
    y1 = N.synthesize(t, ref, workspace=ws)
    y2 = torch.empty_like(y1)
    N.synthesize_(lambda i: N.subband(t, ref, i), y2, t, workspace=ws)
    assert torch.equal(y1, y2)
    with pytest.raises(surfhip.DimensionMismatch):
        N.synthesize_(lambda i: torch.zeros(4, 4, 4, device=dev), torch.empty_like(y1), t, workspace=ws)


@pytest.mark.parametrize("dims", [(16, 16, 16), (32, 32, 64)])
def test_ns_complex_is_two_real_channels_bitwise(dims):
    dev = _dev()
    rng = np.random.default_rng(94)
    xr = rng.standard_normal(dims).astype(np.float32)
    xi = rng.standard_normal(dims).astype(np.float32)
    t = N.NSSurfacelet(scales=2, direction_levels=[1, 1], bernstein_order=6)
    z = from_numpy((xr + 1j * xi).astype(np.complex64), dev)
    cc = N.analyze(t, z)
    cr = N.analyze(t, from_numpy(xr, dev))
    ci = N.analyze(t, from_numpy(xi, dev))
    torch.cuda.synchronize()
    assert torch.equal(cc, torch.complex(cr, ci))
    zr = N.synthesize(t, cc)
    assert _rel(to_numpy(zr), xr + 1j * xi) <= RTOL
    yr = N.synthesize(t, cr)
    yi = N.synthesize(t, ci)
    assert torch.equal(zr, torch.complex(yr, yi))


@pytest.mark.parametrize("n", [16, 32])
def test_ns_workspace_reuse_bitwise_and_validated(n):
    dev = _dev()
    t = N.NSSurfacelet(scales=2, direction_levels=[1, 2], bernstein_order=6)
    x = from_numpy(np.random.default_rng(96).standard_normal((n, n, n)).astype(np.float32), dev)
    ws = N.make_workspace(t, torch.float32, (n, n, n))
    a = N.analyze(t, x)
    b = N.analyze(t, x, workspace=ws)
    c = N.analyze(ws, x)
    assert torch.equal(a, b) and torch.equal(a, c)
    with pytest.raises(surfhip.DimensionMismatch):
        N.analyze(t, torch.zeros(8, 8, 8, device=dev), workspace=ws)
    with pytest.raises(surfhip.ArgumentError, match="element type"):
        N.analyze(t, torch.complex(x, x), workspace=ws)
    wsj = N.make_workspace(N.NSSurfacelet(scales=1, direction_levels=[1]), torch.float32, (n, n, n))
    with pytest.raises(surfhip.ArgumentError, match=r"the workspace was built for NSSurfacelet\(scales = 1"):
        N.analyze(t, x, workspace=wsj)
    with pytest.raises(surfhip.ArgumentError, match="tasks must be ≥ 1"):
        N.make_workspace(t, torch.float32, (n, n, n), tasks=0)


@pytest.mark.parametrize("n", [16, 32])
def test_ns_constant_volume_stays_in_lowpass(n):
    dev = _dev()
    t = N.NSSurfacelet(scales=2, direction_levels=[1, 1], bernstein_order=6)
    c = N.analyze(t, from_numpy(np.full((n, n, n), 1.75, np.float32), dev))
    got = to_numpy(c)
    assert np.max(np.abs(got[..., 0] - 1.75)) < 1e-5
    assert np.max(np.abs(got[..., 1:])) < 1e-5


@pytest.mark.parametrize("n", [16, 32])
def test_ns_shift_equivariance(n):
    dev = _dev()
    t = N.NSSurfacelet(scales=2, direction_levels=[1, 2], bernstein_order=6)
    x = np.random.default_rng(91).standard_normal((n, n, n)).astype(np.float32)
    s = (3, -2, 5)
    a = to_numpy(N.analyze(t, from_numpy(x, dev)))
    b = to_numpy(N.analyze(t, from_numpy(np.roll(x, s, axis=(0, 1, 2)), dev)))
    assert np.max(np.abs(b - np.roll(a, s, axis=(0, 1, 2)))) < 1e-5


@pytest.mark.parametrize("dims,route", [((96, 96, 96), "factored"), ((96, 96, 96), "general"),
                                        ((128, 128, 128), "factored")])
def test_ns_default_transform_matches_oracle(dims, route, monkeypatch):
    # NSSurfacelet(scales = 3, direction_levels = [2, 2, 3]), 49 subbands, on
    # Float32 volumes at full size: 96^3 (the reference's CPU-runnable case)
    # on the factored kernels and, forced, on the general-size kernels, and
    # 128^3 (the headline).
    dev = _dev()
    print(f"case {dims} default transform, {route} kernels", flush=True)
    if route == "general":
        monkeypatch.setenv("SURFHIP_NS_GENERIC", "1")
    x = np.random.default_rng(20260).standard_normal(dims).astype(np.float32)
    t = N.NSSurfacelet(scales=3, direction_levels=[2, 2, 3])
    to = O.NSParams(3, (2, 2, 3))
    tb = O.tables(to, dims)
    xd = from_numpy(x, dev)
    ws = N.make_workspace(t, xd)
    c = N.analyze(t, xd, workspace=ws)
    torch.cuda.synchronize()
    ref = O.analyze(to, x.astype(np.float64), tb)
    _check_coeffs(c, ref, np.linalg.norm(x))
    cn = to_numpy(c).astype(np.float64)
    ea = np.sqrt(sum(np.sum((cn[..., i] - ref[i]) ** 2) for i in range(len(ref))) / sum(np.sum(r ** 2) for r in ref))
    del ref
    # Parseval (tight frame) and perfect reconstruction at full size.
    assert abs(np.sum(cn ** 2) / np.sum(x.astype(np.float64) ** 2) - 1) < 1e-5
    xr = N.synthesize(t, c, workspace=ws)
    assert _rel(to_numpy(xr), x) <= RTOL
    # synthesis of the same coefficients against the oracle's synthesis
    yref = O.synthesize(to, [cn[..., i] for i in range(cn.shape[-1])], tb)
    assert _rel(to_numpy(xr), yref) <= RTOL
    print(f"{dims} {route}: analysis rel L2 vs oracle {ea:.2e}, synthesis vs oracle {_rel(to_numpy(xr), yref):.2e}, "
          f"reconstruction {_rel(to_numpy(xr), x):.2e}", flush=True)
    del cn, yref
    # streaming at full size equals the materialised transform bit for bit
    y2 = torch.empty_like(xr)
    N.synthesize_(lambda i: N.subband(t, c, i), y2, t, workspace=ws)
    assert torch.equal(xr, y2)


@pytest.mark.parametrize("dims,factored", [((96, 96, 96), True), ((112, 128, 96), True), ((80, 32, 112), True),
                                          ((128, 64, 32), True), ((80, 96, 72), False), ((48, 48, 48), False)])
def test_ns_kernel_route_by_size(dims, factored, monkeypatch):
    """Grids whose three lengths are in {32, 64, 80, 96, 112, 128} run the
    factored kernels (plane and pencil classes); others, or any grid with
    SURFHIP_NS_GENERIC set, the general-size kernels (rows and axes classes)."""
    dev = _dev()
    t = N.NSSurfacelet(scales=2, direction_levels=[1, 2], bernstein_order=6)
    x = from_numpy(np.random.default_rng(5).standard_normal(dims).astype(np.float32), dev)
    for generic in (False, True):
        if generic:
            monkeypatch.setenv("SURFHIP_NS_GENERIC", "1")
        ws = N.make_workspace(t, x, prewarm=False)
        ws.set_timing(True)
        N.synthesize(ws, N.analyze(ws, x))
        cls = set(ws.timing())
        ws.set_timing(False)
        assert ("plane_inverse" in cls) == (factored and not generic), (dims, generic, cls)
        assert ("rows_c2r" in cls) == (not factored or generic), (dims, generic, cls)
    monkeypatch.delenv("SURFHIP_NS_GENERIC")


@pytest.mark.parametrize("dims,J,L,tight", [((64, 32, 128), 3, (1, 2, 3), True), ((32, 64, 64), 2, (2, 3), False)])
def test_ns_tile_pruning_changes_nothing(dims, J, L, tight, monkeypatch):
    """Tiles whose masks are exactly zero are skipped; the results must equal
    those of the plan that transforms every tile, bit for bit."""
    dev = _dev()
    print(f"case dims={dims} J={J} L={L} tight={tight}", flush=True)
    rng = np.random.default_rng(97)
    x = from_numpy(rng.standard_normal(dims).astype(np.float32), dev)
    t = N.NSSurfacelet(scales=J, direction_levels=list(L), bernstein_order=12, tight=tight)
    ws = N.make_workspace(t, x)
    monkeypatch.setenv("SURFHIP_NS_NO_PRUNE", "1")
    ws_full = N.make_workspace(t, x)
    monkeypatch.delenv("SURFHIP_NS_NO_PRUNE")
    c = N.analyze(t, x, workspace=ws)
    c_full = N.analyze(t, x, workspace=ws_full)
    assert torch.equal(c, c_full)
    r = from_numpy(rng.standard_normal(tuple(dims) + (N.nsubbands(t, dims),)).astype(np.float32), dev)
    assert torch.equal(N.synthesize(t, r, workspace=ws), N.synthesize(t, r, workspace=ws_full))
