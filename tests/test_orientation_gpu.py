"""Subband orientation of the device path: which subband is which direction,
as the HIP kernels compute it.

Restates test/items/test_hourglass.jl:113-146, test/items/test_ndfb.jl:
118-147 and test/items/test_surfbox_reference.jl:17-155 through the C ABI
(surfhip_ndfb_plan_create banks and surfhip_ns_analyze), and extends the
plane-wave prediction to the NSSurfacelet's finest scale on the factored
kernels (32^3, 128^3) and the general-size kernels (48^3).  Every check
compares the device result with a prediction from the wave's frequency or
with SurfBox's fixtures, not with the oracle.
"""
import numpy as np
import pytest
import torch

import orientation_cases as C

surfhip = pytest.importorskip("surfhip")
from surfhip import directional as B  # noqa: E402
from surfhip import ns as NS  # noqa: E402
from surfhip._arrays import from_numpy, to_numpy  # noqa: E402


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return "cuda"


def _energies(bands):
    e = torch.stack([torch.sum(b.double() ** 2) for b in bands])
    return (e / e.sum()).cpu().numpy()


def _wave(n, k, dev):
    return from_numpy(C.plane_wave(n, k).astype(np.float32), dev)


@pytest.mark.gpu
def test_hourglass_axis_and_diagonal_waves_on_device():
    dev = _dev()
    n = 16
    fb = B.make_hourglass(torch.float32, (n, n, n), bernstein_order=8, halfband_order=1)
    for ax in (1, 2, 3):
        e = _energies(B.hourglass_decompose(from_numpy(C.axis_wave(n, ax, 6).astype(np.float32), dev), fb))
        assert e[ax - 1] > 0.999, (ax, e)
        assert sum(e[j] for j in range(3) if j != ax - 1) < 1e-3, (ax, e)
    e = _energies(B.hourglass_decompose(_wave(n, [5, 5, 0], dev), fb))
    assert abs(e[0] - e[1]) < 2e-6, e
    assert e[2] < 2e-6, e


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["L4", "L3"])
def test_ndfb_plane_waves_land_in_predicted_wedge_on_device(case):
    dev = _dev()
    cs = C.NDFB_L4 if case == "L4" else C.NDFB_L3
    n, L = cs["n"], cs["levels"]
    la, lb = (L + 1) // 2, L // 2
    fb = B.make_ndfb(torch.float32, (n, n, n), levels=L, bernstein_order=cs["bernstein_order"],
                     transition=cs["transition"])
    for k in cs["ks"]:
        e = _energies(B.ndfb_decompose(_wave(n, k, dev), fb))
        want = B.ndfb_subband_index(fb, *C.predict(k, la, lb))
        assert int(np.argmax(e)) + 1 == want, (k, want, int(np.argmax(e)) + 1)
        assert e[want - 1] > 0.3, (k, e[want - 1])


@pytest.mark.gpu
@pytest.mark.parametrize("n", [32, 128, 48])
def test_ns_finest_scale_plane_waves_on_device(n):
    dev = _dev()
    ks = dict(C.ns_finest_cases())[n]
    levels = C.NS_FINEST["direction_levels"]
    t = NS.NSSurfacelet(scales=C.NS_FINEST["scales"], direction_levels=levels)
    ws = NS.make_workspace(t, torch.float32, (n, n, n), prewarm=False)
    off = C.ns_finest_offset(levels)
    la, lb = (levels[-1] + 1) // 2, levels[-1] // 2
    S = NS.nsubbands(t, (n, n, n))
    for k in ks:
        c = NS.analyze(ws, _wave(n, k, dev))
        e = _energies([c[:, :, :, i] for i in range(S)])
        want = off + C.subband_index(*C.predict(k, la, lb), la, lb)
        assert e[off:].sum() > 0.99999, (n, k, e[off:].sum())
        assert int(np.argmax(e)) + 1 == want, (n, k, want, int(np.argmax(e)) + 1)
        assert e[want - 1] > 0.3, (n, k, e[want - 1])


@pytest.mark.gpu
def test_hourglass_matches_surfbox_partition_on_device():
    # The device's weights: the untight bank applied to a unit impulse at the
    # origin gives subband k = irfft(W_k), so rfft of it is W_k as the kernels
    # evaluate it.
    dev = _dev()
    w_ref, _ = C.surfbox_hourglass_weights()
    fb = B.make_hourglass(torch.float32, (32, 32, 32), tight=False)
    d = np.zeros((32, 32, 32), dtype=np.float32)
    d[0, 0, 0] = 1
    s = B.hourglass_decompose(from_numpy(d, dev), fb)
    w_dev = [np.fft.rfftn(to_numpy(b).astype(np.float64), axes=(2, 1, 0)).real for b in s]
    fails = C.check_surfbox_hourglass(w_dev, w_ref, atol_unity=1e-6)
    assert not fails, fails


@pytest.mark.gpu
def test_ndfb_matches_surfbox_direction_partition_on_device():
    dev = _dev()
    dirs, labels = C.surfbox_directions()
    fb = B.make_ndfb(torch.float32, (32, 32, 32), levels=2)
    E = np.stack([_energies(B.ndfb_decompose(from_numpy(C.surfbox_direction_wave(dirs[:, t]).astype(np.float32),
                                                        dev), fb)) for t in range(80)])
    fails = C.check_surfbox_directions(E, labels)
    assert not fails, fails
