"""Subband orientation, pinned on the CPU: which subband is which direction.

The reference's orientation tests (test/items/test_hourglass.jl:113-146,
test/items/test_ndfb.jl:118-147, test/items/test_surfbox_reference.jl:
17-155) are restated on two independent mask builders:
- the FP64 oracle (oracle/filters.py, oracle/directional.py, oracle/ns.py);
- the library's host-built FP32 masks (surfhip_ndfb_mask_host and
  surfhip_ns_mask_host: the same table code and expressions the device
  evaluates, csrc/ns_tables.cpp).
Each case compares the result with a prediction made from the wave's
frequency alone, or with SurfBox's fixtures, never with the other builder, so
an axis or ordering mistake shared by both still fails.  The device path's
own orientation is tested in tests/test_orientation_gpu.py.
"""
import ctypes

import numpy as np
import pytest

import orientation_cases as C
from oracle import directional as D
from oracle import filters as F
from oracle import ns as O


def _lib_ndfb_masks(dims, levels, bernstein_order=12, halfband_order=1, transition=1.0, tight=True):
    surfhip = pytest.importorskip("surfhip")  # noqa: F841
    from surfhip import _lib
    from surfhip.directional import NDFBParams
    p = NDFBParams()
    p.levels, p.bernstein_order, p.halfband_order = levels, bernstein_order, halfband_order
    p.transition, p.tight = transition, int(tight)
    h = (dims[0] // 2 + 1, dims[1], dims[2])
    out = []
    for k in range(1, 3 * (1 << levels) + 1):
        m = np.empty(int(np.prod(h)), dtype=np.float32)
        _lib.call("surfhip_ndfb_mask_host", ctypes.byref(p), dims[0], dims[1], dims[2], k,
                  m.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
        out.append(m.reshape(h, order="F").astype(np.float64))
    return out


def _lib_ns_masks(t, dims):
    surfhip = pytest.importorskip("surfhip")  # noqa: F841
    from surfhip import ns as NS
    nt = NS.NSSurfacelet(scales=t.scales, direction_levels=list(t.direction_levels),
                         bernstein_order=t.bernstein_order, halfband_order=t.halfband_order,
                         transition=t.transition, tight=t.tight)
    S = O.nsubbands(t)
    return [NS.ns_mask_host(nt, dims, i).astype(np.float64) for i in range(1, S + 1)]


def _oracle_ndfb_masks(dims, levels, bernstein_order=12, halfband_order=1, transition=1.0, tight=True):
    return D.ndfb_masks(dims, levels, bernstein_order=bernstein_order, halfband_order=halfband_order,
                        transition=transition, tight=tight)


BUILDERS = {"oracle": _oracle_ndfb_masks, "library": _lib_ndfb_masks}


def _shares(x, masks):
    return C.energy_shares(D.decompose(x, masks))


@pytest.mark.parametrize("builder", sorted(BUILDERS))
def test_hourglass_axis_waves_land_in_their_hourglass(builder):
    # test/items/test_hourglass.jl:113-128
    n = 16
    m = BUILDERS[builder]((n, n, n), 0, bernstein_order=8, halfband_order=1)
    for ax in (1, 2, 3):
        e = _shares(C.axis_wave(n, ax, 6), m)
        assert e[ax - 1] > 0.999, (ax, e)
        assert sum(e[j] for j in range(3) if j != ax - 1) < 1e-3, (ax, e)


@pytest.mark.parametrize("builder", sorted(BUILDERS))
def test_hourglass_diagonal_wave_splits_evenly(builder):
    # test/items/test_hourglass.jl:130-146.  The library's masks are FP32, so
    # its even split holds to FP32 rounding of the squared masks.
    n = 16
    m = BUILDERS[builder]((n, n, n), 0, bernstein_order=8, halfband_order=1)
    e = _shares(C.plane_wave(n, [5, 5, 0]), m)
    tol = 1e-6 if builder == "oracle" else 2e-6
    assert abs(e[0] - e[1]) < tol, e
    assert e[2] < tol, e


@pytest.mark.parametrize("builder", sorted(BUILDERS))
@pytest.mark.parametrize("case", ["L4", "L3"])
def test_ndfb_plane_waves_land_in_predicted_wedge(builder, case):
    # test/items/test_ndfb.jl:118-147 (levels 4), plus levels 3.
    cs = C.NDFB_L4 if case == "L4" else C.NDFB_L3
    n, L = cs["n"], cs["levels"]
    la, lb = F.wedge_levels(L)
    m = BUILDERS[builder]((n, n, n), L, bernstein_order=cs["bernstein_order"], transition=cs["transition"])
    assert len(m) == 3 * 2 ** L
    for k in cs["ks"]:
        e = _shares(C.plane_wave(n, k), m)
        want = C.subband_index(*C.predict(k, la, lb), la, lb)
        assert int(np.argmax(e)) + 1 == want, (k, want, int(np.argmax(e)) + 1)
        assert e[want - 1] > 0.3, (k, e[want - 1])


@pytest.mark.parametrize("builder", ["oracle", "library"])
def test_ns_finest_scale_plane_waves_land_in_predicted_subband(builder):
    # The NSSurfacelet's own masks (group factor x wedges), headline
    # parameters at 32^3: each wave lies wholly in the finest scale and its
    # energy peaks in subband offsets_J + ns_index(predicted i, p, q).
    cs = C.NS_FINEST
    n = cs["n"]
    t = O.NSParams(cs["scales"], tuple(cs["direction_levels"]))
    masks = O.subband_masks(t, (n, n, n)) if builder == "oracle" else _lib_ns_masks(t, (n, n, n))
    off = C.ns_finest_offset(cs["direction_levels"])
    la, lb = F.wedge_levels(cs["direction_levels"][-1])
    for k in cs["ks"]:
        e = _shares(C.plane_wave(n, k), masks)
        want = off + C.subband_index(*C.predict(k, la, lb), la, lb)
        assert e[off:].sum() > 0.999999, (k, e[off:].sum())
        assert int(np.argmax(e)) + 1 == want, (k, want, int(np.argmax(e)) + 1)
        assert e[want - 1] > 0.3, (k, e[want - 1])


@pytest.mark.parametrize("builder", sorted(BUILDERS))
def test_hourglass_matches_surfbox_partition(builder):
    # test/items/test_surfbox_reference.jl:17-92: the weights W_i (the
    # untight masks) against SurfBox's V_k^2 / 3 from the committed fixture.
    w_ref, tight = C.surfbox_hourglass_weights()
    assert np.all(np.abs(tight - 3.0) <= 1e-5)
    w_mine = BUILDERS[builder]((32, 32, 32), 0, tight=False)
    assert all(w.shape == (17, 32, 32) for w in w_mine)
    fails = C.check_surfbox_hourglass(w_mine, w_ref, atol_unity=1e-12 if builder == "oracle" else 1e-6)
    assert not fails, fails
    # The tight masks are the weights' square roots (Parseval), and the loose
    # bank's energy runs from 1/3 to 1 (test_surfbox_reference.jl:79-91).
    tm = BUILDERS[builder]((32, 32, 32), 0, tight=True)
    tol = 1e-12 if builder == "oracle" else 1e-6
    assert np.max(np.abs(sum(m ** 2 for m in tm) - 1)) < tol
    energy = sum(m ** 2 for m in w_mine)
    assert abs(energy.min() - 1 / 3) < 1e-3
    assert abs(energy.max() - 1) < tol


@pytest.mark.parametrize("builder", sorted(BUILDERS))
def test_ndfb_matches_surfbox_direction_partition(builder):
    # test/items/test_surfbox_reference.jl:101-155: levels 2 at 32^3.
    dirs, labels = C.surfbox_directions()
    assert labels.size == 80
    m = BUILDERS[builder]((32, 32, 32), 2)
    assert len(m) == 12
    E = np.array([[float(np.sum(b ** 2)) for b in D.decompose(C.surfbox_direction_wave(dirs[:, t]), m)]
                  for t in range(80)])
    fails = C.check_surfbox_directions(E, labels)
    assert not fails, fails
