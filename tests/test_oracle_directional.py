"""The FP64 restatement of the directional banks (oracle/directional.py) and
the library's host-side masks, on the CPU.

The checks restate the reference's own tests of the banks
(test/items/test_ndfb.jl:1-101, test/items/test_hourglass.jl:59-96): index
round trip and the docstring example (src/directional/ndfb.jl:172-182),
partition of unity of the weights, squares of the tight masks summing to 1,
Parseval and perfect reconstruction, evenness of the masks on the
self-conjugate planes at n = 12, 26, 30, tight reconstruction with Nyquist
bins that are not exactly pi, and the wedges of one hourglass summing to its
weight.  The library's masks (surfhip_ndfb_mask_host, FP64 tables rounded
once to FP32) are compared with the oracle's.
"""
import ctypes

import numpy as np
import pytest

from oracle import directional as D
from oracle import filters as F


def _triples(L):
    na, nb = 1 << F.wedge_levels(L)[0], 1 << F.wedge_levels(L)[1]
    return [(i, p, q) for i in (1, 2, 3) for p in range(1, na + 1) for q in range(1, nb + 1)]


@pytest.mark.parametrize("L", [0, 1, 2, 3])
def test_index_round_trip(L):
    assert D.ndfb_nsubbands(L) == 3 * 2 ** L
    seen = []
    for k, t in enumerate(_triples(L)):
        assert D.ndfb_subband_triple(L, k + 1) == t
        seen.append(k + 1)
    assert sorted(seen) == list(range(1, 3 * 2 ** L + 1))
    with pytest.raises(ValueError):
        D.ndfb_subband_triple(L, 0)


def test_docstring_example():
    # ndfb_subband_index(fb, 2, 1, 2) == 6 and its inverse at levels = 2.
    assert D.ndfb_subband_triple(2, 6) == (2, 1, 2)


@pytest.mark.parametrize("L", [0, 1, 2, 3])
def test_weights_partition_of_unity_and_tight_squares(L):
    dims = (10, 12, 8)
    w = D.ndfb_weights(dims, L, bernstein_order=6)
    assert np.max(np.abs(sum(w) - 1)) < 1e-13
    m = D.ndfb_masks(dims, L, bernstein_order=6, tight=True)
    assert np.max(np.abs(sum(x ** 2 for x in m) - 1)) < 1e-13


@pytest.mark.parametrize("L", [0, 1, 2, 3])
def test_tight_parseval_and_reconstruction(L):
    x = np.random.default_rng(81).standard_normal((16, 16, 16))
    m = D.ndfb_masks(x.shape, L, bernstein_order=6)
    s = D.decompose(x, m)
    assert abs(sum(np.sum(b ** 2) for b in s) / np.sum(x ** 2) - 1) < 1e-12
    assert np.max(np.abs(D.reconstruct(s, m, True) - x)) < 1e-12


def test_loose_bank():
    x = np.random.default_rng(82).standard_normal((16, 16, 16))
    m = D.ndfb_masks(x.shape, 2, bernstein_order=6, tight=False)
    s = D.decompose(x, m)
    assert sum(np.sum(b ** 2) for b in s) / np.sum(x ** 2) < 0.9
    assert np.max(np.abs(D.reconstruct(s, m, False) - x)) < 1e-12


@pytest.mark.parametrize("n", [12, 26, 30])
def test_masks_even_on_self_conjugate_planes(n):
    neg = lambda i: (-i) % n  # noqa: E731
    for M in D.ndfb_masks((n, n, n), 2, bernstein_order=6):
        for i in (0, n // 2):
            P = M[i]
            assert np.max(np.abs(P - P[np.ix_([neg(j) for j in range(n)], [neg(k) for k in range(n)])])) <= 1e-13


@pytest.mark.parametrize("dims", [(30, 16, 16), (16, 30, 16), (16, 16, 30), (26, 26, 26)])
def test_tight_reconstruction_with_inexact_nyquist(dims):
    x = np.random.default_rng(83).standard_normal(dims)
    for L in (1, 2, 3):
        m = D.ndfb_masks(dims, L, bernstein_order=6)
        s = D.decompose(x, m)
        assert np.max(np.abs(D.reconstruct(s, m, True) - x)) < 1e-11


def test_wedges_sum_to_hourglass_weight():
    dims = (12, 12, 12)
    w = D.ndfb_weights(dims, 3, bernstein_order=6)
    hg = D.ndfb_weights(dims, 0, bernstein_order=6)
    for i in (1, 2, 3):
        acc = sum(w[k] for k, t in enumerate(_triples(3)) if t[0] == i)
        assert np.max(np.abs(acc - hg[i - 1])) < 1e-13


surfhip = pytest.importorskip("surfhip")


@pytest.mark.parametrize("dims,L,tight", [((16, 16, 16), 2, True), ((12, 10, 14), 3, False),
                                          ((32, 32, 32), 0, True), ((9, 16, 12), 1, True)])
def test_library_masks_match_oracle(dims, L, tight):
    from surfhip import _lib
    from surfhip.directional import NDFBParams
    p = NDFBParams()
    p.levels, p.bernstein_order, p.halfband_order, p.transition, p.tight = L, 6, 1, 1.0, int(tight)
    ref = D.ndfb_masks(dims, L, bernstein_order=6, tight=tight)
    h = (dims[0] // 2 + 1, dims[1], dims[2])
    for k, R in enumerate(ref):
        out = np.empty(int(np.prod(h)), dtype=np.float32)
        _lib.call("surfhip_ndfb_mask_host", ctypes.byref(p), dims[0], dims[1], dims[2], k + 1,
                  out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
        M = out.reshape(h, order="F")
        assert np.max(np.abs(M - R)) < 2e-7, k
