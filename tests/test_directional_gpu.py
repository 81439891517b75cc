"""The hourglass and directional banks on the GPU (surfhip.directional).

Restates test/items/test_gpu.jl:57-98: at 16³ with bernstein_order = 6 the
device subbands match the CPU bank's (here the FP64 oracle) within 1e-4 and
reconstruction returns the input within 1e-4.  Also: power-of-two sizes on
the factored kernels, a bank that is not tight, complex volumes as two real
passes, the streaming forms equal to the batched ones, and the argument and
size errors of the reference.
"""
import numpy as np
import pytest
import torch

from oracle import directional as D

surfhip = pytest.importorskip("surfhip")
from surfhip import directional as B  # noqa: E402
from surfhip._arrays import from_numpy, to_numpy  # noqa: E402
from surfhip._lib import ArgumentError, DimensionMismatch  # noqa: E402


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return "cuda"


@pytest.mark.gpu
def test_hourglass_matches_oracle_16():
    dev = _dev()
    x = np.random.default_rng(200).standard_normal((16, 16, 16)).astype(np.float32)
    fb = B.make_hourglass(torch.float32, x.shape, bernstein_order=6)
    s = B.hourglass_decompose(from_numpy(x, dev), fb)
    ref = D.decompose(x.astype(np.float64), D.ndfb_masks(x.shape, 0, bernstein_order=6))
    assert len(s) == 3
    for i in range(3):
        assert np.max(np.abs(to_numpy(s[i]) - ref[i])) < 1e-4
    y = B.hourglass_reconstruct(s, fb)
    assert np.max(np.abs(to_numpy(y) - x)) < 1e-4


@pytest.mark.gpu
@pytest.mark.parametrize("dims,L,tight", [((16, 16, 16), 2, True), ((32, 32, 32), 3, True),
                                          ((32, 64, 32), 2, False), ((20, 18, 24), 1, True)])
def test_ndfb_matches_oracle(dims, L, tight):
    dev = _dev()
    x = np.random.default_rng(201).standard_normal(dims).astype(np.float32)
    fb = B.make_ndfb(torch.float32, dims, levels=L, bernstein_order=6, tight=tight)
    s = B.ndfb_decompose(from_numpy(x, dev), fb)
    assert len(s) == B.ndfb_nsubbands(fb) == 3 * 2 ** L
    ref = D.decompose(x.astype(np.float64), D.ndfb_masks(dims, L, bernstein_order=6, tight=tight))
    for k in range(len(s)):
        assert np.max(np.abs(to_numpy(s[k]) - ref[k])) < 1e-4, k
    y = B.ndfb_reconstruct(s, fb)
    assert np.max(np.abs(to_numpy(y) - x)) < 1e-4


@pytest.mark.gpu
def test_streaming_forms_and_complex_volumes():
    dev = _dev()
    dims = (32, 32, 32)
    rng = np.random.default_rng(202)
    xr = rng.standard_normal(dims).astype(np.float32)
    xi = rng.standard_normal(dims).astype(np.float32)
    fb = B.make_ndfb(torch.float32, dims, levels=2, bernstein_order=6)
    s = B.ndfb_decompose(from_numpy(xr, dev), fb)
    got = {}
    B.ndfb_decompose(lambda idx, band: got.__setitem__(idx, band.clone()), from_numpy(xr, dev), fb)
    assert sorted(got) == list(range(1, 13))
    for k in range(12):
        assert torch.equal(got[k + 1], s[k])
    y = torch.empty_like(from_numpy(xr, dev))
    B.ndfb_reconstruct_(lambda k: s[k - 1], y, fb)
    assert torch.equal(y, B.ndfb_reconstruct(s, fb))
    # Complex data: T(x + i y) = T(x) + i T(y) exactly (two real passes).
    fc = B.make_ndfb(torch.complex64, dims, levels=2, bernstein_order=6)
    sc = B.ndfb_decompose(from_numpy(xr + 1j * xi, dev), fc)
    si = B.ndfb_decompose(from_numpy(xi, dev), fb)
    for k in range(12):
        assert torch.equal(sc[k].real, s[k]) and torch.equal(sc[k].imag, si[k])
    yc = B.ndfb_reconstruct(sc, fc)
    assert np.max(np.abs(to_numpy(yc) - (xr + 1j * xi))) < 1e-4


@pytest.mark.gpu
def test_bank_errors():
    dev = _dev()
    fb = B.make_ndfb(torch.float32, (16, 16, 16), levels=2, bernstein_order=4)
    with pytest.raises(ArgumentError):
        B.ndfb_subband_index(fb, 4, 1, 1)
    with pytest.raises(ArgumentError):
        B.ndfb_subband_index(fb, 1, 99, 1)
    with pytest.raises(ArgumentError):
        B.ndfb_subband_triple(fb, 0)
    assert B.ndfb_subband_index(fb, 2, 1, 2) == 6 and B.ndfb_subband_triple(fb, 6) == (2, 1, 2)
    with pytest.raises(DimensionMismatch):
        B.ndfb_decompose(torch.zeros((16, 16, 8), device=dev), fb)
    s = B.ndfb_decompose(torch.zeros((16, 16, 16), device=dev), fb)
    with pytest.raises(DimensionMismatch):
        B.ndfb_reconstruct(s[:-1], fb)
    hg = B.make_hourglass(torch.float32, (16, 16, 16), bernstein_order=4)
    with pytest.raises(ArgumentError):
        B.hourglass_mask(hg, 4)
    w = sum(B.hourglass_weight(hg, i) for i in (1, 2, 3))
    assert np.max(np.abs(w - 1)) < 1e-6
    shared = B.ndfb_sharing(hg, levels=3)
    assert B.ndfb_nsubbands(shared) == 24


def test_bank_argument_errors_on_the_host():
    with pytest.raises(ArgumentError):
        B.make_ndfb(torch.float32, (8, 8, 8), levels=-1)
    with pytest.raises(ArgumentError):
        B.make_ndfb(torch.float32, (8, 8, 8), transition=-1.0)
