"""The FP64 oracle of the decimated Surfacelet against SurfBox's golden data.

Restates test/items/test_irc.jl:68-193 on the oracle: every stage (checkerboard
filters, ritf hourglass, IRC, NDFB scale, pyramid, whole transform) must match
the reference-held SurfBox fingerprints at the reference's tolerances.
"""
import numpy as np
import pytest

from oracle import decimated as D
from oracle import filters as F
from oracle.golden import fp_match, load_golden
from oracle.golden import testvol as make_testvol

G = load_golden()
L111 = np.array([[-1, 1, 1], [1, -1, 1], [1, 1, -1]])
L212 = np.array([[-1, 2, 1], [1, -1, 2], [2, 1, -1]])


@pytest.mark.parametrize("bo", [4, 8, 12])
def test_checkerboard_filters_match_surfbox_dtft(bo):
    # test/items/test_irc.jl:68-90
    t = F.CBD_LADDER[bo]
    ws = np.linspace(0, 2 * np.pi, 13)[:12]

    def dtft(H, c, w1, w2):
        i = np.arange(1, H.shape[0] + 1)[:, None]
        j = np.arange(1, H.shape[1] + 1)[None, :]
        return np.sum(H * np.exp(-1j * (w1 * (i - c[0, 0]) + w2 * (j - c[0, 1]))))

    worst = 0.0
    for a in ws:
        for b in ws:
            for ch in (0, 1):
                tag = "H0" if ch == 0 else "H1"
                ctag = "c0" if ch == 0 else "c1"
                hd = dtft(G[f"cbd_{bo}_dec_{tag}"], G[f"cbd_{bo}_dec_{ctag}"], a, b)
                hr = dtft(G[f"cbd_{bo}_rec_{tag}"], G[f"cbd_{bo}_rec_{ctag}"], a, b)
                worst = max(worst, abs(F.cbd_analysis(t, ch, a, b) - hd), abs(F.cbd_synthesis(t, ch, a, b) - hr))
    assert worst < 1.0e-13


def test_ritf_and_irc_match_surfbox():
    # test/items/test_irc.jl:92-115
    x = make_testvol(32)
    X = np.fft.fftn(x.astype(np.complex128))
    V = F.ritf_masks(x.shape)
    for K in range(3):
        hd = D.half_dims(x.shape, K)
        hg = X[:hd[0], :hd[1], :hd[2]] * V[K]
        assert fp_match(hg, G[f"hg_{K + 1}"], rtol=1.0e-10)
        for nm, L in (("l111", L111), ("l212", L212)):
            levels = tuple(0 if d == K else int(L[K, d]) for d in range(3))
            st = D.irc_stages(hg.shape, 32, K, levels, 4)
            Sk = hg.copy()
            D.irc_analysis(Sk, st)
            assert fp_match(Sk, G[f"irc_{nm}_{K + 1}"], rtol=1.0e-10)
            D.irc_synthesis(Sk, st)
            assert np.max(np.abs(Sk - hg)) < 1.0e-12 * np.max(np.abs(hg))
    Vf = F.ritf_masks((16, 20, 12))
    tot = Vf[0][:9, :11, :7] ** 2 + Vf[1][:9, :11, :7] ** 2 + Vf[2][:9, :11, :7] ** 2
    assert np.max(np.abs(tot - 3)) < 1.0e-12


@pytest.mark.parametrize("nm,L", [("l111", L111), ("l212", L212)])
def test_ndfb_scale_matches_surfbox(nm, L):
    # test/items/test_irc.jl:117-137
    x = make_testvol(32)
    lv = D.surf_levels(L)
    masks = F.ritf_masks(x.shape)
    sc = D.SurfScale(x.shape, lv, masks, D.surf_stages(x.shape, lv, 4))
    bands = D.scale_forward(np.fft.fftn(x.astype(np.complex128)), sc)
    i = 0
    for K in range(3):
        for w in range(1 << sum(lv[K])):
            assert fp_match(bands[i], G[f"ndfb_{nm}_{K + 1}_{w + 1}"])
            i += 1
    rec = np.real(np.fft.ifftn(D.scale_inverse(bands, sc)))
    assert fp_match(rec, G[f"ndfbrec_{nm}"])
    assert np.max(np.abs(rec - x)) < 1.0e-13


@pytest.mark.parametrize("tag,n,mode", [("m15", 48, 1.5), ("m1", 32, 1), ("m2", 32, 2)])
def test_pyramid_matches_surfbox(tag, n, mode):
    # test/items/test_irc.jl:139-161
    x = make_testvol(n)
    levels = D.fourier_pyramid(x.shape, 2, mode)
    X = np.fft.fftn(x.astype(np.complex128))
    orig = X.copy()
    highs = []
    for lv in levels:
        low = D.pyramid_split(X, lv)
        highs.append(X)
        X = low
    assert fp_match(highs[0], G[f"pyr_{tag}_1"])
    assert fp_match(highs[1], G[f"pyr_{tag}_2"])
    assert fp_match(X, G[f"pyr_{tag}_3"])
    for k in (1, 0):
        X = D.pyramid_merge(highs[k], X, levels[k])
    assert np.max(np.abs(X - orig)) < 1.0e-12 * np.max(np.abs(orig))


@pytest.mark.parametrize("tag,n,mode,lev_array,bo", [
    ("m15", 48, 1.5, [L111, L111], 8),
    ("m2", 64, 2, [L212, L111], 4),
])
def test_surfacelet_matches_surfdec_surfrec(tag, n, mode, lev_array, bo):
    # test/items/test_irc.jl:163-184
    x = make_testvol(n)
    t = D.SParams(scales=2, direction_levels=tuple(lev_array[::-1]), pyramid_ratio=mode, checkerboard_order=bo)
    c = D.analyze(t, x)
    lay = D.subband_layout(t, x.shape)
    for i, (s, K, w) in enumerate(lay):
        if i == 0:
            continue
        key = f"surf_{tag}_{t.scales - s + 1}_{K}_{w}"
        assert fp_match(c[i], G[key]), key
    assert fp_match(c[0], G[f"surf_{tag}_low"])
    r = D.synthesize(t, c, dims=x.shape)
    assert fp_match(r, G[f"surfrec_{tag}"])
    assert np.max(np.abs(r - x)) < 1.0e-13


def test_redundancy_matches_surfbox():
    # test/items/test_irc.jl:186-193
    for nm, Lm, J in (("d24", [[-1, 1, 2], [2, -1, 1], [1, 2, -1]], 3), ("d96", [[-1, 2, 3], [3, -1, 2], [2, 3, -1]], 2)):
        t = D.SParams(scales=J, direction_levels=tuple([np.array(Lm)] * J), checkerboard_order=4)
        assert abs(D.redundancy(t, (96, 96, 96)) - G[f"red_96_{nm}"][0, 0]) < 1e-12


@pytest.mark.parametrize("dims,mode,J,L,bo", [
    ((24, 24, 24), 1.5, 1, [1], 12),
    ((32, 32, 32), 2, 2, [1, 2], 4),
    ((32, 32, 32), 1, 2, [1, 1], 8),
    ((48, 24, 36), 1.5, 1, [2], 8),
])
def test_adjoint_dense_probes(dims, mode, J, L, bo):
    # The decimated transform is not tight, so its adjoint is not synthesize
    # (src/transforms/surfacelet.jl:24-26): check <A x, y> = <x, A^T y> on
    # random probes, and that A^T differs from synthesize.
    rng = np.random.default_rng(11)
    t = D.SParams(scales=J, direction_levels=tuple(L), pyramid_ratio=mode, checkerboard_order=bo)
    ws = D.workspace(t, dims)
    for _ in range(3):
        x = rng.standard_normal(dims)
        c = D.analyze(t, x, ws)
        y = [rng.standard_normal(b.shape) for b in c]
        lhs = sum(np.sum(a * b) for a, b in zip(c, y))
        rhs = np.sum(x * D.adjoint(t, y, ws))
        assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), 1.0)
    ys = D.synthesize(t, y, ws)
    assert np.max(np.abs(ys - D.adjoint(t, y, ws))) > 1e-6


def test_adjoint_dense_matrix_tiny():
    # Brute force: the full analysis matrix at 12^3 (mode 1.5, one scale),
    # column by column, against the adjoint applied to unit coefficient vectors.
    t = D.SParams(scales=1, direction_levels=(1,), pyramid_ratio=1.5, checkerboard_order=4)
    dims = (12, 12, 12)
    ws = D.workspace(t, dims)
    sizes = D.layout_sizes(t, dims)
    ncoef = sum(int(np.prod(s)) for s in sizes)
    rng = np.random.default_rng(5)
    picks = rng.choice(ncoef, size=40, replace=False)
    N = int(np.prod(dims))
    cols = rng.choice(N, size=40, replace=False)
    A = np.zeros((40, 40))
    for jj, col in enumerate(cols):
        e = np.zeros(N)
        e[col] = 1
        flat = np.concatenate([b.ravel(order="F") for b in D.analyze(t, e.reshape(dims, order="F"), ws)])
        A[:, jj] = flat[picks]
    offs = np.cumsum([0] + [int(np.prod(s)) for s in sizes])
    for ii, row in enumerate(picks):
        flat = np.zeros(ncoef)
        flat[row] = 1
        c = [flat[offs[q]:offs[q + 1]].reshape(sizes[q], order="F") for q in range(len(sizes))]
        at = D.adjoint(t, c, ws).ravel(order="F")
        assert np.max(np.abs(at[cols] - A[ii, :])) < 1e-13
