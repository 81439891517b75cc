"""Pins of the FP64 NSSurfacelet oracle.

No reference-held vector fixes NS coefficient values (the reference's engine
package is absent), so the oracle is pinned by the reference's own invariant
and known-answer tests (test/items/test_hourglass.jl, test_ndfb.jl,
test_ns_surfacelet.jl, test_literature.jl) and by closed forms derived
independently of the code under test.
"""
from math import comb

import numpy as np
import pytest

from oracle import filters as F
from oracle import ns

rng = np.random.default_rng(1234)


# ------------------------------------------------------------- filters ------

def test_bernstein_partition_and_endpoints():
    # test/items/test_hourglass.jl:1-18
    w = np.linspace(0, 2 * np.pi, 33)
    for N in (1, 4, 12):
        tbl = F.bernstein_table(N, w)
        assert tbl.shape == (N + 1, 33)
        assert np.max(np.abs(tbl.sum(axis=0) - 1)) < 1e-12
        assert np.all(tbl >= -1e-15)
    t0 = F.bernstein_table(6, [0.0])
    assert abs(t0[0, 0] - 1) < 1e-15 and np.max(np.abs(t0[1:, 0])) < 1e-14
    tp = F.bernstein_table(6, [np.pi])
    assert abs(tp[-1, 0] - 1) < 1e-15


def test_mapping_coefficients_sign_pattern():
    # test/items/test_hourglass.jl:20-33
    c = F.mapping_coefficients(4, 2)
    for j1 in range(5):
        for j2 in range(5):
            assert c[j1, j2] == (1 if j1 > j2 else (-1 if j1 < j2 else 0))
    c3 = F.mapping_coefficients(3, 3)
    assert c3[3, 0, 0] == 1 and c3[0, 3, 0] == -1 and c3[1, 1, 1] == 0


def test_halfband_split_complementarity():
    # test/items/test_hourglass.jl:35-57
    for m in range(4):
        assert abs(F.halfband_split(1.0, m) - 1) < 1e-15
        assert abs(F.halfband_split(-1.0, m)) < 1e-15
        assert abs(F.halfband_split(0.0, m) - 0.5) < 1e-15
        for x in np.linspace(-1, 1, 21):
            f = F.halfband_split(x, m)
            assert abs(f + F.halfband_split(-x, m) - 1) < 1e-12
            assert -1e-12 <= f <= 1 + 1e-12
    f = [float(F.halfband_split(0.5, m)) for m in range(61)]
    assert f == sorted(f) and all(0 <= v <= 1 for v in f)
    for m in (31, 45, 60):
        for x in np.linspace(-1, 1, 9):
            assert abs(F.halfband_split(x, m) + F.halfband_split(-x, m) - 1) < 1e-12


def test_halfband_split_closed_form():
    # f0(x) = 1 - I_{(1-x)/2}(m+1, m+1): the regularised incomplete beta, i.e.
    # a binomial tail -- an independent closed form of the maximally flat split.
    for m in (0, 1, 3, 7):
        for x in np.linspace(-1, 1, 17):
            p = (1 - x) / 2
            n = 2 * m + 1
            tail = sum(comb(n, k) * p ** k * (1 - p) ** (n - k) for k in range(m + 1))
            assert abs(F.halfband_split(x, m) - tail) < 1e-13


def _binom_pmf(N, x):
    return np.array([comb(N, j) * x ** j * (1 - x) ** (N - j) for j in range(N + 1)])


def test_mapping_response_binomial_closed_form():
    # The Bernstein basis at x = sin^2(w/2) is a Binomial(N, x) pmf, so
    # M1 = P(J1 > max(J2, J3)) - P(J1 < max(J2, J3)) for independent binomials
    # and M2 = P(J2 > J3) - P(J2 < J3): checks bernstein_table,
    # mapping_coefficients and mapping_response together.
    N = 7
    dims = (10, 7, 9)
    b = [F.bernstein_table(N, F.dft_frequencies(dims[0], half=True)),
         F.bernstein_table(N, F.dft_frequencies(dims[1])),
         F.bernstein_table(N, F.dft_frequencies(dims[2]))]
    M1 = F.mapping_response(F.mapping_coefficients(N, 3), b)
    M2 = F.mapping_response(F.mapping_coefficients(N, 2), b[1:])
    w = [F.dft_frequencies(dims[0], half=True), F.dft_frequencies(dims[1]), F.dft_frequencies(dims[2])]
    for _ in range(25):
        k = [rng.integers(len(w[d])) for d in range(3)]
        p = [_binom_pmf(N, np.sin(w[d][k[d]] / 2) ** 2) for d in range(3)]
        c = [np.cumsum(q) for q in p]
        pm = np.diff(np.concatenate([[0.0], c[1] * c[2]]))      # pmf of max(J2, J3)
        gt = np.array([p[0][m + 1:].sum() for m in range(N + 1)])
        lt = np.array([p[0][:m].sum() for m in range(N + 1)])
        assert abs(M1[k[0], k[1], k[2]] - np.sum(pm * (gt - lt))) < 1e-13
        gt2 = np.array([p[1][m + 1:].sum() for m in range(N + 1)])
        lt2 = np.array([p[1][:m].sum() for m in range(N + 1)])
        assert abs(M2[k[1], k[2]] - np.sum(p[2] * (gt2 - lt2))) < 1e-13


def test_rcos_edges_exact_rationals():
    # src/pyramid/fourier_pyramid.jl:47: P = ceil(n/2 (w+t)), Q = floor(n/2 (w-t)).
    from fractions import Fraction as Fr
    assert F.rcos_edges(128, Fr(1, 2), Fr(1, 6)) == (43, 21)
    assert F.rcos_edges(128, Fr(1, 4), Fr(1, 12)) == (22, 10)
    assert F.rcos_edges(96, Fr(1, 2), Fr(1, 6)) == (32, 16)   # exact integers stay put
    prof, P = F.rcos_axis(16, Fr(1, 2), Fr(1, 6))
    assert P == 6 and prof[0] == 1.0 and prof[6] == 0.0 and prof[10] == 0.0


# -------------------------------------------------------- hourglass/ndfb ----

def test_hourglass_weights_partition_of_unity():
    # test/items/test_hourglass.jl:59-96 (partition of unity in [0, 1])
    m1, m2 = F.hourglass_tables((10, 12, 8), 6, 1)
    W = [F.hourglass_weight(m1, m2, i) for i in (1, 2, 3)]
    assert np.max(np.abs(W[0] + W[1] + W[2] - 1)) < 1e-13
    assert all(np.all((w > -1e-14) & (w < 1 + 1e-14)) for w in W)


def _mae(order, n=24, tight=True):
    m1, m2 = F.hourglass_tables((n, n, n), order, 1)
    W = F.hourglass_weight(m1, m2, 1)
    m = F.tight_sqrt(W) if tight else W
    fold = lambda k: np.minimum(2 * np.pi * k / n, 2 * np.pi - 2 * np.pi * k / n)
    p, j, k = np.meshgrid(np.arange(m.shape[0]), np.arange(n), np.arange(n), indexing="ij")
    w1, w2, w3 = fold(p), fold(j), fold(k)
    mx = np.maximum(w2, w3)
    sel = np.abs(w1 - mx) > 1e-9
    return np.mean(np.abs(m[sel] - (w1[sel] > mx[sel])))


def test_hourglass_literature_figures():
    # test/items/test_hourglass.jl:147-175 (Lu & Do 2008, Fig. 10)
    assert _mae(12) < _mae(4) and _mae(20) < _mae(12)
    assert abs(_mae(4, tight=False) - 0.095) <= 0.001
    assert abs(_mae(20, tight=False) - 0.024) <= 0.001


def _ndfb_masks(dims, L, order=12, transition=1.0):
    m1, m2 = F.hourglass_tables(dims, order, 1)
    lev = F.wedge_levels(L)
    out = []
    for i in (1, 2, 3):
        wa, wb = F.wedge_masks(dims, i, lev, 1, transition)
        W = F.hourglass_weight(m1, m2, i)
        for a in wa:
            for b in wb:
                out.append(F.tight_sqrt(W * a * b))
    return out


def test_ndfb_wedge_center_energy_share():
    # test/items/test_literature.jl:13-32 -- share(1.0) = 0.71 at 32^3, k = (8, 4, 4)
    n = 32
    I = np.meshgrid(*(np.arange(n),) * 3, indexing="ij")
    v = np.cos(np.pi * 2 * (8 * I[0] + 4 * I[1] + 4 * I[2]) / n)
    X = ns.rfft3(v)
    e = []
    for m in _ndfb_masks((n, n, n), 2):
        b = ns.irfft3(X * m, (n, n, n))
        e.append(np.sum(b ** 2))
    e = np.array(e)
    # ndfb_subband_index(fb, 1, 2, 2) with na = nb = 2: ((1-1)*2 + 1)*2 + 2 = 4
    assert int(np.argmax(e)) == 3
    assert abs(e.max() / e.sum() - 0.71) <= 0.01


@pytest.mark.parametrize("n", [12, 26, 30])
def test_masks_even_on_self_conjugate_planes(n):
    # test/items/test_ndfb.jl:51-71 -- the mask must be even on the two
    # self-conjugate planes of the rfft grid.
    neg = (-np.arange(n)) % n
    for m in _ndfb_masks((n, n, n), 2, order=6):
        for i in (0, n // 2):
            pl = m[i]
            assert np.max(np.abs(pl - pl[np.ix_(neg, neg)])) <= 1e-13


def test_wedges_sum_to_hourglass_weight():
    # test/items/test_ndfb.jl:90-101
    dims = (12, 12, 12)
    m1, m2 = F.hourglass_tables(dims, 6, 1)
    lev = F.wedge_levels(3)
    for i in (1, 2, 3):
        wa, wb = F.wedge_masks(dims, i, lev, 1, 1.0)
        W = F.hourglass_weight(m1, m2, i)
        acc = sum(W * a * b for a in wa for b in wb)
        assert np.max(np.abs(acc - W)) < 1e-13


# --------------------------------------------------------- NS transform -----

CASES = [((16, 16, 16), 1, (0,)), ((16, 16, 16), 2, (1, 2)), ((12, 20, 8), 3, (1, 1, 2))]


@pytest.mark.parametrize("dims,J,L", CASES)
def test_ns_perfect_reconstruction(dims, J, L):
    # test/items/test_ns_surfacelet.jl:40-54
    t = ns.NSParams(J, L, bernstein_order=6)
    x = rng.standard_normal(dims)
    c = ns.analyze(t, x)
    assert len(c) == 1 + sum(3 * 2 ** l for l in L)
    assert np.max(np.abs(ns.synthesize(t, c) - x)) < 1e-12


@pytest.mark.parametrize("dims", [(16, 16, 16), (15, 20, 8)])
@pytest.mark.parametrize("J,L", [(1, (1,)), (3, (1, 1, 2))])
def test_ns_parseval_and_adjoint(dims, J, L):
    # test/items/test_ns_surfacelet.jl:218-237
    t = ns.NSParams(J, L, bernstein_order=6)
    x = rng.standard_normal(dims)
    c = ns.analyze(t, x)
    assert abs(sum(np.sum(b * b) for b in c) / np.sum(x * x) - 1) < 1e-12
    r = [rng.standard_normal(dims) for _ in c]
    lhs = sum(np.sum(a * b) for a, b in zip(c, r))
    rhs = np.sum(x * ns.synthesize(t, r))
    assert abs(lhs - rhs) <= 1e-12 * abs(rhs)


def test_ns_shift_equivariance():
    # test/items/test_ns_surfacelet.jl:56-67
    t = ns.NSParams(2, (1, 2), bernstein_order=6)
    x = rng.standard_normal((16, 16, 16))
    s = (3, -2, 5)
    a = ns.analyze(t, x)
    b = ns.analyze(t, np.roll(x, s, axis=(0, 1, 2)))
    for u, v in zip(a, b):
        assert np.max(np.abs(v - np.roll(u, s, axis=(0, 1, 2)))) < 1e-10


def test_ns_constant_goes_to_lowpass():
    # test/items/test_ns_surfacelet.jl:69-77
    t = ns.NSParams(2, (1, 1), bernstein_order=6)
    c = ns.analyze(t, np.full((16, 16, 16), 1.75))
    assert np.max(np.abs(c[0] - 1.75)) < 1e-12
    assert all(np.max(np.abs(b)) < 1e-12 for b in c[1:])


def test_ns_empty_scale_refused():
    # test/items/test_ns_surfacelet.jl:239-252
    t = ns.NSParams(4, (0, 0, 0, 0), bernstein_order=4)
    with pytest.raises(ValueError, match="too small for 4 scales"):
        ns.tables(t, (8, 8, 8))
    q = ns.NSParams(3, (0, 1, 1), bernstein_order=4)
    x = rng.standard_normal((32, 32, 8))
    c = ns.analyze(q, x)
    assert np.max(np.abs(ns.synthesize(q, c) - x)) < 1e-12


def test_ns_not_tight_reconstructs():
    t = ns.NSParams(2, (1, 2), bernstein_order=6, tight=False)
    x = rng.standard_normal((16, 12, 20))
    assert np.max(np.abs(ns.synthesize(t, ns.analyze(t, x)) - x)) < 1e-12


def test_ns_dense_analysis_matrix_is_an_isometry():
    # Brute force on a tiny odd-sized grid: the analysis matrix A of a tight
    # transform satisfies A^T A = I, and synthesize is A^T column by column.
    t = ns.NSParams(2, (1, 1), bernstein_order=4)
    dims = (5, 4, 6)
    N = int(np.prod(dims))
    cols = []
    tb = ns.tables(t, dims)
    for i in range(N):
        e = np.zeros(N)
        e[i] = 1
        cols.append(np.concatenate([b.ravel() for b in ns.analyze(t, e.reshape(dims), tb)]))
    A = np.array(cols).T
    assert np.max(np.abs(A.T @ A - np.eye(N))) < 1e-12
    y = rng.standard_normal(A.shape[0])
    S = len(ns.analyze(t, np.zeros(dims), tb))
    parts = [p.reshape(dims) for p in np.split(y, S)]
    assert np.max(np.abs(ns.synthesize(t, parts, tb).ravel() - A.T @ y)) < 1e-12


def test_ns_scales_run_coarse_to_fine():
    # test/items/test_ns_surfacelet.jl:326-339
    n = 16
    t = ns.NSParams(2, (0, 1), bernstein_order=6)
    lay = ns.subband_layout(t)
    I = np.meshgrid(*(np.arange(n),) * 3, indexing="ij")
    fast = np.cos(np.pi * 2 * (7 * I[0] + 6 * I[1]) / n)
    slow = np.cos(np.pi * 2 * I[0] / n)
    energy = lambda x: [np.sum(b ** 2) for b in ns.analyze(t, x)]
    assert lay[int(np.argmax(energy(fast)))][0] == 2
    assert lay[int(np.argmax(energy(slow)))][0] in (0, 1)


def test_ns_complex_is_two_real_channels():
    # test/items/test_ns_surfacelet.jl:140-152
    t = ns.NSParams(2, (1, 1), bernstein_order=6)
    xr = rng.standard_normal((16, 16, 16))
    xi = rng.standard_normal((16, 16, 16))
    cc = ns.analyze(t, xr + 1j * xi)
    cr, ci = ns.analyze(t, xr), ns.analyze(t, xi)
    assert all(np.array_equal(a, b + 1j * c) for a, b, c in zip(cc, cr, ci))
