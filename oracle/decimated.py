"""FP64 restatement of the decimated Surfacelet (test infrastructure only).

Follows src/transforms/surfacelet.jl, src/directional/surfacelet_scale.jl,
src/directional/irc.jl and src/pyramid/fourier_pyramid.jl (cited per
function).  Volumes are numpy arrays indexed [i1, i2, i3]; all index tables
here are 0-based (the reference's are 1-based).
"""
from dataclasses import dataclass
from fractions import Fraction

import numpy as np

from . import filters as F

MODES = {
    1.0: ((Fraction(1, 2), Fraction(1, 4)), (Fraction(1, 6), Fraction(1, 12)), (Fraction(1), Fraction(2))),
    1.5: ((Fraction(1, 2), Fraction(3, 8)), (Fraction(1, 7), Fraction(1, 9)), (Fraction(3, 2), Fraction(2))),
    2.0: ((Fraction(1, 3), Fraction(1, 3)), (Fraction(1, 7), Fraction(1, 7)), (Fraction(2), Fraction(2))),
}


def pyramid_mode(mode):
    """src/pyramid/fourier_pyramid.jl:35-40."""
    if float(mode) not in MODES:
        raise ValueError(f"pyramid_ratio must be 1, 1.5 or 2, got {mode}")
    return MODES[float(mode)]


@dataclass
class PyramidLevel:
    """src/pyramid/fourier_pyramid.jl:26-33."""
    n: tuple
    nlow: tuple
    D: Fraction
    idx: tuple        # per axis, 0-based bins of n inside the lowpass support
    idx_low: tuple    # where they land in nlow
    prof: tuple       # per axis, the 1-D profile on idx


def pyramid_level(n, w, tbw, D):
    """src/pyramid/fourier_pyramid.jl:68-85."""
    if not all((d * D.denominator) % D.numerator == 0 for d in n):
        raise ValueError(f"a pyramid level with factor {D} cannot divide size {n}")
    nlow = tuple(d * D.denominator // D.numerator for d in n)
    idx, idx_low, prof = [], [], []
    for a in range(3):
        nr = n[a]
        full, P = F.rcos_axis(nr, w, tbw)
        if 2 * P > nlow[a]:
            raise ValueError(f"size {nr} is too small for the pyramid")
        ia = np.array(list(range(0, P + 1)) + list(range(nr - P, nr)))
        il = np.array(list(range(0, P + 1)) + list(range(nlow[a] - P, nlow[a])))
        idx.append(ia)
        idx_low.append(il)
        prof.append(full[ia])
    return PyramidLevel(tuple(n), nlow, D, tuple(idx), tuple(idx_low), tuple(prof))


def fourier_pyramid(dims, J, mode):
    """src/pyramid/fourier_pyramid.jl:92-103."""
    ws, tbws, Ds = pyramid_mode(mode)
    levels = []
    n = tuple(dims)
    for j in range(J):
        k = 0 if j == 0 else 1
        lv = pyramid_level(n, ws[k], tbws[k], Ds[k])
        levels.append(lv)
        n = lv.nlow
    return levels


def _sep(prof):
    return prof[0][:, None, None] * prof[1][None, :, None] * prof[2][None, None, :]


def pyramid_split(X, lv):
    """src/pyramid/fourier_pyramid.jl:108-118 -- in place on X; returns low."""
    L = _sep(lv.prof)
    D15 = float(lv.D) ** 1.5
    ix = np.ix_(*lv.idx)
    low = np.zeros(lv.nlow, dtype=np.complex128)
    low[np.ix_(*lv.idx_low)] = (L / D15) * X[ix]
    X[ix] = X[ix] * np.sqrt(1 - L ** 2)
    return low


def pyramid_merge(X, low, lv):
    """src/pyramid/fourier_pyramid.jl:122-127 -- in place on X."""
    L = _sep(lv.prof)
    D15 = float(lv.D) ** 1.5
    ix = np.ix_(*lv.idx)
    X[ix] = np.sqrt(1 - L ** 2) * X[ix] + (L * D15) * low[np.ix_(*lv.idx_low)]
    return X


# ------------------------------------------------------------------- IRC ----

@dataclass
class IRCStage:
    """src/directional/irc.jl:29-34 (m is 0-based)."""
    m: int
    b: int
    ana: tuple
    syn: tuple


def irc_slice(f, t, chan, hdims, nK, K, m, le, scale):
    """src/directional/irc.jl:49-67 -- filter slice on a half-spectrum of size
    hdims (axis K halved), oriented to broadcast (K and m 0-based)."""
    nKh, nm = hdims[K], hdims[m]
    W = nm >> (le - 1)
    k = np.arange(nKh)[:, None]
    c = np.arange(nm)[None, :]
    g, l = np.divmod(c, W)
    sf = (1 << (le - 1)) - 1 - 2 * g
    wK = 2 * np.pi * k / nK
    wm = 2 * np.pi * l / W + sf * wK
    S = scale * f(t, chan, wK + 0 * wm, wm)
    shape = [1, 1, 1]
    shape[K] = nKh
    shape[m] = nm
    return (S if K < m else S.T).reshape(shape)


def irc_stages(hdims, nK, K, levels, bo):
    """src/directional/irc.jl:76-91 -- stages in analysis order."""
    t = F.CBD_LADDER[bo]
    stages = []
    for m in range(3):
        if m == K:
            continue
        if levels[m] < 0:
            raise ValueError("directional levels must be ≥ 0")
        if hdims[m] % (1 << levels[m]):
            raise ValueError(f"axis {m + 1} of length {hdims[m]} cannot be split {levels[m]} times")
        for le in range(1, levels[m] + 1):
            ana = tuple(irc_slice(F.cbd_analysis, t, c, hdims, nK, K, m, le, 0.5) for c in (0, 1))
            syn = tuple(irc_slice(F.cbd_synthesis, t, c, hdims, nK, K, m, le, 1.0) for c in (0, 1))
            stages.append(IRCStage(m, hdims[m] >> le, ana, syn))
    return stages


def _halves_index(n, b):
    """src/directional/irc.jl:39-45,96-100 -- axis positions of halves A and B
    when an n-bin axis is cut into (b, 2, n/2b) column-major."""
    c = np.arange(n)
    grp = c // (2 * b)
    r = c % (2 * b)
    A = c[r < b]
    B = c[r >= b]
    return A, B


def _take(arr, m, ix):
    if arr.shape[m] == 1:
        return arr
    return np.take(arr, ix, axis=m)


def irc_stage(S, m, b, W11, W12, W21, W22):
    """src/directional/irc.jl:105-111 -- A <- W11 A + W12 B, B <- W21 A + W22 B."""
    iA, iB = _halves_index(S.shape[m], b)
    A = np.take(S, iA, axis=m)
    Bv = np.take(S, iB, axis=m)
    nA = W11 * A + W12 * Bv
    nB = W21 * A + W22 * Bv
    sl = [slice(None)] * 3
    sl[m] = iA
    S[tuple(sl)] = nA
    sl[m] = iB
    S[tuple(sl)] = nB
    return S


def irc_analysis(S, stages):
    """src/directional/irc.jl:127-134."""
    for st in stages:
        iA, iB = _halves_index(S.shape[st.m], st.b)
        F0a, F0b = _take(st.ana[0], st.m, iA), _take(st.ana[0], st.m, iB)
        F1a, F1b = _take(st.ana[1], st.m, iA), _take(st.ana[1], st.m, iB)
        irc_stage(S, st.m, st.b, F0a, F0b, F1a, F1b)
    return S


def irc_synthesis(S, stages):
    """src/directional/irc.jl:139-146 -- transposed filter matrix, stages reversed."""
    for st in reversed(stages):
        iA, iB = _halves_index(S.shape[st.m], st.b)
        G0a, G0b = _take(st.syn[0], st.m, iA), _take(st.syn[0], st.m, iB)
        G1a, G1b = _take(st.syn[1], st.m, iA), _take(st.syn[1], st.m, iB)
        irc_stage(S, st.m, st.b, G0a, G1a, G0b, G1b)
    return S


# --------------------------------------------------------------- levels -----

def wedge_split(L):
    return (L + 1) // 2, L // 2


def surf_levels(L):
    """src/types/transforms.jl:270-284 -- integer L or a 3x3 matrix -> levels[K][m]
    (0-based K, m)."""
    if isinstance(L, (int, np.integer)):
        la, lb = wedge_split(int(L))
        out = []
        for K in range(3):
            a, b = [x - 1 for x in F.transverse_axes(K + 1)]
            out.append(tuple(la if m == a else (lb if m == b else 0) for m in range(3)))
        return tuple(out)
    M = np.asarray(L)
    if M.shape != (3, 3):
        raise ValueError(f"a level matrix must be 3 × 3, got size {M.shape}")
    return tuple(tuple(0 if m == K else int(M[K, m]) for m in range(3)) for K in range(3))


def n_scale_subbands(lv):
    """src/types/transforms.jl:315."""
    return sum(1 << sum(lv[K]) for K in range(3))


@dataclass(frozen=True)
class SParams:
    """src/types/transforms.jl:222-261 -- Surfacelet's parameters."""
    scales: int
    direction_levels: tuple
    pyramid_ratio: float = 1.5
    checkerboard_order: int = 12
    diamond_length: int = 15
    kaiser_beta: float = 2.5
    diamond_power: int = 4

    @property
    def levels(self):
        return [surf_levels(L) for L in self.direction_levels]


# ---------------------------------------------------------------- scale -----

def half_dims(n, K):
    """src/directional/surfacelet_scale.jl:28."""
    return tuple(n[K] // 2 + 1 if d == K else n[d] for d in range(3))


@dataclass
class SurfScale:
    """src/directional/surfacelet_scale.jl:12-17."""
    dims: tuple
    levels: tuple
    masks: list
    stages: list


def surf_stages(n, lv, bo):
    """src/directional/surfacelet_scale.jl:21-25."""
    return [irc_stages(half_dims(n, K), n[K], K, lv[K], bo) for K in range(3)]


def surf_blocks(sc, K):
    """src/directional/surfacelet_scale.jl:32-37."""
    n = sc.dims
    counts = tuple(1 << sc.levels[K][d] for d in range(3))
    bsize = tuple(n[K] // 2 + 1 if d == K else n[d] // counts[d] for d in range(3))
    return counts, bsize


def surf_subband_size(sc, K):
    """src/directional/surfacelet_scale.jl:40-43."""
    _, bsize = surf_blocks(sc, K)
    return tuple(sc.dims[K] if d == K else bsize[d] for d in range(3))


def block_ranges(counts, bsize, i):
    """src/directional/surfacelet_scale.jl:57-60 -- block i (0-based), column
    major over block indices; returns slices."""
    v = np.unravel_index(i, counts, order="F")
    return tuple(slice(v[d] * bsize[d], (v[d] + 1) * bsize[d]) for d in range(3))


def brfft_K(blk, nK, K):
    """Unnormalised c2r over (K, others) with axis K halved."""
    others = [d for d in range(3) if d != K]
    shape = [blk.shape[d] for d in range(3)]
    shape[K] = nK
    s = [shape[d] for d in others] + [nK]
    r = np.fft.irfftn(blk, s=s, axes=others + [K])
    return r * np.prod(shape)


def rfft_K(sb, K):
    others = [d for d in range(3) if d != K]
    return np.fft.rfftn(sb, axes=others + [K])


def scale_forward(X, sc):
    """src/directional/surfacelet_scale.jl:200-213 -- real subbands of one scale."""
    out = []
    Ss = []
    for K in range(3):
        s = 1.0 / np.prod(surf_subband_size(sc, K))
        hd = half_dims(sc.dims, K)
        S = X[:hd[0], :hd[1], :hd[2]] * sc.masks[K] * s
        irc_analysis(S, sc.stages[K])
        Ss.append(S)
    for K in range(3):
        counts, bsize = surf_blocks(sc, K)
        for i in range(int(np.prod(counts))):
            r = block_ranges(counts, bsize, i)
            out.append(brfft_K(Ss[K][r], sc.dims[K], K))
    return out


def hermitian_add(X, S, M, K, c):
    """src/directional/surfacelet_scale.jl:246-257 -- X += c M S on the half of
    axis K and X[-w] += c M conj(S) on the strictly interior mirror."""
    n = X.shape
    h = S.shape[K]
    X[:S.shape[0], :S.shape[1], :S.shape[2]] += c * M * S
    # Mirror: dst index along K in h..n-1 from src h-2 down to 1; along the
    # other axes index 0 -> 0 and 1..n-1 -> n-1..1, i.e. (-i) mod n.
    src_k = np.arange(h - 2, 0, -1)
    dst_k = np.arange(h, n[K])
    Mc = c * M * np.conj(S)
    idx_dst = [None] * 3
    idx_src = [None] * 3
    for d in range(3):
        if d == K:
            idx_dst[d] = dst_k
            idx_src[d] = src_k
        else:
            idx_dst[d] = np.arange(n[d])
            idx_src[d] = (-np.arange(n[d])) % n[d]
    X[np.ix_(*idx_dst)] += Mc[np.ix_(*idx_src)]
    return X


def scale_inverse(subbands, sc):
    """src/directional/surfacelet_scale.jl:220-239 -- the scale's full spectrum."""
    X = np.zeros(sc.dims, dtype=np.complex128)
    n = 0
    Ss = []
    for K in range(3):
        counts, bsize = surf_blocks(sc, K)
        S = np.zeros(half_dims(sc.dims, K), dtype=np.complex128)
        want = surf_subband_size(sc, K)
        for i in range(int(np.prod(counts))):
            sb = subbands[n]
            if tuple(sb.shape) != want:
                raise ValueError(f"subband {n + 1} has size {tuple(sb.shape)}, expected {want}")
            S[block_ranges(counts, bsize, i)] = rfft_K(sb, K)
            n += 1
        Ss.append(S)
    for K in range(3):
        irc_synthesis(Ss[K], sc.stages[K])
        hermitian_add(X, Ss[K], sc.masks[K], K, 1.0 / 3)
    return X


# ------------------------------------------------------- sizes / layout -----

def axis_problem(t, a, n):
    """src/transforms/surfacelet.jl:230-249 (a 0-based)."""
    J = t.scales
    ws, tbws, Ds = pyramid_mode(t.pyramid_ratio)
    levels = t.levels
    for j in range(1, J + 1):
        s = J - j + 1
        where = f"at scale {s} of {J} (1 the coarsest) the length is {n}"
        if n % 2:
            return f"{where}, which must be even"
        c = 1 << max(levels[s - 1][K][a] for K in range(3))
        if n % c:
            return f"{where}, which the directional levels must split into {c} wedge blocks"
        k = 0 if j == 1 else 1
        D = Ds[k]
        if (n * D.denominator) % D.numerator:
            return f"{where}, which resampling by {D.numerator}/{D.denominator} needs to be a multiple of {D.numerator}"
        nlow = n * D.denominator // D.numerator
        P = F.rcos_edges(n, ws[k], tbws[k])[0]
        if 2 * P > nlow:
            return f"{where}, too small for the pyramid's lowpass"
        n = nlow
    return None


def check_size(t, dims):
    """src/transforms/surfacelet.jl:253-269."""
    for a in range(3):
        problem = axis_problem(t, a, dims[a])
        if problem is None:
            continue
        n = dims[a]
        near = 0
        for d in range(1, 4 * n + 65):
            for m in (n - d, n + d):
                if m >= 1 and axis_problem(t, a, m) is None:
                    near = m
                    break
            if near > 0:
                break
        hint = f"; the nearest length that fits is {near}" if near > 0 else ""
        raise ValueError(
            f"a volume of size ({', '.join(map(str, dims))}) does not fit Surfacelet(scales = {t.scales}, "
            f"pyramid_ratio = {float(t.pyramid_ratio)}): axis {a + 1}: {problem}{hint}")
    return tuple(dims)


def layout_sizes(t, dims):
    """src/transforms/surfacelet.jl:275-290."""
    check_size(t, dims)
    pyr = fourier_pyramid(dims, t.scales, t.pyramid_ratio)
    sizes = [pyr[-1].nlow]
    levels = t.levels
    for j in range(1, t.scales + 1):
        n = pyr[t.scales - j].n
        for K in range(3):
            counts = tuple(1 << levels[j - 1][K][d] for d in range(3))
            sb = tuple(n[K] if d == K else n[d] // counts[d] for d in range(3))
            sizes += [sb] * int(np.prod(counts))
    return sizes


def redundancy(t, dims):
    """src/transforms/surfacelet.jl:220-222."""
    return sum(int(np.prod(s)) for s in layout_sizes(t, dims)) / int(np.prod(dims))


def subband_layout(t, dims):
    """src/transforms/surfacelet.jl:301-308."""
    layout_sizes(t, dims)
    lay = [(0, 0, 0)]
    for j in range(1, t.scales + 1):
        for K in range(1, 4):
            for w in range(1, (1 << sum(t.levels[j - 1][K - 1])) + 1):
                lay.append((j, K, w))
    return lay


@dataclass
class SWork:
    dims: tuple
    pyramid: list
    scales: list
    offsets: list


def workspace(t, dims):
    """src/transforms/surfacelet.jl:136-149,73-102 (tables only)."""
    check_size(t, dims)
    pyr = fourier_pyramid(dims, t.scales, t.pyramid_ratio)
    scales = []
    levels = t.levels
    for j in range(t.scales):
        n = pyr[j].n
        lv = levels[t.scales - j - 1]
        masks = F.ritf_masks(n, t.diamond_length, t.kaiser_beta, t.diamond_power)
        scales.append(SurfScale(n, lv, masks, surf_stages(n, lv, t.checkerboard_order)))
    counts = [n_scale_subbands(sc.levels) for sc in scales]
    offsets = [1 + sum(counts[j + 1:]) for j in range(len(counts))]
    return SWork(tuple(dims), pyr, scales, offsets)


def analyze(t, x, ws=None):
    """src/transforms/surfacelet.jl:339-381 -- list of subbands (layout order)."""
    x = np.asarray(x)
    if np.iscomplexobj(x):
        a = analyze(t, x.real, ws)
        b = analyze(t, x.imag, ws)
        return [p + 1j * q for p, q in zip(a, b)]
    ws = ws or workspace(t, x.shape)
    S = 1 + sum(n_scale_subbands(sc.levels) for sc in ws.scales)
    out = [None] * S
    X = np.fft.fftn(x.astype(np.complex128))
    for j, sc in enumerate(ws.scales):
        low = pyramid_split(X, ws.pyramid[j])
        bands = scale_forward(X, sc)
        for n, b in enumerate(bands):
            out[ws.offsets[j] - 1 + n + 1] = b
        X = low
    out[0] = np.real(np.fft.ifftn(X))
    return out


def synthesize(t, c, ws=None, dims=None):
    """src/transforms/surfacelet.jl:354-396."""
    if any(np.iscomplexobj(b) for b in c):
        a = synthesize(t, [np.real(b) for b in c], ws, dims)
        b = synthesize(t, [np.imag(b) for b in c], ws, dims)
        return a + 1j * b
    if ws is None:
        ws = workspace(t, dims)
    C = np.fft.fftn(np.asarray(c[0], dtype=np.complex128))
    for j in reversed(range(len(ws.scales))):
        sc = ws.scales[j]
        nb = n_scale_subbands(sc.levels)
        bands = c[ws.offsets[j]:ws.offsets[j] + nb]
        X = scale_inverse(bands, sc)
        C = pyramid_merge(X, C, ws.pyramid[j])
    return np.real(np.fft.ifftn(C))


# ------------------------------------------------------------ adjoint -------
#
# The decimated transform is a frame but not a tight one, so its adjoint is
# not `synthesize` (src/transforms/surfacelet.jl:24-26).  The adjoint below is
# derived stage by stage from the analysis (real inner products; a complex
# spectrum counts as R^2): fft -> bfft real part, pyramid crop/scale ->
# scatter with L / D^1.5, IRC 2x2 mixing -> conjugate transpose in reverse
# stage order, real masks -> the same masks, and the unnormalised c2r of a
# wedge block -> rfft with the interior bins along K doubled.  It is pinned
# by dense probes <A x, y> = <x, A^T y> (tests/test_oracle_decimated.py).

def pyramid_split_adjoint(Xh, low, lv):
    """Adjoint of pyramid_split: returns X-bar from (highpass-bar, lowpass-bar)."""
    L = _sep(lv.prof)
    D15 = float(lv.D) ** 1.5
    ix = np.ix_(*lv.idx)
    X = Xh.copy()
    X[ix] = np.sqrt(1 - L ** 2) * Xh[ix] + (L / D15) * low[np.ix_(*lv.idx_low)]
    return X


def irc_analysis_adjoint(S, stages):
    """Adjoint of irc_analysis: conjugate-transposed stage matrices, reversed."""
    for st in reversed(stages):
        iA, iB = _halves_index(S.shape[st.m], st.b)
        F0a, F0b = _take(st.ana[0], st.m, iA), _take(st.ana[0], st.m, iB)
        F1a, F1b = _take(st.ana[1], st.m, iA), _take(st.ana[1], st.m, iB)
        irc_stage(S, st.m, st.b, np.conj(F0a), np.conj(F1a), np.conj(F0b), np.conj(F1b))
    return S


def brfft_K_adjoint(sb, K):
    """Adjoint of brfft_K (unnormalised c2r along K after inverse c2c along the
    other axes): rfft over (others, K), interior bins along K doubled."""
    Z = rfft_K(sb, K)
    h = Z.shape[K]
    n = sb.shape[K]
    sl = [slice(None)] * 3
    sl[K] = slice(1, h - 1 if n % 2 == 0 else h)
    Z[tuple(sl)] *= 2
    return Z


def scale_adjoint(subbands, sc):
    """Adjoint of scale_forward: the highpass spectrum-bar of one scale."""
    X = np.zeros(sc.dims, dtype=np.complex128)
    n = 0
    for K in range(3):
        counts, bsize = surf_blocks(sc, K)
        S = np.zeros(half_dims(sc.dims, K), dtype=np.complex128)
        for i in range(int(np.prod(counts))):
            S[block_ranges(counts, bsize, i)] = brfft_K_adjoint(subbands[n], K)
            n += 1
        irc_analysis_adjoint(S, sc.stages[K])
        s = 1.0 / np.prod(surf_subband_size(sc, K))
        hd = half_dims(sc.dims, K)
        X[:hd[0], :hd[1], :hd[2]] += S * sc.masks[K] * s
    return X


def adjoint(t, c, ws=None, dims=None):
    """A^T c for the analysis operator A of the decimated Surfacelet."""
    if any(np.iscomplexobj(b) for b in c):
        a = adjoint(t, [np.real(b) for b in c], ws, dims)
        b = adjoint(t, [np.imag(b) for b in c], ws, dims)
        return a + 1j * b
    if ws is None:
        ws = workspace(t, dims)
    C = np.fft.fftn(np.asarray(c[0], dtype=np.complex128)) / np.prod(c[0].shape)
    for j in reversed(range(len(ws.scales))):
        sc = ws.scales[j]
        nb = n_scale_subbands(sc.levels)
        Xh = scale_adjoint(c[ws.offsets[j]:ws.offsets[j] + nb], sc)
        C = pyramid_split_adjoint(Xh, C, ws.pyramid[j])
    return np.real(np.fft.ifftn(C)) * np.prod(C.shape)
