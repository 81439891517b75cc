"""FP64 restatement of NSSurfacelet analyze/synthesize (test infrastructure only).

Follows src/transforms/ns_surfacelet.jl (cited per function).  Volumes are
numpy arrays indexed [i1, i2, i3]; coefficients are a list of S volumes in
subband_layout order (the lowpass first).  numpy.fft's rfftn/irfftn over axes
(1, 2, 0)... are arranged so that axis 0 (i1) is the halved axis, matching the
reference's rfft along axis 1.
"""
from dataclasses import dataclass, field

import numpy as np

from . import filters as F


@dataclass(frozen=True)
class NSParams:
    """src/types/transforms.jl:86-113 -- NSSurfacelet's parameters."""
    scales: int
    direction_levels: tuple
    bernstein_order: int = 12
    halfband_order: int = 1
    transition: float = 1.0
    tight: bool = True


def nbands(t, j):
    """src/types/transforms.jl:157 -- directional subbands of scale j (1-based)."""
    return 3 * (1 << t.direction_levels[j - 1])


def nsubbands(t):
    """src/transforms/ns_surfacelet.jl:527."""
    return 1 + sum(nbands(t, j) for j in range(1, t.scales + 1))


def subband_layout(t):
    """src/transforms/ns_surfacelet.jl:518-525."""
    lay = [(0, 0)]
    for j in range(1, t.scales + 1):
        lay += [(j, k) for k in range(1, nbands(t, j) + 1)]
    return lay


def rfft3(x):
    """rfft along axis 1 (numpy axis 0), full along axes 2, 3."""
    return np.fft.rfftn(x, axes=(2, 1, 0))


def irfft3(X, dims):
    """irfft normalised by 1/N, axis 1 halved."""
    return np.fft.irfftn(X, s=(dims[2], dims[1], dims[0]), axes=(2, 1, 0))


@dataclass
class NSTables:
    dims: tuple
    m1: np.ndarray
    m2: np.ndarray
    pa: list          # [level 0..J] of per-axis (a1, a2, a3), oriented
    pl: list          # [level 0..J-1] of per-axis (l1, l2, l3), oriented
    wedges: list      # [scale j-1][hourglass i-1] = (wa planes, wb planes)
    groups: list = field(default_factory=list)   # [(j, i, subband indices)]


def tables(t, dims):
    """src/transforms/ns_surfacelet.jl:98-123 (tables of the host workspace)."""
    if any(d < 1 for d in dims):
        raise ValueError(f"volume size must be positive, got {tuple(dims)}")
    n1, n2, n3 = dims
    m1, m2 = F.hourglass_tables(dims, t.bernstein_order, t.halfband_order)
    nb = (n1 // 2 + 1, n2, n3)
    resp = [F.rc_responses(t.scales, dims[d], nb[d]) for d in range(3)]
    J = t.scales
    for j in range(J):
        empty = all(
            all(a == 0 or l == 1 for a, l in zip(resp[d][0][j], resp[d][1][j]))
            for d in range(3)
        )
        if empty:
            raise ValueError(
                f"size ({n1}, {n2}, {n3}) is too small for {J} scales: scale {J - j} would be empty")

    def orient(v, d):
        shape = [1, 1, 1]
        shape[d] = len(v)
        return v.reshape(shape)

    pa = [tuple(orient(resp[d][0][j], d) for d in range(3)) for j in range(J + 1)]
    pl = [tuple(orient(resp[d][1][j], d) for d in range(3)) for j in range(J)]
    wedges = []
    for L in t.direction_levels:
        lev = F.wedge_levels(L)
        per = []
        for i in (1, 2, 3):
            wa, wb = F.wedge_masks(dims, i, lev, t.halfband_order, t.transition)
            if t.tight:
                wa = [F.tight_sqrt(w) for w in wa]
                wb = [F.tight_sqrt(w) for w in wb]
            per.append((wa, wb))
        wedges.append(per)
    tb = NSTables(tuple(dims), m1, m2, pa, pl, wedges)
    # src/transforms/ns_surfacelet.jl:222-242: items and subband indices.
    offsets = np.cumsum([1] + [nbands(t, j) for j in range(1, J)])
    for j in range(1, J + 1):
        for i in (1, 2, 3):
            wa, wb = wedges[j - 1][i - 1]
            na, nbb = len(wa), len(wb)
            idx = []
            for a in range(1, na + 1):
                for b in range(1, nbb + 1):
                    idx.append((int(offsets[j - 1]) + ns_index(i, a, b, na, nbb), a, b))
            tb.groups.append((j, i, idx))
    return tb


def ns_index(i, p, q, na, nb):
    """src/transforms/ns_surfacelet.jl:430 -- 1-based within a scale."""
    return ((i - 1) * na + (p - 1)) * nb + q


def bandpass(tb, level):
    """src/transforms/ns_surfacelet.jl:271-272 -- B_level (0-based pyramid level)."""
    a1, a2, a3 = tb.pa[level]
    l1, l2, l3 = tb.pl[level]
    return a1 * a2 * a3 * np.sqrt(np.maximum(1 - (l1 * l2 * l3) ** 2, 0.0))


def hg_factor(tb, i, tight):
    """src/transforms/ns_surfacelet.jl:265-268."""
    w = F.stage1_weight(tb.m1, i) * F.stage2_weight(tb.m2, i)[None, :, :]
    return F.tight_sqrt(w) if tight else w


def group_mask(t, tb, j, i, analysis=True):
    """src/transforms/ns_surfacelet.jl:226-227,295-303 -- the group mask of
    hourglass i at scale j: B * hg (analysis; synthesis when tight) or B."""
    B = bandpass(tb, t.scales - j)
    if analysis or t.tight:
        return B * hg_factor(tb, i, t.tight)
    return B


def lowpass_mask(t, tb):
    """src/transforms/ns_surfacelet.jl:350-358."""
    a1, a2, a3 = tb.pa[t.scales]
    return a1 * a2 * a3


def subband_masks(t, dims, tb=None):
    """The full analysis mask of every subband, in subband_layout order, on the
    rfft grid (for tests and dense checks; the transform never stores them)."""
    tb = tb or tables(t, dims)
    h = (dims[0] // 2 + 1, dims[1], dims[2])
    S = nsubbands(t)
    masks = [None] * S
    masks[0] = np.broadcast_to(lowpass_mask(t, tb), h).copy()
    for (j, i, idx) in tb.groups:
        G = group_mask(t, tb, j, i)
        wa, wb = tb.wedges[j - 1][i - 1]
        for (k, a, b) in idx:
            masks[k - 1] = np.broadcast_to(G * wa[a - 1] * wb[b - 1], h).copy()
    return masks


def analyze(t, x, tb=None):
    """src/transforms/ns_surfacelet.jl:551-566 + the engine schedule at :23-32.
    Real or complex x; complex runs as two real channels."""
    x = np.asarray(x)
    if np.iscomplexobj(x):
        cr = analyze(t, x.real, tb)
        ci = analyze(t, x.imag, tb)
        return [a + 1j * b for a, b in zip(cr, ci)]
    dims = x.shape
    tb = tb or tables(t, dims)
    X = rfft3(x.astype(np.float64))
    out = [None] * nsubbands(t)
    for (j, i, idx) in tb.groups:
        Y = X * group_mask(t, tb, j, i)
        wa, wb = tb.wedges[j - 1][i - 1]
        for (k, a, b) in idx:
            out[k - 1] = irfft3(Y * wa[a - 1] * wb[b - 1], dims)
    out[0] = irfft3(X * lowpass_mask(t, tb), dims)
    return out


def synthesize(t, c, tb=None):
    """src/transforms/ns_surfacelet.jl:568-581 -- per-subband rfft, wedge
    accumulate per group, group accumulate, lowpass, one irfft."""
    if any(np.iscomplexobj(b) for b in c):
        return synthesize(t, [np.real(b) for b in c], tb) + 1j * synthesize(t, [np.imag(b) for b in c], tb)
    dims = c[0].shape
    tb = tb or tables(t, dims)
    acc = np.zeros((dims[0] // 2 + 1, dims[1], dims[2]), dtype=np.complex128)
    for (j, i, idx) in tb.groups:
        wa, wb = tb.wedges[j - 1][i - 1]
        accg = np.zeros_like(acc)
        for (k, a, b) in idx:
            Fk = rfft3(np.asarray(c[k - 1], dtype=np.float64))
            accg += Fk * (wa[a - 1] * wb[b - 1] if t.tight else 1.0)
        acc += accg * group_mask(t, tb, j, i, analysis=False)
    acc += rfft3(np.asarray(c[0], dtype=np.float64)) * lowpass_mask(t, tb)
    return irfft3(acc, dims)


def stream_order(t):
    """src/transforms/ns_surfacelet.jl:472-477,583-592 -- directional subbands
    in increasing index, then the lowpass (1-based indices)."""
    return list(range(2, nsubbands(t) + 1)) + [1]
