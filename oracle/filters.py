"""FP64 restatement of the reference's filter design (test infrastructure only).

Every function cites the reference file:line it restates.  Arrays are indexed
[i1, i2, i3] with axis 1 the contiguous (rfft) axis, as in the reference.
"""
from fractions import Fraction
from math import comb, pi

import numpy as np


# ---------------------------------------------------------------- Bernstein --

def bernstein_table(N, omega):
    """src/filters/bernstein.jl:36-53 -- T[j, k] = C(N,j) x^j (1-x)^(N-j),
    x = sin(w/2)^2 clamped to [0, 1]."""
    if N < 1:
        raise ValueError("Bernstein order must be ≥ 1")
    omega = np.asarray(omega, dtype=np.float64)
    x = np.clip(np.sin(omega / 2) ** 2, 0.0, 1.0)
    y = 1.0 - x
    j = np.arange(N + 1)[:, None]
    binom = np.array([float(comb(N, int(k))) for k in range(N + 1)])[:, None]
    return binom * x[None, :] ** j * y[None, :] ** (N - j)


def dft_frequencies(n, half=False):
    """src/filters/bernstein.jl:64-68 -- w_k = 2 pi (k-1) / n."""
    m = n // 2 + 1 if half else n
    return np.array([2 * pi * k / n for k in range(m)])


def folded_frequencies(n, half=False):
    """src/filters/bernstein.jl:72-79 -- folded to (-pi, pi]; the half grid
    is returned unfolded in [0, pi]."""
    if half:
        return dft_frequencies(n, half=True)
    w = np.array([2 * pi * k / n for k in range(n)])
    return np.where(w > pi, w - 2 * pi, w)


def mapping_coefficients(N, d):
    """src/filters/bernstein.jl:97-108 for group (1,) -- sign(j1 - max others)."""
    idx = np.indices((N + 1,) * d)
    a = idx[0]
    b = idx[1:].max(axis=0)
    return np.sign(a - b).astype(np.int64)


def mapping_response(coeffs, tables):
    """src/filters/bernstein.jl:123-151 -- M(w) = sum_j c_j prod_d T_d[j_d, k_d]."""
    c = coeffs.astype(np.float64)
    if c.ndim == 2:
        return np.einsum("ab,ai,bj->ij", c, tables[0], tables[1], optimize=True)
    return np.einsum("abc,ai,bj,ck->ijk", c, tables[0], tables[1], tables[2], optimize=True)


def halfband_split(x, m):
    """src/filters/bernstein.jl:171-190 -- f0(x) = ((1+x)/2)^(m+1) *
    sum_k C(m+k,k) ((1-x)/2)^k, the sum formed first and scaled last."""
    if m < 0:
        raise ValueError("polynomial order must be ≥ 0")
    x = np.asarray(x, dtype=np.float64)
    u = (1.0 + x) / 2
    v = (1.0 - x) / 2
    s = np.zeros_like(x)
    c = 1.0
    for k in range(m + 1):
        s = s + c * v ** k
        c = c * (m + k + 1) / (k + 1)
    out = u ** (m + 1) * s
    return np.where(u == 0, 0.0, out)


def step(u, poly):
    """src/filters/bernstein.jl:196 -- halfband_split of the clamped argument."""
    return halfband_split(np.clip(u, -1.0, 1.0), poly)


def soft_step(s, b, w, poly):
    """src/filters/bernstein.jl:200-203 -- hard step when w <= 0."""
    if w <= 0:
        return np.where(s >= b, 1.0, 0.0)
    return step((s - b) / w, poly)


# ------------------------------------------------------- Hourglass tables --

def hourglass_tables(dims, order, poly):
    """src/directional/hourglass.jl:172-189 -- m1 = f0(M1) on the rfft grid
    (h1, n2, n3) and m2 = f0(M2) on the (w2, w3) grid, shape (n2, n3)."""
    n1, n2, n3 = dims
    b1 = bernstein_table(order, dft_frequencies(n1, half=True))
    b2 = bernstein_table(order, dft_frequencies(n2))
    b3 = bernstein_table(order, dft_frequencies(n3))
    m1 = halfband_split(mapping_response(mapping_coefficients(order, 3), (b1, b2, b3)), poly)
    m2 = halfband_split(mapping_response(mapping_coefficients(order, 2), (b2, b3)), poly)
    return m1, m2


def stage1_weight(f0, i):
    """src/directional/hourglass.jl:296."""
    return f0 if i == 1 else 1.0 - f0


def stage2_weight(f0, i):
    """src/directional/hourglass.jl:300-303."""
    if i == 1:
        return np.ones_like(f0)
    return f0 if i == 2 else 1.0 - f0


def tight_sqrt(x):
    """src/directional/hourglass.jl:289."""
    return np.sqrt(np.maximum(x, 0.0))


def hourglass_weight(m1, m2, i):
    """src/directional/hourglass.jl:279-285 -- W_i on the rfft grid."""
    return stage1_weight(m1, i) * stage2_weight(m2, i)[None, :, :]


# ------------------------------------------------------ Raised cosine -------

def rcos(x):
    """src/pyramid/fourier_pyramid.jl:42."""
    x = np.asarray(x, dtype=np.float64)
    return np.where(x <= 0, 0.0, np.where(x >= 1, 1.0, (1 - np.cos(pi * x)) / 2))


def rcos_edges(n, w, t):
    """src/pyramid/fourier_pyramid.jl:47 -- exact rational band edges (P, Q)."""
    a = Fraction(n, 2) * (w + t)
    b = Fraction(n, 2) * (w - t)
    P = -((-a.numerator) // a.denominator)   # ceil
    Q = b.numerator // b.denominator         # floor
    return P, Q


def rcos_axis(n, w, t):
    """src/pyramid/fourier_pyramid.jl:54-61 -- sqrt-rcos lowpass on a full
    n-point DFT axis; returns (profile, P)."""
    P, Q = rcos_edges(n, w, t)
    prof = np.zeros(n)
    for k in range(n):
        f = abs(k if k <= n // 2 else k - n)
        if f <= P:
            prof[k] = np.sqrt(rcos((P - f) / (P - Q)))
    return prof, P


def rc_responses(J, n, nbins):
    """src/pyramid/fourier_pyramid.jl:143-150 -- per-axis lowpass products
    acc[j] = prod_{i<j} l_i (j = 0..J) and profiles l_j (j = 0..J-1), on the
    first nbins DFT bins of an n-point axis."""
    prof = [rcos_axis(n, Fraction(1, 2 ** j), Fraction(1, 3 * 2 ** j))[0][:nbins] for j in range(1, J + 1)]
    acc = [np.ones(nbins)]
    for j in range(J):
        acc.append(acc[-1] * prof[j])
    return acc, prof


# ------------------------------------------------------------- NS wedges ----

def transverse_axes(i):
    """src/directional/ndfb_wedges.jl:14 (1-based axes)."""
    return (2, 3) if i == 1 else ((1, 3) if i == 2 else (1, 2))


def wedge_levels(L):
    """src/directional/ndfb_wedges.jl:19 -- (cld(L,2), fld(L,2))."""
    return (L + 1) // 2, L // 2


def wedge_weight(s, idx, nw, transition, poly):
    """src/directional/ndfb_wedges.jl:25-32 (idx 1-based)."""
    if nw == 1:
        return np.ones_like(s)
    delta = 2.0 / nw
    width = transition * delta
    lo = 1.0 if idx == 1 else soft_step(s, -1.0 + (idx - 1) * delta, width, poly)
    hi = 0.0 if idx == nw else soft_step(s, -1.0 + idx * delta, width, poly)
    return lo - hi


def self_conjugate_bin(k, n):
    """src/directional/ndfb_wedges.jl:78 with 0-based k: mod(2k, n) == 0."""
    return (2 * k) % n == 0


def wedge_ratio(wr, wt, ref_self, tr_self):
    """src/directional/ndfb_wedges.jl:97-106."""
    if wr == 0:
        return 0.0
    if ref_self:
        return abs(wt) / abs(wr)
    if tr_self:
        return wt / abs(wr)
    return wt / wr


def wedge_plane(dims, ref, tr, l, poly, transition):
    """src/directional/ndfb_wedges.jl:45-69 -- the 2^l weight planes of one
    transverse axis, each indexed [p (ref axis bin), q (tr axis bin)] on the
    rfft grid (axis 1 halved).  Orientation (ndfb_wedges.jl:112-113) only
    places value [p, q] at (axis ref = p, axis tr = q)."""
    wref = folded_frequencies(dims[ref - 1], half=ref == 1)
    wtr = folded_frequencies(dims[tr - 1], half=tr == 1)
    nw = 1 << l
    if l == 0:
        return [np.ones((len(wref), len(wtr)))]
    nref, ntr = dims[ref - 1], dims[tr - 1]
    s = np.zeros((len(wref), len(wtr)))
    for q, wt in enumerate(wtr):
        for p, wr in enumerate(wref):
            s[p, q] = wedge_ratio(wr, wt, self_conjugate_bin(p, nref), self_conjugate_bin(q, ntr))
    return [wedge_weight(s, k, nw, transition, poly) for k in range(1, nw + 1)]


def orient_plane(plane, ref, tr, dims):
    """Broadcastable 3-D view of a (ref, tr) plane on the rfft grid."""
    h = (dims[0] // 2 + 1, dims[1], dims[2])
    shape = [1, 1, 1]
    shape[ref - 1] = h[ref - 1]
    shape[tr - 1] = h[tr - 1]
    m = plane if ref < tr else plane.T
    return m.reshape(shape)


def wedge_masks(dims, i, levels, poly, transition):
    """src/directional/ndfb_wedges.jl:35-43 -- ((wa planes), (wb planes)),
    each plane already oriented to broadcast on the rfft grid."""
    a, b = transverse_axes(i)
    pa = [orient_plane(p, i, a, dims) for p in wedge_plane(dims, i, a, levels[0], poly, transition)]
    pb = [orient_plane(p, i, b, dims) for p in wedge_plane(dims, i, b, levels[1], poly, transition)]
    return pa, pb


# ------------------------------------------------------ Checkerboard (IRC) --

CBD_LADDER = {
    4: [-0.6262169131293146, -0.2391935764853579],
    6: [-0.6263429050738657, -0.18302382681237633, -0.10753046674888095],
    8: [-0.627819441355453, -0.1868110527281001, -0.08821345831338342, -0.05299463507171143],
    10: [-0.6290391996524649, -0.19023436487159193, -0.09316419099049718,
         -0.04787638089257991, -0.02729049477362672],
    12: [-0.6299909650161403, -0.19294788893803155, -0.09719811683520764,
         -0.05257765509290852, -0.027224347419537415, -0.014421093056370137],
    14: [-0.6307470366053054, -0.19509552613399547, -0.10042251574577052,
         -0.056418179282558024, -0.03113763249657735, -0.015819995468636647,
         -0.007750295209451851],
    16: [-0.631356163548133, -0.19683521105934682, -0.10304471724680653,
         -0.05956342644490844, -0.03441276144847804, -0.018849533868796398,
         -0.009297418074817033, -0.004215553479603451],
    18: [-0.6318535467876545, -0.19826135647461407, -0.10520738123671233,
         -0.06218709792109189, -0.037184557113685186, -0.021472356755448885,
         -0.011535270028700011, -0.00549634585408209, -0.0023119823375261405],
}
"""The ladder taps t_q (q = 1, 3, ..., bo-1) of SurfBox's checkerboard filters,
a table that must be identical: src/filters/checkerboard.jl:26-53."""


def cbd_b(t, w):
    """src/filters/checkerboard.jl:71-77 -- b(w) = 2 sum t_q sin(q w), q odd."""
    w = np.asarray(w, dtype=np.float64)
    s = np.zeros_like(w)
    for i, tq in enumerate(t):
        s = s + tq * np.sin((2 * i + 1) * w)
    return 2 * s


def cbd_analysis(t, chan, w1, w2):
    """src/filters/checkerboard.jl:80-84."""
    B = -cbd_b(t, w1) * cbd_b(t, w2)
    r = 1 / np.sqrt(2.0)
    if chan == 0:
        return (r * (1 + B)).astype(np.complex128)
    return np.exp(1j * np.asarray(w2)) * (r * (2 - B - B ** 2))


def cbd_synthesis(t, chan, w1, w2):
    """src/filters/checkerboard.jl:87-91."""
    B = -cbd_b(t, w1) * cbd_b(t, w2)
    r = 1 / np.sqrt(2.0)
    if chan == 0:
        return (r * (2 + B - B ** 2)).astype(np.complex128)
    return np.exp(-1j * np.asarray(w2)) * (r * (1 - B))


# -------------------------------------------------------- ritf hourglass ----

def ritf_I0(x):
    """src/directional/ritf.jl:24-37 -- SurfBox's truncated I0 series."""
    x = x / 2
    y = 1.0
    z = 1.0
    t = 1.0
    k = 1
    while t > 1.0e-10 * y:
        z = z * x / k
        k += 1
        t = z * z
        y += t
    return y


def ritf_window(n, beta):
    """src/directional/ritf.jl:40-52 -- Kaiser window x half-band sinc."""
    if n % 2 == 0:
        raise ValueError("msize must be odd")
    den = ritf_I0(beta)
    half = (n - 1) // 2
    val = [2.0 * i / (n - 1) - 1 for i in range(half + 1)]
    wh = [ritf_I0(beta * np.sqrt(1 - v * v)) / den for v in val]
    w = np.array(wh + wh[:-1][::-1])
    for i, u in enumerate(range(-half, half + 1)):
        w[i] *= np.sinc(u / 2)
    return w


def ritf_diamond(w, a, b):
    """src/directional/ritf.jl:58-73 -- zero-phase diamond response."""
    half = (len(w) - 1) // 2
    al = (a + b) / 2
    be = (a - b) / 2
    ce_a = ce_b = co_a = co_b = 0.0
    for i, u in enumerate(range(-half, half + 1)):
        if u % 2 == 0:
            ce_a = ce_a + w[i] * np.cos(u * al)
            ce_b = ce_b + w[i] * np.cos(u * be)
        else:
            co_a = co_a + w[i] * np.cos(u * al)
            co_b = co_b + w[i] * np.cos(u * be)
    return ce_a * ce_b + co_a * co_b


def ritf_plane(w, nK, nm, lam):
    """src/directional/ritf.jl:76-84 -- D(w_K + pi, w_m)^lambda, [k, l]."""
    k = np.arange(nK)[:, None]
    l = np.arange(nm)[None, :]
    a = 2 * pi * k / nK + pi
    b = 2 * pi * l / nm
    return ritf_diamond(w, a + 0 * b, b + 0 * a) ** lam


def ritf_masks(dims, msize=15, beta=2.5, lam=4):
    """src/directional/ritf.jl:89-108 -- V_K = sqrt(3 hg_K / sum hg) on the
    half-spectrum of axis K."""
    if lam % 2 or lam <= 0:
        raise ValueError("lambda must be a positive even integer")
    if any(d % 2 for d in dims):
        raise ValueError(f"the hourglass stage needs even dimensions, got {tuple(dims)}")
    w = ritf_window(msize, beta)
    hg = []
    for K in range(3):
        acc = np.ones(dims)
        for m in range(3):
            if m == K:
                continue
            P = ritf_plane(w, dims[K], dims[m], lam)
            shape = [1, 1, 1]
            shape[K] = dims[K]
            shape[m] = dims[m]
            acc = acc * (P if K < m else P.T).reshape(shape)
        hg.append(acc)
    total = hg[0] + hg[1] + hg[2]
    out = []
    for K in range(3):
        sl = [slice(None)] * 3
        sl[K] = slice(0, dims[K] // 2 + 1)
        out.append(np.sqrt(3 * hg[K][tuple(sl)] / total[tuple(sl)]))
    return out
