"""FP64 restatement of the directional filter banks (test infrastructure only).

make_hourglass / hourglass_decompose / hourglass_reconstruct
(src/directional/hourglass.jl:149-496) and make_ndfb / ndfb_decompose /
ndfb_reconstruct (src/directional/ndfb.jl:91-468): nonsubsampled banks whose
subband k is irfft(rfft(x) M_k) on the full grid.  The hourglass bank is the
directional bank with levels = 0 (one wedge per plane, the unit mask:
src/directional/ndfb_wedges.jl:55-58).  Volumes are numpy arrays indexed
[i1, i2, i3], as in oracle.ns.

The hourglass weights and the wedge planes are oracle.filters'
(hourglass_weight, wedge_masks), which the NS oracle tests pin; the checks
of this module's own structure (partition of unity, perfect reconstruction,
Parseval when tight) are in tests/test_oracle_directional.py.
"""
import numpy as np

from . import filters as F
from .ns import irfft3, rfft3


def ndfb_nsubbands(levels):
    """src/directional/ndfb.jl:171-174: 3 * 2^levels."""
    return 3 * (1 << levels)


def ndfb_subband_triple(levels, idx):
    """src/directional/ndfb.jl:147-152,199-203: idx (1-based) -> (hourglass, p, q),
    hourglass first, then p, then q."""
    na, nb = 1 << F.wedge_levels(levels)[0], 1 << F.wedge_levels(levels)[1]
    if not 1 <= idx <= 3 * na * nb:
        raise ValueError("subband index out of range")
    k = idx - 1
    return k // (na * nb) + 1, (k // nb) % na + 1, k % nb + 1


def ndfb_weights(dims, levels, bernstein_order=12, halfband_order=1, transition=1.0):
    """src/directional/ndfb.jl:230-241: W_i w_a,p w_b,q of every subband, in
    index order, on the rfft grid (h1, n2, n3); they sum to 1."""
    m1, m2 = F.hourglass_tables(dims, bernstein_order, halfband_order)
    lev = F.wedge_levels(levels)
    h = (dims[0] // 2 + 1, dims[1], dims[2])
    out = []
    for i in (1, 2, 3):
        W = F.hourglass_weight(m1, m2, i)
        wa, wb = F.wedge_masks(dims, i, lev, halfband_order, transition)
        for a in wa:
            for b in wb:
                out.append(np.broadcast_to(W * a * b, h).copy())
    return out


def ndfb_masks(dims, levels, bernstein_order=12, halfband_order=1, transition=1.0, tight=True):
    """src/directional/hourglass.jl:400-412 (with the wedges of ndfb.jl:222-223):
    the analysis masks, sqrt of the whole weight when tight."""
    w = ndfb_weights(dims, levels, bernstein_order, halfband_order, transition)
    return [F.tight_sqrt(m) for m in w] if tight else w


def decompose(x, masks):
    """src/directional/ndfb.jl:263-346 / hourglass.jl:330-398: subband k =
    irfft(rfft(x) M_k); complex x runs as two real passes."""
    x = np.asarray(x)
    if np.iscomplexobj(x):
        re, im = decompose(x.real, masks), decompose(x.imag, masks)
        return [r + 1j * m for r, m in zip(re, im)]
    X = rfft3(x)
    return [irfft3(X * m, x.shape) for m in masks]


def reconstruct(subbands, masks, tight):
    """src/directional/ndfb.jl:412-468 / hourglass.jl:435-496: a plain sum when
    the bank is not tight (the masks are a partition of unity); otherwise
    irfft(sum_k M_k rfft(s_k))."""
    if not tight:
        return sum(subbands[1:], subbands[0].copy())
    if np.iscomplexobj(subbands[0]):
        return reconstruct([s.real for s in subbands], masks, True) + \
            1j * reconstruct([s.imag for s in subbands], masks, True)
    acc = sum(rfft3(s) * m for s, m in zip(subbands, masks))
    return irfft3(acc, subbands[0].shape)
