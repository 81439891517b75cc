"""Shared cases of the subband-orientation checks (plain data and helpers; no
test functions, so pytest collects nothing here).

The cases restate the reference's own orientation tests:
- test/items/test_hourglass.jl:113-146: an axis-aligned plane wave lands in
  the hourglass of its axis, and a wave on the (1, 2) diagonal splits evenly
  between hourglasses 1 and 2;
- test/items/test_ndfb.jl:118-147: a plane wave lands in the wedge
  ndfb_subband_index(fb, i, p, q) that its frequency ratios predict;
- test/items/test_surfbox_reference.jl:17-155: the hourglass geometry and the
  direction partition agree with SurfBox's (the two fixtures in
  tests/golden/).
The levels = 3 wedge case and the NSSurfacelet finest-scale case are this
build's own extensions of test_ndfb.jl:118-147: levels = 3 exercises the
(cld, fld) = (2, 1) split of src/directional/ndfb_wedges.jl:19 that the
headline's finest scale uses.
"""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def transverse(i):
    """src/directional/ndfb_wedges.jl:14 (1-based axes)."""
    return (2, 3) if i == 1 else ((1, 3) if i == 2 else (1, 2))


def predict(k, la, lb):
    """test/items/test_ndfb.jl:125-131: (hourglass, p, q) of a plane wave of
    integer frequency k, from its dominant axis and two transverse ratios."""
    i = int(np.argmax(np.abs(k))) + 1
    a, b = transverse(i)
    ki = k[i - 1]
    p = min(max(int(np.floor((k[a - 1] / ki + 1) / (2 / (1 << la)))) + 1, 1), 1 << la)
    q = min(max(int(np.floor((k[b - 1] / ki + 1) / (2 / (1 << lb)))) + 1, 1), 1 << lb)
    return i, p, q


def subband_index(i, p, q, la, lb):
    """src/directional/ndfb.jl:184-190 (1-based)."""
    na, nb = 1 << la, 1 << lb
    return ((i - 1) * na + (p - 1)) * nb + q


def plane_wave(n, k):
    """cos(2 pi k . (I - 1) / n) on an n^3 grid, indexed [i1, i2, i3]."""
    i = np.arange(n)
    g = np.meshgrid(i, i, i, indexing="ij")
    return np.cos(2 * np.pi * (k[0] * g[0] + k[1] * g[1] + k[2] * g[2]) / n)


def axis_wave(n, ax, f):
    """test/items/test_hourglass.jl:119-123: cos(2 pi f t / n) along axis ax."""
    k = [0, 0, 0]
    k[ax - 1] = f
    return plane_wave(n, k)


# test/items/test_ndfb.jl:121,135: n = 32, levels = 4, bernstein_order = 12,
# transition = 0.5, and its six frequencies.
NDFB_L4 = dict(n=32, levels=4, bernstein_order=12, transition=0.5,
               ks=[[2, 5, 9], [5, 9, -12], [-2, 3, -13], [-13, -9, 7], [9, 7, -12], [3, 11, -7]])

# levels = 3 ((la, lb) = (2, 1)), same bank parameters: frequencies with
# distinct magnitudes whose ratios stay off the wedge edges (multiples of 1/2
# on plane a, 0 on plane b), at least one per hourglass and per sign pattern.
NDFB_L3 = dict(n=32, levels=3, bernstein_order=12, transition=0.5,
               ks=[[2, 5, 9], [5, 9, -12], [-2, 3, -13], [-13, -9, 7], [9, 7, -12], [3, 11, -7],
                   [12, -3, 7], [-11, 2, -8]])

# NSSurfacelet(scales = 3, direction_levels = [2, 2, 3]) at 32^3: every
# frequency has its dominant component at |k| >= 11, where the level-1
# pyramid lowpass is exactly zero (its edge P = ceil(16 (1/2 + 1/6)) = 11,
# src/pyramid/fourier_pyramid.jl:47), so the wave is wholly in the finest
# scale (levels 3: (la, lb) = (2, 1)).
NS_FINEST = dict(n=32, scales=3, direction_levels=[2, 2, 3],
                 ks=[[5, 9, -12], [-2, 3, -13], [-13, -9, 7], [9, 7, -12], [12, -3, 7], [-11, 2, -7],
                     [3, 13, 6], [-7, -12, 2]])


def ns_finest_cases():
    """(n, frequencies) of the finest-scale NS check per volume size: 32^3 and
    128^3 run the factored power-of-two kernels (the same ratios scaled by
    4; the level-1 lowpass edge at 128 is 43), 48^3 the general-size kernels
    (edge 16)."""
    ks = NS_FINEST["ks"]
    return [(32, ks), (128, [[4 * v for v in k] for k in ks]),
            (48, [[8, 14, -18], [-3, 5, -20], [-20, -14, 10], [14, 10, -18], [18, -5, 10], [-17, 3, -11],
                  [5, 20, 9], [-10, -18, 3]])]


def ns_finest_offset(direction_levels):
    """offsets_J of src/transforms/ns_surfacelet.jl:230-236,430 (offsets =
    cumsum([1; 3 * 2^L_j])): subband (i, p, q) of the finest scale has the
    1-based index offsets_J + ns_index(i, p, q)."""
    return 1 + sum(3 * (1 << L) for L in direction_levels[:-1])


def energy_shares(bands):
    e = np.array([float(np.sum(np.abs(np.asarray(b, dtype=np.float64)) ** 2)) for b in bands])
    return e / e.sum()


# ------------------------------------------------------------- SurfBox --

def surfbox_hourglass_weights():
    """test/items/test_surfbox_reference.jl:19-40: SurfBox's V_1 octant (17^3
    float32) expanded evenly to 32^3, V_2 and V_3 as its axis permutations,
    and W_k = V_k^2 / 3 on the rfft grid (17, 32, 32)."""
    oct_ = np.fromfile(os.path.join(GOLDEN, "surfbox_hourglass_32.bin"), dtype="<f4").reshape((17, 17, 17), order="F")
    n = 32
    mirror = np.array([i if i < 17 else n - i for i in range(n)])
    v1 = oct_.astype(np.float64)[np.ix_(mirror, mirror, mirror)]
    v = (v1, np.transpose(v1, (1, 0, 2)), np.transpose(v1, (2, 1, 0)))
    tight = v[0] ** 2 + v[1] ** 2 + v[2] ** 2
    return [vk[:17] ** 2 / 3 for vk in v], tight


def check_surfbox_hourglass(w_mine, w_ref, atol_unity=1e-12):
    """test/items/test_surfbox_reference.jl:42-83 on weights w_mine (three
    (17, 32, 32) arrays).  Returns the list of failed checks (empty: pass)."""
    fails = []
    for k in range(3):
        for m in range(3):
            d = float(np.max(np.abs(w_mine[k] - w_ref[m])))
            if (k == m and not d < 0.5) or (k != m and not d > 0.9):
                fails.append(f"max |w_mine[{k + 1}] - w_ref[{m + 1}]| = {d}")
    wm = np.stack(w_mine)
    wr = np.stack(w_ref)
    dm = np.argmax(wm, axis=0)
    dr = np.argmax(wr, axis=0)
    srt = -np.sort(-wr, axis=0)
    gap = srt[0] - srt[1]
    dis = dm != dr
    agree = 1 - dis.sum() / dm.size
    if not agree > 0.94:
        fails.append(f"dominant-hourglass agreement {agree}")
    gaps = np.sort(gap[dis])
    if gaps.size:
        # Julia's gaps[cld(length(gaps), 2)], 1-based.
        if not gaps[(gaps.size + 1) // 2 - 1] < 1e-6:
            fails.append(f"median disagreement gap {gaps[(gaps.size + 1) // 2 - 1]}")
        if not gaps[-1] < 0.48:
            fails.append(f"largest disagreement gap {gaps[-1]}")
    mean_gap = float(gap.mean())
    if not abs(mean_gap - 0.887) <= 0.001:
        fails.append(f"mean reference gap {mean_gap}")
    mean_abs = float(np.abs(wm - wr).sum() / (3 * wm[0].size))
    if not mean_abs < 0.04:
        fails.append(f"mean |w_mine - w_ref| {mean_abs}")
    unity = float(np.max(np.abs(wm.sum(axis=0) - 1)))
    if not unity <= atol_unity:
        fails.append(f"partition of unity off by {unity}")
    return fails


def surfbox_directions():
    """test/items/test_surfbox_reference.jl:103-109: 80 unit directions and
    SurfBox's dominant-subband label of each."""
    raw = np.fromfile(os.path.join(GOLDEN, "surfbox_directions_32.bin"), dtype="<f8")
    nd = 80
    dirs = raw[:3 * nd].reshape((3, nd), order="F")
    labels = raw[3 * nd:].astype(np.int64)
    return dirs, labels


def surfbox_direction_wave(d):
    """test/items/test_surfbox_reference.jl:117-122: the 32^3 plane wave of
    frequency round(d * 0.6 * 16) (Julia rounds half to even, as numpy)."""
    k = np.round(d * (0.6 * 16)).astype(np.int64)
    if np.all(k == 0):
        k = np.array([1, 0, 0])
    return plane_wave(32, k)


def check_surfbox_directions(energies, ref_label):
    """test/items/test_surfbox_reference.jl:111-155 on an (80, 12) energy
    matrix.  Returns the list of failed checks."""
    fails = []
    nd = energies.shape[0]
    if len(np.unique(ref_label)) != 12:
        fails.append("reference labels are not 12 distinct subbands")
    my = np.argmax(energies, axis=1) + 1
    if len(np.unique(my)) != 12:
        fails.append(f"{len(np.unique(my))} distinct dominant subbands, expected 12")
    same = np.sum((ref_label[:, None] == ref_label[None, :]) == (my[:, None] == my[None, :]))
    if not same / nd ** 2 > 0.93:
        fails.append(f"pairwise partition agreement {same / nd ** 2}")
    conf = np.zeros((12, 12), dtype=np.int64)
    for t in range(nd):
        conf[my[t] - 1, ref_label[t] - 1] += 1
    relabel = np.argmax(conf, axis=1) + 1
    if len(np.unique(relabel)) != 12:
        fails.append("confusion matrix induces no bijection")
        return fails
    hit = np.sum(relabel[my - 1] == ref_label) / nd
    if not hit > 0.8:
        fails.append(f"relabelled agreement {hit}")
    inverse = np.zeros(12, dtype=np.int64)
    for m in range(12):
        inverse[relabel[m] - 1] = m + 1
    for t in range(nd):
        top2 = np.argsort(-energies[t], kind="stable")[:2] + 1
        if inverse[ref_label[t] - 1] not in top2:
            fails.append(f"direction {t + 1}: SurfBox's subband is not in the top two here")
    return fails
