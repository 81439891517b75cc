"""SurfBox golden data, test volume and fingerprints (test infrastructure only).

Restates the loader and helpers of test/items/test_irc.jl:7-66; the data files
are the reference's own fixtures (test/reference/surfbox_irc.{txt,bin}),
copied unchanged into tests/golden/.
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN_DIR = os.path.join(os.path.dirname(HERE), "tests", "golden")


def _is_complex(name):
    # test/items/test_irc.jl:35 -- spectra (hg_, irc_, pyr_) are complex.
    return name.startswith("hg_") or name.startswith("irc_") or name.startswith("pyr_")


def load_golden(directory=GOLDEN_DIR):
    """test/items/test_irc.jl:15-31 -- records of surfbox_irc.bin in the order
    of surfbox_irc.txt: 'full' arrays (column-major) or fingerprints."""
    out = {}
    raw = np.fromfile(os.path.join(directory, "surfbox_irc.bin"), dtype="<f8")
    pos = 0
    with open(os.path.join(directory, "surfbox_irc.txt")) as fh:
        for ln in fh:
            f = ln.split()
            if not f:
                continue
            name, kind, nd = f[0], f[1], int(f[2])
            sz = tuple(int(v) for v in f[3:3 + nd])
            if kind == "full":
                n = int(np.prod(sz))
                out[name] = raw[pos:pos + n].reshape(sz, order="F")
                pos += n
            else:
                n = 5 if _is_complex(name) else 3
                out[name] = {"size": sz, "fp": raw[pos:pos + n]}
                pos += n
    assert pos == raw.size, (pos, raw.size)
    return out


def testvol(n):
    """test/items/test_irc.jl:39-44 -- exact broadband volume, 1-based i, j, k."""
    i, j, k = np.meshgrid(np.arange(1, n + 1), np.arange(1, n + 1), np.arange(1, n + 1), indexing="ij")
    v = (73 * i ** 2 + 179 * j ** 2 + 283 * k ** 2 + 31 * i * j + 47 * j * k + 61 * k * i) % 1031
    return v / 1031 - 0.5


def fingerprint(A):
    """test/items/test_irc.jl:46-56."""
    A = np.asarray(A)
    i, j, k = np.meshgrid(*(np.arange(1, s + 1) for s in A.shape), indexing="ij")
    s1 = np.sum(A * np.cos(0.37 * i + 0.91 * j + 1.53 * k))
    s2 = np.sum(A * np.sin(1.11 * i - 0.29 * j + 0.77 * k))
    e = np.sum(np.abs(A) ** 2)
    if np.iscomplexobj(A):
        return np.array([s1.real, s1.imag, s2.real, s2.imag, e])
    return np.array([s1, s2, e])


def fp_match(A, ref, rtol=1.0e-11):
    """test/items/test_irc.jl:59-65 -- every entry within rtol of the norm."""
    if tuple(np.shape(A)) != tuple(ref["size"]):
        return False
    f = fingerprint(A)
    fp = ref["fp"]
    if len(f) != len(fp):
        return False
    scale = np.sqrt(fp[-1])
    return bool(np.all(np.abs(f[:-1] - fp[:-1]) <= rtol * scale) and abs(f[-1] - fp[-1]) <= rtol * fp[-1])


def fp_error(A, ref):
    """Worst fingerprint deviation relative to the norm (for reporting)."""
    f = fingerprint(A)
    fp = ref["fp"]
    scale = np.sqrt(fp[-1])
    return float(max(np.max(np.abs(f[:-1] - fp[:-1])) / scale, abs(f[-1] - fp[-1]) / fp[-1]))
