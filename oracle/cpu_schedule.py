"""FP32 CPU restatement of the reference's NS schedule, timed as the CPU
baseline (test/bench infrastructure only; never a product path).

The reference's CPU path (MultiscaleFilterBanks' MaskBankEngine, as described
at src/transforms/ns_surfacelet.jl:23-32) splits the subbands into contiguous
chunks, one task per FFT buffer with single-threaded FFT plans: rfft once, a
group product B_j W_i when a chunk enters a group, then per subband a wedge
product and an irfft; synthesis mirrors it with per-task partial sums.  This
module runs that schedule with scipy.fft (pocketfft) in FP32 on a thread
pool (scipy's FFTs release the GIL).  It is a restatement, not the reference:
Julia and the reference's sibling packages are not available on the machine.
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import scipy.fft as sfft

from . import ns as O


class CPUSchedule:
    def __init__(self, t, dims, workers):
        self.t = t
        self.dims = tuple(dims)
        self.workers = max(1, int(workers))
        tb = O.tables(t, dims)
        self.S = O.nsubbands(t)
        self.items = []          # (subband index 0-based, group id, wa, wb)
        self.gmask_a = {}        # group id -> analysis group mask (FP32)
        self.gmask_s = {}
        for g, (j, i, idx) in enumerate(tb.groups):
            self.gmask_a[g] = O.group_mask(t, tb, j, i, True).astype(np.float32)
            self.gmask_s[g] = O.group_mask(t, tb, j, i, False).astype(np.float32)
            wa, wb = tb.wedges[j - 1][i - 1]
            for (k, a, b) in idx:
                self.items.append((k - 1, g, wa[a - 1].astype(np.float32), wb[b - 1].astype(np.float32)))
        self.low = O.lowpass_mask(t, tb).astype(np.float32)
        n = len(self.items)
        w = min(self.workers, n)
        bounds = np.linspace(0, n, w + 1).astype(int)
        self.chunks = [self.items[bounds[q]:bounds[q + 1]] for q in range(w)]
        self.pool = ThreadPoolExecutor(max_workers=w)

    def _rfft(self, x):
        return sfft.rfftn(x, axes=(2, 1, 0), workers=1)

    def _irfft(self, X):
        d = self.dims
        return sfft.irfftn(X, s=(d[2], d[1], d[0]), axes=(2, 1, 0), workers=1).astype(np.float32)

    def analyze(self, x, out):
        X = self._rfft(x).astype(np.complex64)

        def run(chunk):
            gcur, Y = None, None
            for (k, g, wa, wb) in chunk:
                if g != gcur:
                    Y = X * self.gmask_a[g]
                    gcur = g
                out[k] = self._irfft(Y * wa * wb)

        list(self.pool.map(run, self.chunks))
        out[0] = self._irfft(X * self.low)
        return out

    def synthesize(self, c):
        tight = self.t.tight

        def run(chunk):
            acc = None
            gcur, accg = None, None
            for (k, g, wa, wb) in chunk:
                if g != gcur:
                    if accg is not None:
                        acc = accg * self.gmask_s[gcur] if acc is None else acc + accg * self.gmask_s[gcur]
                    gcur, accg = g, None
                F = self._rfft(c[k]).astype(np.complex64)
                term = F * (wa * wb) if tight else F
                accg = term if accg is None else accg + term
            if accg is not None:
                acc = accg * self.gmask_s[gcur] if acc is None else acc + accg * self.gmask_s[gcur]
            return acc

        parts = [p for p in self.pool.map(run, self.chunks) if p is not None]
        acc = parts[0]
        for p in parts[1:]:
            acc = acc + p
        acc = acc + self._rfft(c[0]).astype(np.complex64) * self.low
        return self._irfft(acc)
