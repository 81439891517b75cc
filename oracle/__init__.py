"""FP64 CPU oracle of the surfacelet analyze/synthesize path.

TEST INFRASTRUCTURE ONLY.  This package is a plain NumPy restatement of the
reference algorithms (hakkelt/Surfacelets.jl, cited file:line per function).
It exists to check the HIP library: only ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` may import it, and only as the
checker or the timed CPU baseline -- never as a product code path.

Modules
-------
filters     Bernstein mapping kernels, halfband split, raised-cosine profiles,
            NS wedge planes, checkerboard ladder filters, ritf hourglass.
ns          NSSurfacelet tables and analyze/synthesize (FP64, numpy.fft).
decimated   Surfacelet (decimated) pyramid, IRC stages, blocks, transform.
golden      SurfBox golden-data loader, test volume and fingerprints.
cpu_schedule  FP32 multi-worker restatement of the reference's CPU schedule,
            used only as the timed CPU baseline.

Parity pins (what checks each function, beyond retyping it):
- decimated.*: SurfBox golden fingerprints at rtol 1e-11 / 1e-10
  (tests/golden/surfbox_irc.*), checked in tests/test_oracle_decimated.py.
- filters.cbd_*: DTFT of SurfBox's shipped 2-D filters at 1e-13.
- filters.bernstein_table / halfband_split / mapping_coefficients: the
  reference's own known-answer tests (partition of unity, endpoints, sign
  pattern, complementarity) plus an independent closed form (binomial CDF).
- ns.*: PR, Parseval/adjoint, shift equivariance, constant->lowpass, evenness
  on self-conjugate planes at n in {12, 26, 30}, wedge sums, the literature
  figures (mae(4) = 0.095, mae(20) = 0.024, share(1.0) = 0.71), and a dense
  analysis matrix at 6^3 whose A^T A = I.  No reference-held vector pins NS
  coefficient values bit for bit (the reference's engine package is absent),
  so NS values rest on these invariants and the closed forms.
"""
