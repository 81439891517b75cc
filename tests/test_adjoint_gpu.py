"""Adjoints and reverse-mode rules on the GPU.

- NS: <A x, y> = <x, A^T y> and <S c, z> = <c, S^T z> for tight and not
  tight banks (dense random probes, FP32, relative 1e-5); for a tight bank
  A^T equals synthesize.
- Decimated: A^T against the FP64 oracle adjoint (itself pinned by dense
  probes and a brute-force matrix) and the probe identity; A^T differs from
  synthesize (the transform is not tight).
- Gradient of sum Huber(|c|, 0.1) over the directional coefficients of a
  complex volume (phantom x smooth random phase) through torch.autograd
  against 16 random-direction central finite differences of the FP64 oracle,
  relative 1e-4.
"""
import numpy as np
import pytest
import torch

from oracle import decimated as D
from oracle import ns as O

pytestmark = pytest.mark.gpu

surfhip = pytest.importorskip("surfhip")
from surfhip import autograd as AG  # noqa: E402
from surfhip import decimated as S  # noqa: E402
from surfhip import ns as N  # noqa: E402
from surfhip._arrays import from_numpy, to_numpy  # noqa: E402
from surfhip.phantom import ellipsoid_phantom  # noqa: E402


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return "cuda"


def _dot(a, b):
    return float(torch.sum(a.double() * b.double()))


@pytest.mark.parametrize("dims,tight", [((32, 32, 32), True), ((32, 32, 32), False), ((16, 12, 20), False),
                                        ((128, 128, 128), True)])
def test_ns_adjoint_identities(dims, tight):
    dev = _dev()
    print(f"case ns adjoint dims={dims} tight={tight}", flush=True)
    J, L = (3, [2, 2, 3]) if dims[0] == 128 else (2, [1, 2])
    t = N.NSSurfacelet(scales=J, direction_levels=L, bernstein_order=12 if dims[0] == 128 else 6, tight=tight)
    rng = np.random.default_rng(sum(dims))
    x = from_numpy(rng.standard_normal(dims).astype(np.float32), dev)
    ws = N.make_workspace(t, x)
    c = N.analyze(ws, x)
    y = from_numpy(rng.standard_normal(tuple(c.shape)).astype(np.float32), dev)
    lhs, rhs = _dot(c, y), _dot(x, N.adjoint(ws, y))
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)
    z = from_numpy(rng.standard_normal(dims).astype(np.float32), dev)
    sc = N.synthesize(ws, y)
    lhs2, rhs2 = _dot(sc, z), _dot(y, N.synthesize_adjoint(ws, z))
    assert abs(lhs2 - rhs2) <= 1e-5 * abs(lhs2)
    if tight:
        assert torch.equal(N.adjoint(ws, y), N.synthesize(ws, y))


@pytest.mark.parametrize("dims,mode,J,L,bo", [
    ((32, 32, 32), 2, 2, [1, 2], 4),
    ((48, 24, 36), 1.5, 1, [2], 8),
    ((32, 32, 32), 1, 1, [5], 6),  # per-stage IRC kernel (3 + 2 stages)
])
def test_decimated_adjoint_matches_oracle(dims, mode, J, L, bo):
    dev = _dev()
    t = S.Surfacelet(scales=J, direction_levels=L, pyramid_ratio=mode, checkerboard_order=bo)
    to = D.SParams(scales=J, direction_levels=tuple(L), pyramid_ratio=mode, checkerboard_order=bo)
    ws = S.make_workspace(t, torch.float32, dims)
    rng = np.random.default_rng(3)
    r = rng.standard_normal(ws.layout.length).astype(np.float32)
    c = S.SubbandVector(torch.from_numpy(r).to(dev), ws.layout)
    got = to_numpy(S.adjoint(t, c, workspace=ws)).astype(np.float64)
    bands = [r[o:o + int(np.prod(s))].reshape(s, order="F").astype(np.float64)
             for o, s in zip(ws.layout.offsets, ws.layout.sizes)]
    ref = D.adjoint(to, bands, D.workspace(to, dims))
    assert np.linalg.norm(got - ref) <= 1e-5 * np.linalg.norm(ref)
    x = from_numpy(rng.standard_normal(dims).astype(np.float32), dev)
    a = S.analyze(t, x, workspace=ws)
    lhs, rhs = _dot(a.data, c.data), _dot(x, S.adjoint(t, c, workspace=ws))
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)
    assert np.linalg.norm(to_numpy(S.synthesize(t, c, workspace=ws)) - got) > 1e-3 * np.linalg.norm(got)


def _phase(dims, seed):
    # a smooth random 4th-order polynomial phase
    rng = np.random.default_rng(seed)
    g = np.meshgrid(*(np.linspace(-1, 1, d) for d in dims), indexing="ij")
    phi = np.zeros(dims)
    for _ in range(12):
        p = rng.integers(0, 3, size=3)
        while p.sum() > 4:
            p = rng.integers(0, 3, size=3)
        phi += rng.normal(0, 1) * g[0] ** p[0] * g[1] ** p[1] * g[2] ** p[2]
    return phi


def _huber_np(c, d):
    r = np.abs(c)
    return np.sum(np.where(r <= d, 0.5 * r ** 2, d * (r - 0.5 * d)))


def test_ns_gradient_matches_finite_differences():
    # Config-3 style, at a size the FP64 oracle evaluates quickly.
    dev = _dev()
    dims = (32, 32, 32)
    delta = 0.1
    x0 = ellipsoid_phantom(dims, seed=30, axes=(2.0, 10.0)).astype(np.float64)
    x = (x0 * np.exp(1j * _phase(dims, 31))).astype(np.complex64)
    t = N.NSSurfacelet(scales=3, direction_levels=[2, 2, 3])
    to = O.NSParams(3, (2, 2, 3))
    tb = O.tables(to, dims)
    xd = from_numpy(x, dev).requires_grad_(True)
    ws = N.make_workspace(t, xd.detach())
    c = AG.analyze(ws, xd)
    r = torch.abs(c[..., 1:])
    loss = torch.sum(torch.where(r <= delta, 0.5 * r ** 2, delta * (r - 0.5 * delta)))
    loss.backward()
    g = to_numpy(xd.grad).astype(np.complex128)

    def L(v):
        cc = O.analyze(to, v, tb)
        return sum(_huber_np(b, delta) for b in cc[1:])

    rng = np.random.default_rng(32)
    eps = 1e-4
    x64 = x.astype(np.complex128)
    for k in range(16):
        d = rng.standard_normal(dims) + 1j * rng.standard_normal(dims)
        d /= np.linalg.norm(d)
        fd = (L(x64 + eps * d) - L(x64 - eps * d)) / (2 * eps)
        gd = float(np.real(np.sum(np.conj(g) * d)))
        assert abs(gd - fd) <= 1e-4 * max(abs(fd), 1e-3 * np.linalg.norm(g)), (k, gd, fd)


def _huber_grad(cr, ci, d):
    """dH/dRe + i dH/dIm of H(|c|) (Huber, threshold d): c min(1, d / |c|)."""
    r = np.sqrt(cr ** 2 + ci ** 2)
    f = np.where(r <= d, 1.0, d / np.maximum(r, 1e-300))
    return cr * f, ci * f


def _huber_sum(cr, ci, d):
    r = np.sqrt(cr ** 2 + ci ** 2)
    return float(np.sum(np.where(r <= d, 0.5 * r ** 2, d * (r - 0.5 * d))))


def test_ns_gradient_128_complex_matches_synthesize_and_finite_differences():
    # The gradient workload at its stated size: NSSurfacelet(scales=3,
    # direction_levels=[2,2,3]) on a 128^3 ComplexF32 volume, the ellipsoid
    # phantom (8 ellipsoids, semi-axes U[8, 40], N(0, 0.05^2)) times
    # exp(i phi) with phi a random 4th-order polynomial; loss = sum of
    # Huber(|c|, 0.1) over the directional coefficients.  The rrule gradient
    # (surfhip_ns_adjoint of the loss's gradient) is checked against
    #  - the FP64 oracle's synthesize (= A^T for the tight bank,
    #    src/transforms/ns_surfacelet.jl:21-25) of the oracle's Huber
    #    gradient, relative L2 1e-4;
    #  - 16 random-direction central finite differences of the FP64 oracle's
    #    loss, relative 1e-4.
    # The oracle work runs subband by subband over threads (numpy's FFTs
    # release the GIL): c_k = irfft(X M_k), the Huber gradient's synthesis
    # sum_k rfft(psi(c_k)) M_k, and for each direction the losses at
    # c_k +/- eps irfft(D M_k).
    dev = _dev()
    dims = (128, 128, 128)
    delta, eps = 0.1, 1e-3
    print("case ns gradient 128^3 complex", flush=True)
    x0 = ellipsoid_phantom(dims, seed=33, noise=0.05).astype(np.float64)
    x = (x0 * np.exp(1j * _phase(dims, 34))).astype(np.complex64)
    t = N.NSSurfacelet(scales=3, direction_levels=[2, 2, 3])
    xd = from_numpy(x, dev).requires_grad_(True)
    ws = N.make_workspace(t, xd.detach())
    c = AG.analyze(ws, xd)
    r = torch.abs(c[..., 1:])
    loss = torch.sum(torch.where(r <= delta, 0.5 * r ** 2, delta * (r - 0.5 * delta)))
    loss.backward()
    g = to_numpy(xd.grad).astype(np.complex128)
    loss_gpu = float(loss.detach())
    del c, r, loss

    to = O.NSParams(3, (2, 2, 3))
    tb = O.tables(to, dims)
    masks = O.subband_masks(to, dims, tb)
    x64 = x.astype(np.complex128)
    Xr, Xi = O.rfft3(x64.real), O.rfft3(x64.imag)
    rng = np.random.default_rng(35)
    D = []
    dirs = []
    for _ in range(16):
        d = rng.standard_normal(dims) + 1j * rng.standard_normal(dims)
        d /= np.linalg.norm(d)
        dirs.append(d)
        D.append((O.rfft3(d.real), O.rfft3(d.imag)))

    def subband(k):
        M = masks[k]
        cr, ci = O.irfft3(Xr * M, dims), O.irfft3(Xi * M, dims)
        pr, pi = _huber_grad(cr, ci, delta)
        acc = (O.rfft3(pr) * M, O.rfft3(pi) * M)
        L0 = _huber_sum(cr, ci, delta)
        Lpm = np.zeros((16, 2))
        for j, (Dr, Di) in enumerate(D):
            ar, ai = O.irfft3(Dr * M, dims), O.irfft3(Di * M, dims)
            Lpm[j, 0] = _huber_sum(cr + eps * ar, ci + eps * ai, delta)
            Lpm[j, 1] = _huber_sum(cr - eps * ar, ci - eps * ai, delta)
        return L0, Lpm, acc

    import os
    from concurrent.futures import ThreadPoolExecutor
    accr = np.zeros_like(Xr)
    acci = np.zeros_like(Xi)
    L = 0.0
    Lpm = np.zeros((16, 2))
    with ThreadPoolExecutor(min(16, os.cpu_count() or 1)) as ex:
        for L0, lpm, (ar, ai) in ex.map(subband, range(1, len(masks))):
            L += L0
            Lpm += lpm
            accr += ar
            acci += ai
    g_ref = O.irfft3(accr, dims) + 1j * O.irfft3(acci, dims)
    assert abs(loss_gpu - L) <= 1e-5 * L, (loss_gpu, L)
    err = np.linalg.norm(g - g_ref) / np.linalg.norm(g_ref)
    print(f"gradient vs oracle synthesize: rel L2 {err:.2e}", flush=True)
    assert err <= 1e-4, err
    worst = 0.0
    for j, d in enumerate(dirs):
        fd = (Lpm[j, 0] - Lpm[j, 1]) / (2 * eps)
        gd = float(np.real(np.sum(np.conj(g) * d)))
        rel = abs(gd - fd) / max(abs(fd), 1e-3 * np.linalg.norm(g))
        worst = max(worst, rel)
        assert rel <= 1e-4, (j, gd, fd)
    print(f"16 central differences: worst relative deviation {worst:.2e}", flush=True)
