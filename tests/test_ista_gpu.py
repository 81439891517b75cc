"""Unrolled ISTA with recomputation checkpointing (surfhip.ista).

- The checkpointed, streaming steps give the same denoised volume and the
  same threshold gradient as plain autograd through the unrolled iteration
  (every intermediate stored), for real and complex volumes.
- Any checkpoint interval gives the same volume and gradient bit for bit
  (the recomputed iterates are the forward's).
- The threshold gradient matches central finite differences of the same
  unrolled ISTA run on the FP64 oracle (oracle.ns analyze/synthesize).
- The unrolled-reconstruction workload at its stated size: 10 iterations,
  NSSurfacelet(scales=4, direction_levels=[2,2,3,3]) (73 subbands), a 256^3
  ellipsoid phantom plus N(0, 0.1^2), lambda = 0.05: the denoised volume is
  closer to the clean phantom than the noisy one, the gradient with respect
  to the four thresholds equals plain autograd's at 2 iterations, and the
  peak of device memory held by coefficient arrays is reported.
"""
import numpy as np
import pytest
import torch

from oracle import ns as O

surfhip = pytest.importorskip("surfhip")
from surfhip import autograd as AG  # noqa: E402
from surfhip import ista as I  # noqa: E402
from surfhip import ns as N  # noqa: E402
from surfhip._arrays import from_numpy, to_numpy  # noqa: E402
from surfhip._lib import ArgumentError  # noqa: E402
from surfhip.phantom import ellipsoid_phantom  # noqa: E402


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return "cuda"


def _volumes(dims, seed, complex_=False):
    clean = ellipsoid_phantom(dims, seed=seed, axes=(2.0, 10.0), noise=0.0).astype(np.float64)
    rng = np.random.default_rng(seed + 1)
    noisy = clean + rng.normal(0, 0.1, dims)
    if complex_:
        ph = np.exp(1j * rng.uniform(0, 2 * np.pi) * np.ones(dims))
        clean, noisy = clean * ph, noisy * ph + 1j * rng.normal(0, 0.1, dims)
    return clean, noisy


def _unrolled(ws, y, theta, iters):
    """Plain autograd through the same iteration (every intermediate kept)."""
    scales = torch.tensor(I.scale_of_subbands(ws.transform, ws.dims), device=y.device)
    th = torch.cat([torch.zeros(1, device=y.device), theta])[scales]
    c = None
    for _ in range(iters):
        r = y if c is None else y - AG.synthesize(ws, c)
        z = AG.analyze(ws, r) if c is None else c + AG.analyze(ws, r)
        shrunk = I.soft_threshold(z, th)
        c = torch.where((scales == 0).view(1, 1, 1, -1), z, shrunk)
    return AG.synthesize(ws, c)


@pytest.mark.gpu
@pytest.mark.parametrize("complex_", [False, True])
def test_ista_matches_unrolled_autograd(complex_):
    dev = _dev()
    dims = (32, 32, 32)
    clean, noisy = _volumes(dims, 50, complex_)
    dt = np.complex64 if complex_ else np.float32
    y = from_numpy(noisy.astype(dt), dev)
    xc = from_numpy(clean.astype(dt), dev)
    t = N.NSSurfacelet(scales=3, direction_levels=[2, 2, 3])
    ws = N.make_workspace(t, y)
    th0 = torch.tensor([0.05, 0.06, 0.07], device=dev)
    out = []
    for fn in (lambda th: I.ista(ws, y, th, iters=4), lambda th: _unrolled(ws, y, th, 4)):
        th = th0.clone().requires_grad_(True)
        x = fn(th)
        loss = torch.sum(torch.abs(x - xc) ** 2)
        loss.backward()
        out.append((to_numpy(x.detach()), th.grad.cpu().numpy().astype(np.float64), float(loss.detach())))
    (xa, ga, la), (xb, gb, lb) = out
    assert np.linalg.norm(xa - xb) <= 1e-5 * np.linalg.norm(xb)
    assert abs(la - lb) <= 1e-5 * abs(lb)
    assert np.all(np.abs(ga - gb) <= 1e-4 * np.max(np.abs(gb))), (ga, gb)
    # The thresholds matter: the gradient is not zero.
    assert np.max(np.abs(gb)) > 0


def _ista_np(to, tb, y, theta, iters):
    S = O.nsubbands(to)
    lay = O.subband_layout(to)
    c = [np.zeros_like(y) for _ in range(S)]
    for k in range(iters):
        r = y - O.synthesize(to, c, tb) if k else y
        ar = O.analyze(to, r, tb)
        for i in range(S):
            z = c[i] + ar[i]
            j = lay[i][0]
            c[i] = z if j == 0 else np.sign(z) * np.maximum(np.abs(z) - theta[j - 1], 0)
    return O.synthesize(to, c, tb)


@pytest.mark.gpu
def test_ista_gradient_matches_oracle_finite_differences():
    dev = _dev()
    dims = (32, 32, 32)
    clean, noisy = _volumes(dims, 60)
    t = N.NSSurfacelet(scales=3, direction_levels=[2, 2, 3])
    to = O.NSParams(3, (2, 2, 3))
    tb = O.tables(to, dims)
    y = from_numpy(noisy.astype(np.float32), dev)
    ws = N.make_workspace(t, y)
    theta = np.array([0.05, 0.06, 0.07])
    th = torch.tensor(theta, dtype=torch.float32, device=dev, requires_grad=True)
    x = I.ista(ws, y, th, iters=3)
    loss = torch.sum((x - from_numpy(clean.astype(np.float32), dev)) ** 2)
    loss.backward()
    g = th.grad.cpu().numpy().astype(np.float64)

    def L(tv):
        return float(np.sum((_ista_np(to, tb, noisy, tv, 3) - clean) ** 2))

    l0 = L(theta)
    assert abs(float(loss.detach()) - l0) <= 1e-4 * l0
    h = 1e-5
    for j in range(3):
        e = np.zeros(3)
        e[j] = h
        fd = (L(theta + e) - L(theta - e)) / (2 * h)
        assert abs(g[j] - fd) <= 1e-3 * max(abs(fd), 1e-3 * np.max(np.abs(g))), (j, g[j], fd)


def test_ista_rejects_bad_arguments():
    with pytest.raises(ArgumentError):
        I.ista(None, None, torch.zeros(3), iters=1)


@pytest.mark.gpu
def test_ista_checkpoint_interval_changes_nothing():
    dev = _dev()
    dims = (32, 32, 32)
    clean, noisy = _volumes(dims, 55)
    y = from_numpy(noisy.astype(np.float32), dev)
    xc = from_numpy(clean.astype(np.float32), dev)
    t = N.NSSurfacelet(scales=3, direction_levels=[2, 2, 3])
    ws = N.make_workspace(t, y)
    out = []
    for every in (1, 2, 5, None):
        th = torch.tensor([0.05, 0.06, 0.07], device=dev, requires_grad=True)
        x = I.ista(ws, y, th, iters=5, checkpoint_every=every)
        torch.sum((x - xc) ** 2).backward()
        out.append((x.detach().clone(), th.grad.clone()))
    for x, g in out[1:]:
        assert torch.equal(x, out[0][0])
        assert torch.equal(g, out[0][1])


@pytest.mark.gpu
def test_ista_256_four_scales_ten_iterations():
    dev = _dev()
    dims = (256, 256, 256)
    print("case ista 256^3, scales 4, 10 iterations", flush=True)
    clean = ellipsoid_phantom(dims, seed=500, noise=0.0).astype(np.float64)
    noisy = clean + np.random.default_rng(501).normal(0, 0.1, dims)
    y = from_numpy(noisy.astype(np.float32), dev)
    xc = from_numpy(clean.astype(np.float32), dev)
    del noisy
    t = N.NSSurfacelet(scales=4, direction_levels=[2, 2, 3, 3])
    ws = N.make_workspace(t, y)
    err0 = float(torch.sum((y - xc) ** 2))
    torch.cuda.synchronize()
    torch.cuda.reset_peak_memory_stats()
    base = torch.cuda.memory_allocated()
    th = torch.full((4,), 0.05, device=dev, requires_grad=True)
    x = I.ista(ws, y, th, iters=10)
    loss = torch.sum((x - xc) ** 2)
    loss.backward()
    torch.cuda.synchronize()
    peak = (torch.cuda.max_memory_allocated() - base) / 2 ** 30
    g = th.grad.detach().cpu().numpy()
    err = float(loss.detach())
    print(f"ista 256^3: squared error {err0:.4e} -> {err:.4e}; d loss / d theta = {g}; "
          f"peak of coefficient arrays {peak:.1f} GiB (one array: "
          f"{256 ** 3 * N.nsubbands(t, dims) * 4 / 2 ** 30:.2f} GiB)", flush=True)
    assert err < 0.5 * err0
    assert np.all(np.isfinite(g)) and np.max(np.abs(g)) > 0
    # Ten saved iterates alone would take 45.6 GiB.
    assert peak < 36.0, peak
    del x, loss
    # At 2 iterations, the checkpointed gradient against plain autograd
    # through the unrolled loop (every intermediate stored).
    res = []
    for fn in (lambda v: I.ista(ws, y, v, iters=2), lambda v: _unrolled(ws, y, v, 2)):
        v = torch.full((4,), 0.05, device=dev, requires_grad=True)
        xo = fn(v)
        lo = torch.sum((xo - xc) ** 2)
        lo.backward()
        res.append((float(lo.detach()), v.grad.detach().cpu().numpy().astype(np.float64)))
        del xo, lo
        torch.cuda.empty_cache()
    (la, ga), (lb, gb) = res
    assert abs(la - lb) <= 1e-5 * abs(lb)
    assert np.all(np.abs(ga - gb) <= 1e-4 * np.max(np.abs(gb))), (ga, gb)
