"""Batched analyze+synthesize of uneven sizes and the plan autotuner.

- A batch spread over several lanes (streams) gives each volume exactly the
  coefficients and reconstruction that the same workspace gives it alone.
- A captured batch graph, replayed several times with other eager work and
  synchronisations in between, reproduces the eager results bit for bit on
  every replay.
- The autotuner keeps results within FP32 rounding of the untuned plan and
  of the FP64 oracle.
"""
import numpy as np
import pytest
import torch

from oracle import ns as O

surfhip = pytest.importorskip("surfhip")
from surfhip import ns as N  # noqa: E402
from surfhip._arrays import from_numpy, to_numpy  # noqa: E402
from surfhip._lib import ArgumentError  # noqa: E402
from surfhip.batch import NSBatch  # noqa: E402

SHAPES = [(40, 48, 32), (32, 32, 32), (48, 40, 36), (32, 32, 32), (40, 48, 32), (64, 32, 48)]


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return "cuda"


def _inputs(shapes, dev, seed=70):
    rng = np.random.default_rng(seed)
    return [from_numpy(rng.standard_normal(s).astype(np.float32), dev) for s in shapes]


@pytest.mark.gpu
def test_batch_matches_per_volume_transforms():
    dev = _dev()
    t = N.NSSurfacelet(scales=3, direction_levels=[2, 2, 3])
    b = NSBatch(t, SHAPES, lanes=3, tune=True, device=dev)
    xs = _inputs(SHAPES, dev)
    cs = b.coefficients()
    ys = [torch.empty_like(x) for x in xs]
    b.run(xs, cs, ys)
    torch.cuda.synchronize()
    for k, x in enumerate(xs):
        ws = b.workspace(k)
        c1 = N.analyze(ws, x)
        y1 = N.synthesize(ws, c1)
        assert torch.equal(cs[k], c1), k
        assert torch.equal(ys[k], y1), k
        rel = float(torch.linalg.vector_norm(ys[k] - x) / torch.linalg.vector_norm(x))
        assert rel < 1e-5, (k, rel)


@pytest.mark.gpu
def test_batch_graph_replays_match_eager():
    dev = _dev()
    t = N.NSSurfacelet(scales=3, direction_levels=[2, 2, 3])
    b = NSBatch(t, SHAPES, lanes=3, tune=False, device=dev)
    xs = _inputs(SHAPES, dev, seed=71)
    cs_e = b.coefficients()
    ys_e = [torch.empty_like(x) for x in xs]
    b.run(xs, cs_e, ys_e)
    torch.cuda.synchronize()
    cs = b.coefficients()
    ys = [torch.empty_like(x) for x in xs]
    g = b.capture(xs, cs, ys)
    other = _inputs([(32, 32, 32)], dev, seed=72)[0]
    ws_other = N.make_workspace(t, other)
    for rep in range(3):
        for c in cs:
            c.fill_(float("nan"))
        for y in ys:
            y.fill_(float("nan"))
        g.replay()
        N.synthesize(ws_other, N.analyze(ws_other, other))  # eager work between replays
        torch.cuda.synchronize()
        for k in range(len(SHAPES)):
            assert torch.equal(cs[k], cs_e[k]), (rep, k)
            assert torch.equal(ys[k], ys_e[k]), (rep, k)


@pytest.mark.gpu
@pytest.mark.parametrize("dims", [(40, 48, 36), (48, 40, 24)])
def test_tuned_plan_matches_untuned_and_oracle(dims):
    dev = _dev()
    t = N.NSSurfacelet(scales=2, direction_levels=[2, 3])
    to = O.NSParams(2, (2, 3))
    x_np = np.random.default_rng(73).standard_normal(dims)
    x = from_numpy(x_np.astype(np.float32), dev)
    ws0 = N.make_workspace(t, x)
    ws1 = N.make_workspace(t, x, tune=True)
    c0, c1 = N.analyze(ws0, x), N.analyze(ws1, x)
    ref = np.stack(O.analyze(to, x_np), axis=-1)
    for c in (c0, c1):
        err = np.linalg.norm(to_numpy(c).astype(np.float64) - ref) / np.linalg.norm(ref)
        assert err < 1e-5, err
    assert float(torch.linalg.vector_norm(c1 - c0) / torch.linalg.vector_norm(c0)) < 1e-5
    y1 = N.synthesize(ws1, c1)
    assert float(torch.linalg.vector_norm(y1 - x) / torch.linalg.vector_norm(x)) < 1e-5


def test_batch_rejects_bad_arguments():
    t = N.NSSurfacelet(scales=3, direction_levels=[2, 2, 3])
    with pytest.raises(ArgumentError):
        NSBatch(t, [], lanes=2)
    with pytest.raises(ArgumentError):
        NSBatch(t, [(32, 32, 32)], lanes=0)
