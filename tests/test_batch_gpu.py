"""Batched analyze+synthesize of uneven sizes (surfhip_ns_batch_*) and the
plan autotuner.

- A batch spread over several lanes gives each volume exactly the
  coefficients and reconstruction that a lone plan of the same size gives it,
  bit for bit, in every mode.
- The recorded graph, replayed several times with other eager work and
  synchronisations in between, reproduces direct launches bit for bit; so
  does a batch run captured into a caller's graph.
- The batch workload at its stated sizes: 64 volumes drawn uniformly from
  {64^3, 80x96x72, 96^3, 112x128x96, 128^3}, ellipsoid phantoms plus
  N(0, 0.05^2), NSSurfacelet(scales=3, direction_levels=[2,2,3]): every
  volume equals its lone transform bit for bit, and one volume of each size
  matches the FP64 oracle (relative L2 1e-5, analysis and reconstruction).
- The autotuner keeps results within FP32 rounding of the untuned plan and
  of the FP64 oracle.
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest
import torch

from oracle import ns as O

surfhip = pytest.importorskip("surfhip")
from surfhip import ns as N  # noqa: E402
from surfhip._arrays import from_numpy, to_numpy  # noqa: E402
from surfhip._lib import ArgumentError  # noqa: E402
from surfhip.batch import NSBatch  # noqa: E402
from surfhip.phantom import ellipsoid_phantom  # noqa: E402

SHAPES = [(40, 48, 32), (32, 32, 32), (48, 40, 36), (32, 32, 32), (40, 48, 32), (64, 32, 48)]
BATCH_SIZES = [(64, 64, 64), (80, 96, 72), (96, 96, 96), (112, 128, 96), (128, 128, 128)]


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return "cuda"


def _inputs(shapes, dev, seed=70):
    rng = np.random.default_rng(seed)
    return [from_numpy(rng.standard_normal(s).astype(np.float32), dev) for s in shapes]


def _lone(t, shapes, xs):
    """Each volume through a lone untuned workspace of its size."""
    wss, cs, ys = {}, [], []
    for s, x in zip(shapes, xs):
        if s not in wss:
            wss[s] = N.make_workspace(t, torch.float32, s, prewarm=False)
        c = N.analyze(wss[s], x)
        cs.append(c)
        ys.append(N.synthesize(wss[s], c))
    return cs, ys


@pytest.mark.gpu
@pytest.mark.parametrize("graph", [True, False])
def test_batch_matches_lone_transforms(graph):
    dev = _dev()
    t = N.NSSurfacelet(scales=3, direction_levels=[2, 2, 3])
    b = NSBatch(t, SHAPES, lanes=3, graph=graph, device=dev)
    assert sorted(set(b.lane_of)) == [0, 1, 2]
    xs = _inputs(SHAPES, dev)
    cs = b.coefficients()
    ys = [torch.empty_like(x) for x in xs]
    b.run(xs, cs, ys)
    torch.cuda.synchronize()
    assert b.graph_recorded() == graph
    cl, yl = _lone(t, SHAPES, xs)
    for k, x in enumerate(xs):
        assert torch.equal(cs[k], cl[k]), k
        assert torch.equal(ys[k], yl[k]), k
        rel = float(torch.linalg.vector_norm(ys[k] - x) / torch.linalg.vector_norm(x))
        assert rel < 1e-5, (k, rel)
    # Analysis alone, then synthesis alone, give the same arrays.
    cs2 = b.coefficients()
    ys2 = [torch.empty_like(x) for x in xs]
    b.run(xs, cs2, mode="analyze")
    b.run(cs=cs2, ys=ys2, mode="synthesize")
    torch.cuda.synchronize()
    for k in range(len(xs)):
        assert torch.equal(cs2[k], cl[k]) and torch.equal(ys2[k], yl[k]), k


@pytest.mark.gpu
def test_batch_graph_replays_match_direct_launches():
    dev = _dev()
    t = N.NSSurfacelet(scales=3, direction_levels=[2, 2, 3])
    xs = _inputs(SHAPES, dev, seed=71)
    bd = NSBatch(t, SHAPES, lanes=3, graph=False, device=dev)
    cs_d = bd.coefficients()
    ys_d = [torch.empty_like(x) for x in xs]
    bd.run(xs, cs_d, ys_d)
    torch.cuda.synchronize()
    b = NSBatch(t, SHAPES, lanes=3, graph=True, device=dev)
    cs = b.coefficients()
    ys = [torch.empty_like(x) for x in xs]
    other = _inputs([(32, 32, 32)], dev, seed=72)[0]
    ws_other = N.make_workspace(t, other)
    for rep in range(3):
        for c in cs:
            c.fill_(float("nan"))
        for y in ys:
            y.fill_(float("nan"))
        b.run(xs, cs, ys)
        N.synthesize(ws_other, N.analyze(ws_other, other))  # eager work between replays
        torch.cuda.synchronize()
        assert b.graph_recorded()
        for k in range(len(SHAPES)):
            assert torch.equal(cs[k], cs_d[k]), (rep, k)
            assert torch.equal(ys[k], ys_d[k]), (rep, k)


@pytest.mark.gpu
def test_batch_inside_a_caller_graph():
    # A batch run on a stream the caller is capturing launches directly and
    # joins the caller's capture; replays reproduce direct launches.
    dev = _dev()
    t = N.NSSurfacelet(scales=3, direction_levels=[2, 2, 3])
    xs = _inputs(SHAPES, dev, seed=73)
    b = NSBatch(t, SHAPES, lanes=3, graph=True, device=dev)
    cs_e = b.coefficients()
    ys_e = [torch.empty_like(x) for x in xs]
    b.run(xs, cs_e, ys_e)
    torch.cuda.synchronize()
    cs = b.coefficients()
    ys = [torch.empty_like(x) for x in xs]
    side = torch.cuda.Stream(dev)
    side.wait_stream(torch.cuda.current_stream(dev))
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=side):
        b.run(xs, cs, ys)
    torch.cuda.synchronize()
    for rep in range(3):
        for c in cs:
            c.fill_(float("nan"))
        g.replay()
        torch.cuda.synchronize()
        for k in range(len(SHAPES)):
            assert torch.equal(cs[k], cs_e[k]) and torch.equal(ys[k], ys_e[k]), (rep, k)


def _batch_set(seed=400, n=64):
    rng = np.random.default_rng(seed)
    idx = rng.integers(0, len(BATCH_SIZES), size=n)
    return [BATCH_SIZES[i] for i in idx], [int(s) for s in rng.integers(0, 2 ** 31, size=n)]


@pytest.mark.gpu
def test_batch_workload_64_uneven_volumes():
    dev = _dev()
    shapes, seeds = _batch_set()
    print("case batch of 64 volumes: " + ", ".join(f"{sum(s == z for s in shapes)} x {z}" for z in BATCH_SIZES),
          flush=True)
    assert set(shapes) == set(BATCH_SIZES)
    vols = [ellipsoid_phantom(s, seed=sd, noise=0.05) for s, sd in zip(shapes, seeds)]
    xs = [from_numpy(v, dev) for v in vols]
    t = N.NSSurfacelet(scales=3, direction_levels=[2, 2, 3])
    b = NSBatch(t, shapes, lanes=4, device=dev)
    cs = b.coefficients()
    ys = [torch.empty_like(x) for x in xs]
    b.run(xs, cs, ys)
    torch.cuda.synchronize()
    cl, yl = _lone(t, shapes, xs)
    for k in range(len(xs)):
        assert torch.equal(cs[k], cl[k]), k
        assert torch.equal(ys[k], yl[k]), k
    # One volume of each size against the FP64 oracle, the five in parallel
    # threads (numpy's FFTs run without the GIL).
    to = O.NSParams(3, (2, 2, 3))
    first = {s: shapes.index(s) for s in BATCH_SIZES}

    def check(s):
        k = first[s]
        x = vols[k].astype(np.float64)
        ref = O.analyze(to, x)
        got = to_numpy(cs[k])
        num = sum(np.sum((got[..., i] - ref[i]) ** 2) for i in range(len(ref)))
        den = sum(np.sum(ref[i] ** 2) for i in range(len(ref)))
        rec = np.linalg.norm(to_numpy(ys[k]) - x) / np.linalg.norm(x)
        return s, float(np.sqrt(num / den)), float(rec)

    with ThreadPoolExecutor(len(BATCH_SIZES)) as ex:
        for s, err, rec in ex.map(check, BATCH_SIZES):
            print(f"batch volume {s}: analysis rel err {err:.2e}, reconstruction rel err {rec:.2e}", flush=True)
            assert err <= 1e-5 and rec <= 1e-5, (s, err, rec)


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


@pytest.mark.gpu
def test_batch_rejects_bad_arrays():
    dev = _dev()
    t = N.NSSurfacelet(scales=3, direction_levels=[2, 2, 3])
    b = NSBatch(t, SHAPES[:2], lanes=2, device=dev)
    xs = _inputs(SHAPES[:2], dev)
    cs = b.coefficients()
    with pytest.raises(ArgumentError):
        b.run(xs[:1], cs, [torch.empty_like(x) for x in xs])
    with pytest.raises(ArgumentError):
        b.run(xs, cs[::-1], mode="analyze")
    with pytest.raises(ArgumentError):
        b.run(xs, cs, mode="synthesize")


def test_batch_rejects_bad_arguments():
    t = N.NSSurfacelet(scales=3, direction_levels=[2, 2, 3])
    with pytest.raises(ArgumentError):
        NSBatch(t, [], lanes=2)
    with pytest.raises(ArgumentError):
        NSBatch(t, [(32, 32, 32)], lanes=0)
