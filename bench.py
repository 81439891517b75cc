"""Throughput of the NSSurfacelet analyze+synthesize pair on one MI355X.

Workload: NSSurfacelet(scales = 3, direction_levels = [2, 2, 3]) (defaults
bernstein_order = 12, halfband_order = 1, transition = 1, tight) on a
128 x 128 x 128 Float32 volume of i.i.d. N(0, 1) voxels (numpy PCG64, seed
20260), 49 subbands.  One step = analyze! + synthesize! with a prebuilt
workspace; the input and the coefficients stay resident in HBM.

metric: voxel-subbands per second = n1 n2 n3 S / (seconds per step).
roofline: the slower of the two plane kernels, timed live with HIP events
that the library records around each of its launches on the stream they run
on (a second, instrumented loop; the headline ms_per_step comes from an
uninstrumented loop of K steps between two events; step_ms, the per-step
median, p10 and p90, comes from another K-step loop with an event between
consecutive steps).  The plane kernels are the ones that touch the user's
data: the inverse plane pass writes the S subbands and the volume (4 N (S+1)
bytes per step), the forward plane pass reads them; achieved = those bytes
over the kernel's summed launch time, against 8 TB/s.  The whole step's
algorithmic bytes, 8 N (S + 1), over ms_per_step are reported beside it.
cpu_baseline: the FP32 CPU restatement of the reference's schedule
(oracle/cpu_schedule.py) on the host cores, on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "surfacelets.jl_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# HBM bytes per launch (read + written, averaged over the class's launches in
# one step) come from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE passes over
# this bench, each scaled by what the counter reads for a plain 256 MiB device
# copy.  The per-class table of that summary is read at run time, so a
# re-measured summary is what the bench line carries.
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "ns128_hbm_bytes.md")
DIMS = (128, 128, 128)
SEED = 20260


def traffic_per_launch(path=TRAFFIC_FILE):
    """{class: HBM bytes per launch} from the per-class table of a counter
    summary (rows '| class | launches | read MiB | written MiB | bytes per
    launch | ratio |'); {} when the file is absent."""
    out = {}
    try:
        with open(path) as fh:
            for ln in fh:
                f = [v.strip() for v in ln.strip().strip("|").split("|")]
                if len(f) == 6 and f[0] and f[0][0].isalpha() and f[1].isdigit():
                    try:
                        out[f[0]] = float(f[4])
                    except ValueError:
                        pass
    except OSError:
        pass
    return out


def step_stats(ms_list):
    """Median, p10, p90 and mean of per-step times (ms)."""
    v = sorted(ms_list)

    def q(f):
        x = f * (len(v) - 1)
        i = int(x)
        return v[i] if i + 1 >= len(v) else v[i] + (x - i) * (v[i + 1] - v[i])
    return {"median_ms": q(0.5), "p10_ms": q(0.1), "p90_ms": q(0.9), "mean_ms": sum(v) / len(v), "steps": len(v)}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="approximate CPU work budget of the cpu_baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    ap.add_argument("--dims", type=lambda v: tuple(int(d) for d in v.split(",")), default=None,
                    metavar="N1,N2,N3",
                    help="ns workload on another volume size (default 128,128,128, the headline); sizes off "
                         "the power-of-two grids run the general-size kernels")
    ap.add_argument("--workload", choices=("ns", "decimated", "batch"), default="ns",
                    help="ns: the NSSurfacelet headline (default); decimated: the decimated Surfacelet "
                         "(scales=3, direction_levels=[2,2,3], pyramid_ratio=2) on the same volume; batch: 64 "
                         "volumes of uneven sizes, the batch against one volume per call")
    ap.add_argument("--lanes", type=int, default=4, help="batch workload: streams of the batch")
    ap.add_argument("--batch-mode", choices=("both", "plain", "batch"), default="both",
                    help="batch workload: time both forms (default), or only one of them (for tracing)")
    return ap.parse_args()


def copy_calibration(mib):
    """One plain device-to-device copy of a known size, for scaling the HBM
    byte counters of rocprofv3 --pmc runs."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": n * 4, "read_bytes": n * 4, "write_bytes": n * 4}), flush=True)


def cpu_baseline(x_np, budget_s, dims=DIMS):
    from oracle import ns as O
    from oracle.cpu_schedule import CPUSchedule
    # The box gives one GPU's job a share of 16 host cores, whatever
    # os.cpu_count() shows; the worker count used is reported beside it.
    workers = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    workers = max(1, min(workers, 16))
    t = O.NSParams(3, (2, 2, 3))
    sch = CPUSchedule(t, dims, workers)
    S = sch.S
    out = [None] * S
    x = x_np.astype(np.float32)
    sch.analyze(x, out)          # warm-up: plans, page faults
    sch.synthesize(out)
    pairs, t0 = 0, time.perf_counter()
    while True:
        sch.analyze(x, out)
        sch.synthesize(out)
        pairs += 1
        if time.perf_counter() - t0 >= budget_s:
            break
    dt = (time.perf_counter() - t0) / pairs
    N = int(np.prod(dims))
    shape = "x".join(str(d) for d in dims)
    return {"value": N * S / dt, "unit": "voxel·subbands/s", "cores": workers, "os_cpu_count": os.cpu_count(),
            "kind": "port",
            "sample": f"{pairs} analyze+synthesize pairs of the same {shape} workload (FP32 scipy.fft, "
                      f"{workers} worker threads, one single-threaded FFT per task), {dt * 1e3:.0f} ms per pair"}


def decimated_cpu_baseline(t_gpu, x_np, budget_s):
    """The FP64 NumPy restatement (oracle/decimated.py) of the decimated pair,
    single-threaded, on whole pairs until the budget is spent."""
    from oracle import decimated as OD
    t = OD.SParams(t_gpu.scales, tuple(t_gpu.direction_levels), pyramid_ratio=t_gpu.pyramid_ratio)
    x = x_np.astype(np.float64)
    pairs, t0 = 0, time.perf_counter()
    while True:
        c = OD.analyze(t, x)
        OD.synthesize(t, c, dims=x.shape)
        pairs += 1
        if time.perf_counter() - t0 >= budget_s:
            break
    dt = (time.perf_counter() - t0) / pairs
    return {"value": 1.0 / dt, "unit": "pairs/s", "cores": 1, "os_cpu_count": os.cpu_count(), "kind": "port",
            "sample": f"{pairs} analyze+synthesize pairs of the same 128^3 workload (FP64 NumPy oracle, "
                      f"one thread), {dt * 1e3:.0f} ms per pair"}


def bench_decimated(args):
    """Decimated Surfacelet(scales=3, direction_levels=[2,2,3], pyramid_ratio=2)
    on the 128^3 volume: one step = analyze! + synthesize! with a prebuilt
    workspace.  Algorithmic bytes per step: the volume and the coefficient
    vector read or written once per direction, 2 (4 N + 4 ncoef)."""
    global np
    import numpy as np
    import torch
    from surfhip import decimated as SD
    from surfhip._arrays import from_numpy

    torch.cuda.set_device(0)
    t = SD.Surfacelet(scales=3, direction_levels=[2, 2, 3], pyramid_ratio=2)
    x_np = np.random.default_rng(SEED).standard_normal(DIMS).astype(np.float32)
    x = from_numpy(x_np, "cuda")
    ws = SD.make_workspace(t, x)
    c = SD.similar_coefficients(t, x)
    y = torch.empty_like(x)

    def step():
        SD.analyze_(c, t, x, workspace=ws)
        SD.synthesize_(y, t, c, workspace=ws)

    for _ in range(max(args.warmup, 20)):
        step()
    torch.cuda.synchronize()
    # The pair issues about 100 launches from the host, about as fast as the
    # GPU runs them, so a host stall shows in a plain mean: the steps are
    # timed in 5 chunks and the median chunk is reported (the mean beside it).
    stream = torch.cuda.current_stream()
    chunks, hosts, per = [], [], max(1, args.steps // 5)
    for _ in range(5):
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        ev0.record(stream)
        h0 = time.perf_counter()
        for _ in range(per):
            step()
        hosts.append((time.perf_counter() - h0) * 1e3 / per)
        ev1.record(stream)
        torch.cuda.synchronize()
        chunks.append(ev0.elapsed_time(ev1) / per)
    ms = sorted(chunks)[2]
    ms_mean = sum(chunks) / 5
    rel = float(torch.linalg.vector_norm(y - x) / torch.linalg.vector_norm(x))
    N = int(np.prod(DIMS))
    ncoef = int(c.data.numel())
    bytes_per_step = 2 * (4 * N + 4 * ncoef)
    achieved = bytes_per_step / (ms * 1e-3) / 1e9
    result = {
        "metric": "decimated Surfacelet analyze+synth pairs/s, 128³ FP32, pyramid_ratio 2, 49 subbands",
        "value": 1e3 / ms,
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": 5 * per,
        "warmup": max(args.warmup, 20),
        "ms_per_step": ms,
        "timing": {"chunks": 5, "steps_per_chunk": per, "chunk_ms_per_step": chunks, "mean_ms_per_step": ms_mean,
                   "host_issue_ms_per_step": hosts, "reported": "median chunk"},
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: i.i.d. N(0,1) Float32 volume, numpy PCG64 seed 20260",
        "config": {"workload": "Surfacelet(scales=3, direction_levels=[2,2,3], pyramid_ratio=2) "
                               "analyze!+synthesize!, 128x128x128 Float32, prebuilt workspace",
                   "dims": list(DIMS), "coefficients": ncoef, "parallelism": "single GPU"},
        "step_roofline": {"achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": achieved / HBM_PEAK_GBS, "bytes_per_step": bytes_per_step},
        "check": {"reconstruction_rel_l2": rel},
    }
    if not args.no_cpu:
        result["cpu_baseline"] = decimated_cpu_baseline(t, x_np, args.cpu_seconds)
    print(json.dumps(result), flush=True)


BATCH_SIZES = [(64, 64, 64), (80, 96, 72), (96, 96, 96), (112, 128, 96), (128, 128, 128)]


def bench_batch(args):
    """64 Float32 volumes drawn uniformly (seed 400) from BATCH_SIZES,
    ellipsoid phantoms plus N(0, 0.05^2), NSSurfacelet(scales=3,
    direction_levels=[2,2,3]).  One step = analyze + synthesize of all 64.
    plain: one volume per call on one stream, one plan per size; batch: the
    library's batch (lanes, one recorded graph replayed per step).  Both are
    timed with HIP events on the current stream, per step."""
    global np
    import numpy as np
    import torch
    from surfhip import ns as NS
    from surfhip._arrays import from_numpy
    from surfhip.batch import NSBatch
    from surfhip.phantom import ellipsoid_phantom

    torch.cuda.set_device(0)
    rng = np.random.default_rng(400)
    idx = rng.integers(0, len(BATCH_SIZES), size=64)
    shapes = [BATCH_SIZES[i] for i in idx]
    seeds = [int(v) for v in rng.integers(0, 2 ** 31, size=64)]
    xs = [from_numpy(ellipsoid_phantom(s, seed=sd, noise=0.05), "cuda") for s, sd in zip(shapes, seeds)]
    t = NS.NSSurfacelet(scales=3, direction_levels=[2, 2, 3])
    wss = {s: NS.make_workspace(t, torch.float32, s, prewarm=False) for s in BATCH_SIZES}
    cs = [NS.similar_coefficients(t, x) for x in xs]
    ys = [torch.empty_like(x) for x in xs]
    b = NSBatch(t, shapes, lanes=args.lanes)
    bc = b.coefficients()
    by = [torch.empty_like(x) for x in xs]
    stream = torch.cuda.current_stream()

    def plain():
        for k, x in enumerate(xs):
            ws = wss[shapes[k]]
            NS.analyze_(cs[k], ws, x)
            NS.synthesize_(ys[k], ws, cs[k])

    def batched():
        b.run(xs, bc, by)

    out = {}
    forms = [f for f in (("plain", plain), ("batch", batched)) if args.batch_mode in ("both", f[0])]
    for name, fn in forms:
        for _ in range(max(2, args.warmup)):
            fn()
        torch.cuda.synchronize()
        evs = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
        evs[0].record(stream)
        for i in range(args.steps):
            fn()
            evs[i + 1].record(stream)
        torch.cuda.synchronize()
        per = [evs[i].elapsed_time(evs[i + 1]) for i in range(args.steps)]
        out[name] = step_stats(per)
    if args.batch_mode != "both":
        print(json.dumps({"batch_mode": args.batch_mode, "step_ms": out[args.batch_mode]}), flush=True)
        return
    same = all(torch.equal(bc[k], cs[k]) and torch.equal(by[k], ys[k]) for k in range(64))
    voxsub = sum(int(np.prod(s)) * NS.nsubbands(t, s) for s in shapes)
    ms = out["batch"]["median_ms"]
    result = {
        "metric": "batched analyze+synth voxel·subbands/s, 64 uneven Float32 volumes, 49 subbands",
        "value": voxsub / (ms * 1e-3),
        "unit": "voxel·subbands/s",
        "n_gpus": 1, "steps": args.steps, "warmup": max(2, args.warmup), "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic: ellipsoid phantoms + N(0, 0.05^2), sizes and phantoms from numpy PCG64 seed 400",
        "config": {"workload": "64 volumes from {64^3, 80x96x72, 96^3, 112x128x96, 128^3}, "
                               "NSSurfacelet(scales=3, direction_levels=[2,2,3]) analyze+synthesize",
                   "counts": {"x".join(map(str, z)): sum(s == z for s in shapes) for z in BATCH_SIZES},
                   "lanes": args.lanes, "parallelism": "single GPU"},
        "plain_step_ms": out["plain"], "batch_step_ms": out["batch"],
        "speedup_vs_plain": out["plain"]["median_ms"] / ms,
        "check": {"batch_equals_plain_bitwise": same},
    }
    print(json.dumps(result), flush=True)


def main():
    args = parse()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the surfacelet path here runs on a single MI355X "
              "(no multi-GPU path is built); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.workload == "decimated":
        bench_decimated(args)
        return
    if args.workload == "batch":
        bench_batch(args)
        return
    global np
    import numpy as np
    import torch
    import surfhip
    from surfhip import ns as NS
    from surfhip._arrays import from_numpy, to_numpy

    dev = "cuda"
    torch.cuda.set_device(0)
    t = NS.NSSurfacelet(scales=3, direction_levels=[2, 2, 3])
    dims = args.dims or DIMS
    x_np = np.random.default_rng(SEED).standard_normal(dims).astype(np.float32)
    x = from_numpy(x_np, dev)
    ws = NS.make_workspace(t, x)
    S = NS.nsubbands(t, dims)
    c = NS.similar_coefficients(t, x)
    y = torch.empty_like(x)
    N = int(np.prod(dims))

    def step():
        NS.analyze_(c, ws, x)
        NS.synthesize_(y, ws, c)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    # The plan's kernels run on torch's current stream (bind_stream), so
    # events recorded on it bracket them.  The headline is K steps between
    # two events, with nothing on the stream between the steps.
    stream = torch.cuda.current_stream()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for _ in range(args.steps):
        step()
    ev1.record(stream)
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / args.steps
    # Per-step times come from a second K-step loop with an event between
    # consecutive steps (each event is one more command on the stream, which
    # the GPU processes between the steps, so these steps read slightly
    # longer than the headline's).
    evs = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
    evs[0].record(stream)
    for i in range(args.steps):
        step()
        evs[i + 1].record(stream)
    torch.cuda.synchronize()
    per_step = [evs[i].elapsed_time(evs[i + 1]) for i in range(args.steps)]
    rel = float(torch.linalg.vector_norm(y - x) / torch.linalg.vector_norm(x))

    # Instrumented loop: per-class kernel times from events around each launch.
    ws.set_timing(True)
    for _ in range(args.steps):
        step()
    cls = ws.timing()
    ws.set_timing(False)
    kernels = {k: {"launches_per_step": n / args.steps, "ms_per_step": t / args.steps} for k, (n, t) in cls.items()}
    # Algorithmic HBM bytes per step of the two kernel classes that touch the
    # user's data: the one that writes the S subbands and the volume (factored
    # path: the inverse plane pass; general path: the axis-1 c2r) and the one
    # that reads the volume and the S subbands (forward plane pass; axis-1 r2c).
    if "plane_inverse" in cls:
        alg = {"plane_inverse": 4 * N * (S + 1), "plane_forward": 4 * N * (S + 1)}
    else:
        alg = {"rows_c2r": 4 * N * (S + 1), "rows_r2c": 4 * N * (S + 1)}
    TRAFFIC = traffic_per_launch()
    dom = max(cls, key=lambda k: cls[k][1])
    dom_plane = max(alg, key=lambda k: cls.get(k, (0, 0.0))[1])
    launches, tms = cls[dom_plane]
    dom_achieved = alg[dom_plane] * args.steps / (tms * 1e-3) / 1e9 if tms > 0 else None

    bytes_per_step = 8 * N * (S + 1)
    achieved = bytes_per_step / (ms * 1e-3) / 1e9
    headline = tuple(dims) == DIMS
    shape = "x".join(str(d) for d in dims)
    result = {
        "metric": ("analyze+synth voxel·subbands/s, 128³ FP32, 49 subbands; % HBM peak" if headline else
                   f"analyze+synth voxel·subbands/s, {shape} FP32, {S} subbands; % HBM peak"),
        "value": N * S / (ms * 1e-3),
        "unit": "voxel·subbands/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "step_ms": dict(step_stats(per_step), source="a second K-step loop, one event between consecutive steps"),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: i.i.d. N(0,1) Float32 volume, numpy PCG64 seed 20260",
        "config": {"workload": "NSSurfacelet(scales=3, direction_levels=[2,2,3]) analyze!+synthesize!, "
                               f"{shape} Float32, {S} subbands, prebuilt workspace",
                   "dims": list(dims), "subbands": S, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": dom_achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": (dom_achieved / HBM_PEAK_GBS) if dom_achieved else None,
                     "traffic": TRAFFIC.get(dom_plane) if headline else None,
                     "traffic_source": "profiles/ns128_hbm_bytes.md" if headline and dom_plane in TRAFFIC else None,
                     "kernel": dom_plane, "launches_per_step": launches / args.steps,
                     "avg_launch_us": tms / launches * 1e3 if launches else None,
                     "dominant_class": dom},
        "step_roofline": {"achieved": achieved, "frac": achieved / HBM_PEAK_GBS,
                          "bytes_per_step": bytes_per_step},
        "kernels": kernels,
        "check": {"reconstruction_rel_l2": rel},
    }
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(x_np, args.cpu_seconds, dims)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
