#!/usr/bin/env python3
"""Benchmark: voxels/s to zero topology mismatch within eps, one MI355X.

Workload (default): 256^3 float32 Gaussian mixture with 64 bumps + 0.01 N(0,1) noise, eps =
1e-2 * range, g = SZ-style quantisation of f, corrected until critical cells and separatrices
(path geometry) match f. A step = one full correction of the uncorrected g (reset from a device
copy at the start of the step) with f and g resident in HBM. value = voxels * steps / wall time.

The JSON line also carries the HBM roofline of the dominant kernel (dense mismatch detection),
measured with HIP events inside the library on the stream the kernel runs on, and the CPU
baseline: the oracle (C++/OpenMP restatement of the reference routines) timed on this host on a
bounded crop of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def measured_traffic(workload, seed):
    """HBM bytes per k_detect launch from the committed rocprofv3 --pmc summary of this same
    workload (FETCH_SIZE + WRITE_SIZE per dispatch, each scaled by the counter's reading on a plain
    copy of known size), or None when no summary for this workload exists."""
    path = os.path.join(ROOT, "profiles", "k_detect_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        t = json.load(fh)
    if t.get("workload") != workload or t.get("seed") != seed:
        return None
    return int(t["traffic_bytes_per_launch"])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--bumps", type=int, default=64)
    ap.add_argument("--field", default="gm", choices=("gm", "fbm"),
                    help="gm: Gaussian mixture + noise; fbm: Gaussian random field with P(k) ~ k^-3 on [0, 1]")
    ap.add_argument("--rel", type=float, default=1e-2)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--geometry", type=int, default=1, help="1: separatrix geometry, 0: critical cells only")
    ap.add_argument("--cpu-sample", type=int, default=64, help="edge of the CPU-baseline crop, 0 = skip")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--dtype", default="f32", choices=("f32", "f64"),
                    help="arithmetic of the correction (f64 = the reference's double precision)")
    ap.add_argument("--max-sweeps", type=int, default=5000)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB four times (the known-size "
                         "reference for scaling the FETCH_SIZE / WRITE_SIZE counters) and exit")
    ap.add_argument("--batch", type=int, default=0, metavar="K",
                    help="time the batched form instead: K fields with sides drawn from 32..192, half "
                         "Gaussian mixtures and half fBm, eps log-uniform in [1e-4, 1e-2]*range, corrected "
                         "by one dmtz_correct_batch call per step")
    ap.add_argument("--learned", default=os.path.join(ROOT, "dmtz_amd", "net_weights.npy"), metavar="NPY",
                    help="also time the learned warm start with these network parameters (from "
                         "python -m dmtz_amd.learned_train) on the same field, and report it beside the plain "
                         "form; 'none' skips it")
    ap.add_argument("--finish", type=int, default=1,
                    help="1: resolve stalls (fixed points where the reference loops forever) by snapping "
                         "growing boxes around the stalled edit set")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (--gpus %d requested); the path would shard "
              "by z-slabs with one-plane halos" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    if args.copy_calibration:
        n = args.copy_calibration * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        for _ in range(4):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": n * 4, "copies": 4}))
        return

    import dmtz_amd
    from dmtz_amd import fields

    if args.batch:
        return bench_batch(args)

    n = args.size
    shape = (n, n, n)
    if args.field == "fbm":
        log("generating %s fBm field" % (shape,))
        f = fields.fbm(shape, seed=args.seed, device="cuda:0").cpu().numpy()
    else:
        log("generating %s Gaussian mixture (%d bumps)" % (shape, args.bumps))
        f = fields.gaussian_mixture(shape, args.bumps, seed=args.seed, noise=0.01)
    eps, rng = fields.eps_from_rel(f, args.rel)
    g0 = fields.quantize(f, eps)
    flags = dmtz_amd.CorrectionFlags(preserve_saddles=1, preserve_vpath=1 if args.geometry else 0,
                                     preserve_geometry=1 if args.geometry else 0, max_sweeps=args.max_sweeps,
                                     finish=args.finish)
    dev = torch.device("cuda:0")
    npdt = np.float64 if args.dtype == "f64" else np.float32
    f_d = torch.from_numpy(f.astype(npdt)).to(dev)
    g0_d = torch.from_numpy(g0.astype(npdt)).to(dev)
    g_d = torch.empty_like(g0_d)
    ctx = dmtz_amd.Context(0)

    def step():
        g_d.copy_(g0_d)
        return dmtz_amd.correct_device(f_d, g_d, eps, rng, flags, ctx=ctx)

    for w in range(args.warmup):
        r = step()
        log("warmup %d: %d sweeps, counts %s, %.3f s, %s" % (w, r.sweeps, r.counts, r.phases["total"], r.info))
    ctx.kernel_stats(reset=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = None
    for s in range(args.steps):
        res = step()
        log("step %d: %d sweeps, counts %s, %s" % (s, res.sweeps, res.counts, res.info))
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    assert res.converged, res.counts
    err = (g_d.double() - f_d.double()).abs().max().item()
    # float32: the snap is the smallest float within eps, so the bound is exact; float64: the
    # reference's snap f - eps (main_3d.cu:1140) is one rounded subtraction
    tol = eps if args.dtype == "f32" else eps + float(np.spacing(np.abs(f).max() + eps))
    assert err <= tol, (err, eps)
    N = n * n * n
    value = N * args.steps / elapsed
    ks = ctx.kernel_stats()
    launches = max(ks["launches"], 1.0)
    avg_ms = ks["ms"] / launches
    bytes_per_launch = ks["bytes"] / launches
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    kind = "fBm (P(k)~k^-3, [0,1])" if args.field == "fbm" else "gaussian-mixture"
    workload = "%s %d^3 %s, %seps=%g*range, %s" % (
        kind, n, "float32" if args.dtype == "f32" else "float64",
        "" if args.field == "fbm" else "%d bumps, " % args.bumps, args.rel,
        "critical cells + separatrix geometry" if args.geometry else "critical cells")
    roofline = dict(bound="hbm", kernel="k_detect (dense mismatch detection)", achieved=round(achieved, 2),
                    peak=HBM_PEAK_GBS, unit="GB/s", frac=round(achieved / HBM_PEAK_GBS, 5),
                    traffic=measured_traffic(workload, args.seed), bytes_per_launch=int(bytes_per_launch),
                    avg_launch_ms=round(avg_ms, 4), launches=int(ks["launches"]))

    # The whole correction priced like one dense sweep per sweep (f and g read once, 8 bytes per
    # mismatch record emitted): the path's HBM fraction end to end, beside the k_detect kernel's.
    traj = np.asarray(res.trajectory).reshape(-1, 8) if res.trajectory is not None else np.zeros((0, 8))
    records = float(traj[:, :4].sum()) if len(traj) else 0.0
    b_run = res.sweeps * N * 2.0 * (8 if args.dtype == "f64" else 4) + 8.0 * records
    run_gbs = b_run / (elapsed / args.steps) / 1e9
    sweep_roofline = dict(bound="hbm", achieved=round(run_gbs, 2), peak=HBM_PEAK_GBS, unit="GB/s",
                          frac=round(run_gbs / HBM_PEAK_GBS, 5), bytes_per_correction=int(b_run),
                          basis="sweeps * (f and g once) + 8 B per mismatch record, over the whole correction")

    learned = None
    if args.learned and args.learned != "none" and os.path.exists(args.learned):
        learned = bench_learned(args, ctx, f_d, g0_d, g_d, eps, rng, flags, value)

    cpu = None
    if args.cpu_sample > 0:
        from oracle import oracle as O

        O.build()
        threads = O.set_threads(min(args.cpu_threads, os.cpu_count() or 1))
        s = min(args.cpu_sample, n)
        fc = np.ascontiguousarray(f[:s, :s, :s])
        gc = np.ascontiguousarray(g0[:s, :s, :s])
        log("cpu baseline: oracle on a %d^3 crop, %d threads" % (s, threads))
        c0 = time.perf_counter()
        rc = O.correct(fc, gc, eps, rng, O.flags(preserve_saddles=1, preserve_vpath=flags.preserve_vpath,
                                                  preserve_geometry=flags.preserve_geometry,
                                                  max_sweeps=args.max_sweeps, finish=args.finish))
        c1 = time.perf_counter()
        cpu = dict(value=round(s ** 3 / (c1 - c0), 1), unit="voxels/s", cores=threads, kind="port",
                   sample="%d^3 crop (corner) of the same field, same eps and flags: %d sweeps in %.2f s"
                          % (s, rc["sweeps"], c1 - c0))

    out = {
        "metric": "voxels/s to zero topology mismatch within eps (plain GPU form, 1 GPU)",
        "value": round(value, 1),
        "unit": "voxels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.dtype,
        "data": "synthetic",
        "config": {"workload": workload,
                   "size": [n, n, n], "bumps": args.bumps, "rel_eps": args.rel, "seed": args.seed,
                   "sweeps": res.sweeps, "form": "plain", "finish": args.finish,
                   "stalls": res.info.get("stalls"), "finisher_vertices": res.info.get("finished")},
        "roofline": roofline,
        "sweep_roofline": sweep_roofline,
        "cpu_baseline": cpu,
    }
    if learned is not None:
        out["learned"] = learned
    print(json.dumps(out), flush=True)


def bench_learned(args, ctx, f_d, g0_d, g_d, eps, rng, flags, plain_value):
    """The learned warm start on the same field: features, forward pass, warm edits and the plain
    loop from there, timed like the plain steps."""
    import numpy as np
    import torch

    from dmtz_amd import learned as L

    net = L.Net(seed=0, ctx=ctx)
    net.set_params(np.load(args.learned))

    def step():
        g_d.copy_(g0_d)
        return L.correct_learned(net, f_d, g_d, eps, rng, flags)

    for _ in range(max(1, args.warmup)):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(args.steps):
        res, edits = step()
        log("learned step %d: %d sweeps, %d warm edits, counts %s" % (s, res.sweeps, edits, res.counts))
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    assert res.converged, res.counts
    assert (g_d.double() - f_d.double()).abs().max().item() <= eps
    value = f_d.numel() * args.steps / elapsed
    net.close()
    return dict(value=round(value, 1), unit="voxels/s", ms_per_step=round(elapsed / args.steps * 1e3, 3),
                sweeps=res.sweeps, warm_edits=edits, speedup_vs_plain=round(value / plain_value, 3),
                params=os.path.basename(args.learned))


def batch_fields(k, seed):
    """k fields: sides drawn independently from 32..192, alternately a Gaussian mixture (8 bumps,
    0.01 noise) and an fBm field generated on the next power-of-two cube and cropped; eps drawn
    log-uniform in [1e-4, 1e-2]*range."""
    import numpy as np
    from dmtz_amd import fields

    r = np.random.default_rng([seed, 48])
    out = []
    for i in range(k):
        shape = tuple(int(x) for x in r.integers(32, 193, size=3))
        rel = float(10.0 ** r.uniform(-4.0, -2.0))
        if i % 2 == 0:
            f = fields.gaussian_mixture(shape, 8, seed=seed * 1000 + i, noise=0.01)
        else:
            p = 1 << int(np.ceil(np.log2(max(shape))))
            full = fields.fbm((p, p, p), seed=seed * 1000 + i, device="cuda:0").cpu().numpy()
            f = np.ascontiguousarray(full[:shape[0], :shape[1], :shape[2]], dtype=np.float32)
        eps, rng = fields.eps_from_rel(f, rel)
        out.append((f, fields.quantize(f, eps), eps, rng, rel))
    return out


def bench_batch(args):
    """The batched form: one dmtz_correct_batch call per step over K device-resident fields."""
    import numpy as np
    import torch

    import dmtz_amd

    log("generating %d batch fields" % args.batch)
    items = batch_fields(args.batch, args.seed)
    dev = torch.device("cuda:0")
    fd = [torch.from_numpy(f).to(dev) for f, _, _, _, _ in items]
    g0 = [torch.from_numpy(np.ascontiguousarray(g, dtype=np.float32)).to(dev) for _, g, _, _, _ in items]
    gd = [torch.empty_like(x) for x in g0]
    flags = dmtz_amd.CorrectionFlags(preserve_saddles=1, preserve_vpath=1 if args.geometry else 0,
                                     preserve_geometry=1 if args.geometry else 0, max_sweeps=args.max_sweeps,
                                     finish=args.finish)
    ctx = dmtz_amd.Context(0)
    batch = [(fd[i], gd[i], items[i][2], items[i][3]) for i in range(len(items))]

    def step():
        for a, b in zip(gd, g0):
            a.copy_(b)
        return dmtz_amd.correct_batch(batch, flags, ctx=ctx)

    for w in range(args.warmup):
        rs = step()
        log("warmup %d: sweeps %s" % (w, [r.sweeps for r in rs]))
    ctx.kernel_stats(reset=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(args.steps):
        rs = step()
        log("step %d: total sweeps %d" % (s, sum(r.sweeps for r in rs)))
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    for i, r in enumerate(rs):
        assert r.converged, (i, r.counts)
        err = (gd[i].double() - fd[i].double()).abs().max().item()
        assert err <= items[i][2], (i, err)
    voxels = sum(int(f.size) for f, _, _, _, _ in items)
    ks = ctx.kernel_stats()
    launches = max(ks["launches"], 1.0)
    avg_ms = ks["ms"] / launches
    bytes_per_launch = ks["bytes"] / launches
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    workload = ("batch of %d float32 fields, sides 32..192, Gaussian mixture / fBm, eps log-uniform "
                "[1e-4, 1e-2]*range, %s" % (args.batch, "critical cells + separatrix geometry"
                                             if args.geometry else "critical cells"))
    cpu = None
    if args.cpu_sample > 0:
        from oracle import oracle as O

        O.build()
        threads = O.set_threads(min(args.cpu_threads, os.cpu_count() or 1))
        k = int(np.argmin([f.size for f, _, _, _, _ in items]))
        f, g, eps, rng, _ = items[k]
        log("cpu baseline: oracle on the smallest field %s, %d threads" % (f.shape, threads))
        c0 = time.perf_counter()
        rc = O.correct(f, g, eps, rng, O.flags(preserve_saddles=1, preserve_vpath=flags.preserve_vpath,
                                               preserve_geometry=flags.preserve_geometry,
                                               max_sweeps=args.max_sweeps, finish=args.finish))
        c1 = time.perf_counter()
        cpu = dict(value=round(f.size / (c1 - c0), 1), unit="voxels/s", cores=threads, kind="port",
                   sample="the batch's smallest field %s: %d sweeps in %.2f s" % (f.shape, rc["sweeps"], c1 - c0))
    out = {
        "metric": "voxels/s to zero topology mismatch within eps (batched form, 1 GPU)",
        "value": round(voxels * args.steps / elapsed, 1),
        "unit": "voxels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": workload, "fields": args.batch, "voxels": voxels, "seed": args.seed,
                   "sweeps": [r.sweeps for r in rs], "form": "batched (one call, fields in up to 8 concurrent lanes)",
                   "finish": args.finish},
        # lanes overlap on the device, so k_detect's event-timed launches include other lanes' work
        "roofline": dict(bound="hbm", kernel="k_detect (dense mismatch detection; concurrent lanes)",
                         achieved=round(achieved, 2),
                         peak=HBM_PEAK_GBS, unit="GB/s", frac=round(achieved / HBM_PEAK_GBS, 5), traffic=None,
                         bytes_per_launch=int(bytes_per_launch), avg_launch_ms=round(avg_ms, 4),
                         launches=int(ks["launches"])),
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
