#!/usr/bin/env python3
"""Benchmark: voxels/s to zero topology mismatch within eps, one MI355X.

Workload (default): 512^3 float32 fBm field (Gaussian random field with P(k) ~ k^-3, normalised
to [0, 1]), eps = 1e-2 * range, g = SZ-style quantisation of f (2 eps bins), corrected until
critical cells and separatrices (path geometry) match f; stalls (fixed points where the
reference loops forever) are resolved by the finisher. A step = one full correction of the
uncorrected g (reset from a device copy at the start of the step) with f and g resident in HBM.
value = voxels * steps / wall time.

The JSON line also carries
  roofline        the per-sweep HBM fraction: the algorithmic bytes of every sweep (f and g read
                  once, 8 B per mismatch record) over the correction's wall time, against 8 TB/s,
                  with the HBM traffic per sweep measured by rocprofv3 when a summary of this
                  workload is committed under profiles/;
  kernel_roofline the dense detection kernel (k_detect) alone, priced by the bytes it must move,
                  timed with HIP events on its stream;
  edit_ratio      edited voxels / N (the reference's edit_ratio, main_3d.cu:1553);
  other_eps       the same field at the other eps of the scale configuration;
  learned         the learned warm start on the same field, and the uniform warm start beside it;
  batched         the batched configuration (48 fields, sides 32..192): plain (one field after
                  another), batched, batched+learned and batched with the uniform warm start, and
                  each one's speed-up over plain, the metric's batched+learned vs plain ratio;
  cpu_baseline    the oracle (C++/OpenMP restatement of the reference routines) on the CPU-runnable
                  workload (64^3 Gaussian mixture, 8 bumps, eps = 1e-3 * range) on all the host cores
                  it may use and on one thread.
Device buffers come from the HIP runtime the library runs on (dmtz_amd.devmem); torch is not used.
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def committed_profile(name, workload, seed):
    """A committed rocprofv3 summary (JSON under profiles/) of this same workload, or None."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        t = json.load(fh)
    if t.get("workload") != workload or t.get("seed") != seed:
        return None
    return t


def workload_name(field, n, dtype, rel, bumps, geometry, finish):
    kind = "fBm (P(k)~k^-3, [0,1])" if field == "fbm" else "gaussian-mixture"
    return "%s %d^3 %s, %seps=%g*range, %s%s" % (
        kind, n, "float32" if dtype == "f32" else "float64", "" if field == "fbm" else "%d bumps, " % bumps, rel,
        "critical cells + separatrix geometry" if geometry else "critical cells",
        ", stalls resolved by the finisher" if finish else "")


def make_field(field, n, seed, bumps):
    from dmtz_amd import fields

    shape = (n, n, n)
    if field == "fbm":
        log("generating %s fBm field" % (shape,))
        return fields.fbm(shape, seed=seed)
    log("generating %s Gaussian mixture (%d bumps)" % (shape, bumps))
    return fields.gaussian_mixture(shape, bumps, seed=seed, noise=0.01)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def usable_cores():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))  # the host share of one GPU is 16 cores


def cpu_baseline(one_thread=True):
    """The oracle on the CPU-runnable workload: 64^3 Gaussian mixture (8 bumps, 0.01 noise), eps =
    1e-3 * range, critical cells + separatrix geometry, seed 0; on all usable cores and on one."""
    from dmtz_amd import fields
    from oracle import oracle as O

    O.build()
    f = fields.gaussian_mixture((64, 64, 64), 8, seed=0, noise=0.01)
    eps, rng = fields.eps_from_rel(f, 1e-3)
    g = fields.quantize(f, eps)
    fl = O.flags(preserve_saddles=1, preserve_vpath=1, preserve_geometry=1)
    runs = []
    for want in [usable_cores()] + ([1] if one_thread else []):
        threads = O.set_threads(want)
        log("cpu baseline: oracle, 64^3 Gaussian mixture, eps=1e-3*range, %d threads" % threads)
        c0 = time.perf_counter()
        r = O.correct(f, g, eps, rng, fl)
        dt = time.perf_counter() - c0
        assert r["status"] == 0 and tuple(r["counts"]) == (0,) * 8
        runs.append((threads, 64 ** 3 / dt, r["sweeps"], dt, r["phases"]))
    threads, v, sw, dt, ph = runs[0]
    # phases under the reference's timer names (main_3d.cu:7187-7188), in seconds
    res = dict(value=round(v, 1), unit="voxels/s", cores=threads, kind="port",
               sample="64^3 Gaussian mixture, 8 bumps, eps=1e-3*range, geometry, seed 0 (the CPU-runnable "
                      "workload), full correction: %d sweeps in %.2f s" % (sw, dt),
               cpu_model=cpu_model(), nproc=os.cpu_count(),
               phases_s={k: round(x, 4) for k, x in ph.items()})
    if one_thread:
        _, v1, _, dt1, ph1 = runs[1]
        res["one_thread"] = dict(value=round(v1, 1), unit="voxels/s", seconds=round(dt1, 2),
                                 phases_s={k: round(x, 4) for k, x in ph1.items()})
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--bumps", type=int, default=64)
    ap.add_argument("--field", default="fbm", choices=("gm", "fbm"),
                    help="gm: Gaussian mixture + noise; fbm: Gaussian random field with P(k) ~ k^-3 on [0, 1]")
    ap.add_argument("--rel", type=float, default=1e-2)
    ap.add_argument("--other-rel", default="1e-3,1e-4",
                    help="comma-separated further eps/range values timed on the same field ('' = none)")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--geometry", type=int, default=1, help="1: separatrix geometry, 0: critical cells only")
    ap.add_argument("--cpu", default="full", choices=("full", "all-cores", "none"),
                    help="CPU baseline: oracle on the 64^3 workload on all cores and one thread, all cores only, "
                         "or none")
    ap.add_argument("--dtype", default="f32", choices=("f32", "f64"),
                    help="arithmetic of the correction (f64 = the reference's double precision)")
    ap.add_argument("--max-sweeps", type=int, default=5000)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB four times (the known-size "
                         "reference for scaling the FETCH_SIZE / WRITE_SIZE counters) and exit")
    ap.add_argument("--batch-leg", type=int, default=48, metavar="K",
                    help="fields of the batched leg reported in the main line (0: none)")
    ap.add_argument("--batch", type=int, default=0, metavar="K",
                    help="time the batched form instead: K fields with sides drawn from 32..192, half "
                         "Gaussian mixtures and half fBm, eps log-uniform in [1e-4, 1e-2]*range")
    ap.add_argument("--learned", default=os.path.join(ROOT, "dmtz_amd", "net_weights.npy"), metavar="NPY",
                    help="also time the learned warm start with these network parameters (from "
                         "python -m dmtz_amd.learned_train) on the same field; 'none' skips it")
    ap.add_argument("--finish", type=int, default=1,
                    help="1: resolve stalls (fixed points where the reference loops forever) by snapping "
                         "growing boxes around the stalled edit set")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (--gpus %d requested); the path would shard "
              "by z-slabs with one-plane halos" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import dmtz_amd
    from dmtz_amd import devmem

    if args.copy_calibration:
        n = args.copy_calibration * (1 << 20) // 4
        src = devmem.DeviceArray((n,), np.float32)
        dst = devmem.DeviceArray((n,), np.float32)
        src.fill_zero()
        devmem.synchronize()
        for _ in range(4):
            dst.copy_from(src)
        devmem.synchronize()
        print(json.dumps({"copy_bytes": n * 4, "copies": 4}))
        return
    if args.batch:
        return bench_batch(args)

    from dmtz_amd import fields

    n = args.size
    N = n * n * n
    f = make_field(args.field, n, args.seed, args.bumps)
    npdt = np.float64 if args.dtype == "f64" else np.float32
    f = f.astype(npdt)
    flags = dmtz_amd.CorrectionFlags(preserve_saddles=1, preserve_vpath=1 if args.geometry else 0,
                                     preserve_geometry=1 if args.geometry else 0, max_sweeps=args.max_sweeps,
                                     finish=args.finish)
    ctx = dmtz_amd.Context(0)
    f_d = devmem.DeviceArray.from_numpy(f)
    g0_d = devmem.DeviceArray(f.shape, npdt)
    g_d = devmem.DeviceArray(f.shape, npdt)

    def run_plain(rel, steps, warmup, tag):
        eps, rng = fields.eps_from_rel(f, rel)
        g0 = fields.quantize(f, eps)
        g0_d.upload(g0)

        def step():
            g_d.copy_from(g0_d)
            return dmtz_amd.correct_device(f_d, g_d, eps, rng, flags, ctx=ctx)

        for w in range(warmup):
            r = step()
            log("%s warmup %d: %d sweeps, counts %s, %.3f s, %s" % (tag, w, r.sweeps, r.counts, r.phases["total"],
                                                                    r.info))
        ctx.kernel_stats(reset=True)
        devmem.synchronize()
        t0 = time.perf_counter()
        res = None
        for s in range(steps):
            res = step()
            log("%s step %d: %d sweeps, counts %s, %s" % (tag, s, res.sweeps, res.counts, res.info))
        devmem.synchronize()
        elapsed = time.perf_counter() - t0
        ks = ctx.kernel_stats()
        assert res.converged, res.counts
        gout = g_d.numpy()
        err = float(np.abs(gout.astype(np.float64) - f.astype(np.float64)).max())
        # float32: the snap is the smallest float within eps, so the bound is exact; float64: the
        # reference's snap f - eps (main_3d.cu:1140) is one rounded subtraction
        tol = eps if args.dtype == "f32" else eps + float(np.spacing(np.abs(f).max() + eps))
        assert err <= tol, (err, eps)
        edited = int(np.count_nonzero(gout != g0))
        return dict(eps=eps, rng=rng, g0=g0, res=res, elapsed=elapsed, ks=ks, edited=edited)

    main_run = run_plain(args.rel, args.steps, args.warmup, "plain")
    res, elapsed, ks = main_run["res"], main_run["elapsed"], main_run["ks"]
    value = N * args.steps / elapsed
    t_corr = elapsed / args.steps
    workload = workload_name(args.field, n, args.dtype, args.rel, args.bumps, args.geometry, args.finish)

    # Per-sweep roofline: every sweep's algorithmic bytes (f and g read once, 8 B per mismatch
    # record emitted) over the correction's wall time (per-run set-up included).
    elem = 8 if args.dtype == "f64" else 4
    traj = np.asarray(res.trajectory).reshape(-1, 8) if res.trajectory is not None else np.zeros((0, 8))
    records = float(traj.sum()) if len(traj) else 0.0
    b_run = res.sweeps * N * 2.0 * elem + 8.0 * records
    sweep_gbs = b_run / t_corr / 1e9
    sweep_prof = committed_profile("sweep_traffic.json", workload, args.seed)
    roofline = dict(bound="hbm", achieved=round(sweep_gbs, 2), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(sweep_gbs / HBM_PEAK_GBS, 5),
                    traffic=int(sweep_prof["traffic_bytes_per_sweep"]) if sweep_prof else None,
                    unit_of_work="one correction sweep over N voxels",
                    bytes_per_sweep=int(b_run / max(res.sweeps, 1)),
                    ms_per_sweep=round(t_corr / max(res.sweeps, 1) * 1e3, 4), sweeps=res.sweeps,
                    basis="algorithmic bytes of every sweep (N*2*sizeof(T) + 8 B per mismatch record) over the "
                          "wall time of the whole correction",
                    dominant_kernel=sweep_prof.get("dominant_kernel") if sweep_prof else None)

    launches = max(ks["launches"], 1.0)
    avg_ms = ks["ms"] / launches
    bytes_per_launch = ks["bytes"] / launches
    det_gbs = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    det_prof = committed_profile("k_detect_traffic.json", workload, args.seed)
    kernel_roofline = dict(bound="hbm", kernel="k_detect (mismatch detection, once per sweep)",
                           achieved=round(det_gbs, 2), peak=HBM_PEAK_GBS, unit="GB/s",
                           frac=round(det_gbs / HBM_PEAK_GBS, 5),
                           traffic=int(det_prof["traffic_bytes_per_launch"]) if det_prof else None,
                           bytes_per_launch=int(bytes_per_launch), avg_launch_ms=round(avg_ms, 4),
                           launches=int(ks["launches"]),
                           basis="4 B count per 2048-voxel tile + the 16-byte mismatch masks of the tiles with a mismatch read + 8 B per record written")

    other = {}
    for r in [x for x in args.other_rel.split(",") if x.strip()]:
        rel = float(r)
        o = run_plain(rel, max(1, min(args.steps, 3)), 1, "eps=%g" % rel)
        t = o["elapsed"] / max(1, min(args.steps, 3))
        other["%g" % rel] = dict(value=round(N / t, 1), unit="voxels/s", ms_per_step=round(t * 1e3, 3),
                                 sweeps=o["res"].sweeps, stalls=o["res"].info.get("stalls"),
                                 edit_ratio=round(o["edited"] / N, 6),
                                 sweep_frac=round((o["res"].sweeps * N * 2.0 * elem
                                                   + 8.0 * float(np.asarray(o["res"].trajectory).sum()))
                                                  / t / 1e9 / HBM_PEAK_GBS, 5))

    learned = None
    if args.learned and args.learned != "none" and os.path.exists(args.learned) and args.dtype == "f32":
        learned = bench_learned(args, ctx, f, f_d, main_run, g0_d, g_d, flags, value)

    batched = None
    if args.batch_leg > 0 and args.dtype == "f32":
        forms, voxels, bworkload = batch_forms(args, args.batch_leg, max(1, min(args.steps, 3)), 1)
        plain_v = forms["plain"]["value"]
        batched = dict(workload=bworkload, voxels=voxels, forms=forms,
                       speedup_vs_plain={k: round(v["value"] / plain_v, 3) for k, v in forms.items()})

    cpu = None if args.cpu == "none" else cpu_baseline(one_thread=args.cpu == "full")

    out = {
        "metric": "voxels/s to zero topology mismatch within eps (plain GPU form, 1 GPU)",
        "value": round(value, 1),
        "unit": "voxels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_corr * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.dtype,
        "data": "synthetic",
        "config": {"workload": workload, "size": [n, n, n], "field": args.field, "rel_eps": args.rel,
                   "seed": args.seed, "sweeps": res.sweeps, "form": "plain", "finish": args.finish,
                   "stalls": res.info.get("stalls"), "finisher_vertices": res.info.get("finished")},
        "roofline": roofline,
        "kernel_roofline": kernel_roofline,
        "edit_ratio": round(main_run["edited"] / N, 6),
        "other_eps": other,
        "cpu_baseline": cpu,
    }
    if learned is not None:
        out["learned"] = learned
    if batched is not None:
        out["batched"] = batched
    print(json.dumps(out), flush=True)


def bench_learned(args, ctx, f, f_d, main_run, g0_d, g_d, flags, plain_value):
    """The learned warm start on the same field: features, forward pass, warm edits and the plain
    loop from there, timed like the plain steps."""
    import numpy as np

    from dmtz_amd import devmem
    from dmtz_amd import learned as L

    eps, rng, g0 = main_run["eps"], main_run["rng"], main_run["g0"]
    g0_d.upload(g0)
    net = L.Net(seed=0, ctx=ctx)
    net.set_params(np.load(args.learned))

    def step():
        g_d.copy_from(g0_d)
        return L.correct_learned(net, f_d, g_d, eps, rng, flags)

    N = f.size

    def timed(tag, steps):
        for _ in range(max(1, args.warmup)):
            step()
        devmem.synchronize()
        t0 = time.perf_counter()
        for s in range(steps):
            res, edits = step()
            log("%s step %d: %d sweeps, %d warm edits, counts %s" % (tag, s, res.sweeps, edits, res.counts))
        devmem.synchronize()
        elapsed = time.perf_counter() - t0
        assert res.converged, res.counts
        gout = g_d.numpy()
        assert float(np.abs(gout.astype(np.float64) - f.astype(np.float64)).max()) <= eps
        value = N * steps / elapsed
        return dict(value=round(value, 1), unit="voxels/s", ms_per_step=round(elapsed / steps * 1e3, 3),
                    sweeps=res.sweeps, warm_edits=edits, edit_ratio=round(int(np.count_nonzero(gout != g0)) / N, 6),
                    speedup_vs_plain=round(value / plain_value, 3))

    out = timed("learned", args.steps)
    out.update(plain_edit_ratio=round(main_run["edited"] / N, 6), params=os.path.basename(args.learned))
    # the trivial warm start the network is measured against: every vertex lowered by eps
    net.set_params(L.uniform_params(1.0))
    out["uniform_warm_start"] = timed("uniform warm start", max(1, min(args.steps, 3)))
    net.close()
    return out


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
            full = fields.fbm((p, p, p), seed=seed * 1000 + i)
            f = np.ascontiguousarray(full[:shape[0], :shape[1], :shape[2]], dtype=np.float32)
        eps, rng = fields.eps_from_rel(f, rel)
        out.append((f, fields.quantize(f, eps), eps, rng, rel))
    return out


def batch_forms(args, k, steps, warmup):
    """The batched forms on k device-resident fields, timed against the plain form on the same
    fields (one field after another): plain, batched (one dmtz_correct_batch call),
    batched+learned (dmtz_correct_batch_learned_f32: the warm start on every field, then the
    batch) and the batch with the uniform warm start. Returns (forms, voxels, workload)."""
    import numpy as np

    import dmtz_amd
    from dmtz_amd import devmem
    from dmtz_amd import learned as L

    log("generating %d batch fields" % k)
    items = batch_fields(k, args.seed)
    fd = [devmem.DeviceArray.from_numpy(f) for f, _, _, _, _ in items]
    g0 = [devmem.DeviceArray.from_numpy(np.ascontiguousarray(g, dtype=np.float32)) for _, g, _, _, _ in items]
    gd = [devmem.DeviceArray(x.shape, np.float32) for x in g0]
    flags = dmtz_amd.CorrectionFlags(preserve_saddles=1, preserve_vpath=1 if args.geometry else 0,
                                     preserve_geometry=1 if args.geometry else 0, max_sweeps=args.max_sweeps,
                                     finish=args.finish)
    ctx = dmtz_amd.Context(0)
    batch = [(fd[i], gd[i], items[i][2], items[i][3]) for i in range(len(items))]
    net = uniform = None
    if args.learned and args.learned != "none" and os.path.exists(args.learned):
        net = L.Net(seed=0, ctx=ctx)
        net.set_params(np.load(args.learned))
        uniform = L.Net(seed=0, ctx=ctx)  # the trivial warm start, for comparison
        uniform.set_params(L.uniform_params(1.0))

    def reset():
        for a, b in zip(gd, g0):
            a.copy_from(b)

    def plain():
        return [dmtz_amd.correct_device(f, g, eps, rng, flags, ctx=ctx) for f, g, eps, rng in batch]

    def batched():
        return dmtz_amd.correct_batch(batch, flags, ctx=ctx)

    def batched_learned():
        return L.correct_batch_learned(net, batch, flags)[0]

    def batched_uniform():
        return L.correct_batch_learned(uniform, batch, flags)[0]

    voxels = sum(int(f.size) for f, _, _, _, _ in items)
    forms = [("plain", plain), ("batched", batched)]
    if net:
        forms += [("batched_learned", batched_learned), ("batched_uniform_warm_start", batched_uniform)]
    out_forms = {}
    for name, fn in forms:
        for w in range(warmup):
            reset()
            fn()
        t = 0.0
        for s in range(steps):
            reset()
            devmem.synchronize()
            t0 = time.perf_counter()
            rs = fn()
            devmem.synchronize()
            t += time.perf_counter() - t0
        edited = 0
        for i, r in enumerate(rs):
            assert r.converged, (name, i, r.counts)
            gi = gd[i].numpy()
            err = float(np.abs(gi.astype(np.float64) - items[i][0].astype(np.float64)).max())
            assert err <= items[i][2], (name, i, err)
            edited += int(np.count_nonzero(gi != items[i][1]))
        out_forms[name] = dict(value=round(voxels * steps / t, 1), ms_per_step=round(t / steps * 1e3, 3),
                               sweeps=int(sum(r.sweeps for r in rs)), edit_ratio=round(edited / voxels, 6))
        log("%s: %.1f ms per batch, %d sweeps" % (name, t / steps * 1e3, out_forms[name]["sweeps"]))
    if net:
        net.close()
        uniform.close()
    workload = ("batch of %d float32 fields, sides 32..192, Gaussian mixture / fBm, eps log-uniform "
                "[1e-4, 1e-2]*range, %s" % (k, "critical cells + separatrix geometry"
                                         if args.geometry else "critical cells"))
    return out_forms, voxels, workload


def bench_batch(args):
    """The batched forms as the line: value is the batched+learned rate; the ratio to the plain
    form is the metric's."""
    out_forms, voxels, workload = batch_forms(args, args.batch, args.steps, args.warmup)
    best = "batched_learned" if "batched_learned" in out_forms else "batched"
    plain_v = out_forms["plain"]["value"]
    out = {
        "metric": "voxels/s to zero topology mismatch within eps (%s vs plain, 1 GPU)" % best.replace("_", "+"),
        "value": out_forms[best]["value"],
        "unit": "voxels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": out_forms[best]["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": workload, "fields": args.batch, "voxels": voxels, "seed": args.seed,
                   "form": best, "finish": args.finish},
        "forms": out_forms,
        "speedup_vs_plain": {k: round(v["value"] / plain_v, 3) for k, v in out_forms.items()},
        "edit_ratio": out_forms[best]["edit_ratio"],
        "roofline": None,
        "cpu_baseline": None if args.cpu == "none" else cpu_baseline(one_thread=False),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
