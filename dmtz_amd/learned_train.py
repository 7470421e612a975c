"""Train the warm-start network on pairs from the plain correction and evaluate its gain.

    python -m dmtz_amd.learned_train --fields 200 --size 128 --epochs 12 --pos-weight 0.35 \
        --out dmtz_amd/net_weights.npy --eval-size 384 --eval 4

Training fields alternate fBm (P(k) ~ k^-3) and Gaussian mixtures (16 bumps, 0.01 noise) with eps
drawn log-uniform in [1e-3, 1e-2] * range; each field's labels come from one plain correction
(separatrix geometry and finisher). Evaluation corrects held-out fields of --eval-size both
ways and prints one JSON line per field and a summary line.
"""
import argparse
import json
import sys
import time

import numpy as np

from .learned import EDIT_LOGIT_BIAS as LOGIT_BIAS


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_field(kind, shape, seed, device):
    from . import fields

    if kind == "fbm":
        return fields.fbm(shape, seed=seed, device=device).contiguous()
    import torch

    return torch.from_numpy(fields.gaussian_mixture(shape, 16, seed=seed, noise=0.01)).to(device)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--fields", type=int, default=200)
    ap.add_argument("--size", type=int, default=128)
    ap.add_argument("--epochs", type=int, default=12)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--lam", type=float, default=0.5)
    ap.add_argument("--pos-weight", type=float, default=0.35,
                    help="weight of the edited class in the loss; below 1 a false edit costs more than a missed one, "
                         "which keeps the warm edits near the vertices the plain loop edits "
                         "(0: sqrt(1 / edited fraction), capped to [1, 20])")
    ap.add_argument("--logit-shift", type=float, default=0.0,
                    help="subtract this from the edit logit's bias of the saved parameters: the warm start then "
                         "edits only vertices with logit above it")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default="")
    ap.add_argument("--init", default="", help="start from these parameters (.npy)")
    ap.add_argument("--eval", type=int, default=4)
    ap.add_argument("--eval-size", type=int, default=384)
    args = ap.parse_args(argv)

    import torch

    dev = torch.device("cuda:0")
    torch.cuda.init()
    import dmtz_amd
    from . import fields
    from . import learned as L

    ctx = dmtz_amd.Context(0)
    net = L.Net(seed=args.seed, ctx=ctx)
    if args.init:
        net.set_params(np.load(args.init))
    flags = dmtz_amd.CorrectionFlags(preserve_vpath=1, preserve_geometry=1, finish=1)
    rs = np.random.default_rng([args.seed, 5])
    data = []
    t0 = time.perf_counter()
    shape = (args.size,) * 3
    for i in range(args.fields):
        kind = "fbm" if i % 2 == 0 else "gm"
        rel = float(10.0 ** rs.uniform(-3.0, -2.0))
        f = make_field(kind, shape, 10000 + args.seed * 1000 + i, dev)
        eps, rng = fields.eps_from_rel(f, rel)
        g = fields.quantize(f, eps)
        feat, lab = L.training_pair(net, f, g, eps, rng, flags, ctx=ctx)
        data.append((feat, lab))
        if i % 10 == 9:
            log("pairs %d/%d (%.1f s)" % (i + 1, args.fields, time.perf_counter() - t0))
    ypos = float(np.mean([lab[..., 0].mean().item() for _, lab in data]))
    pos = args.pos_weight if args.pos_weight > 0 else float(min(20.0, max(1.0, np.sqrt(1.0 / max(ypos, 1e-4)))))
    log("edited fraction %.4f, pos_weight %.2f" % (ypos, pos))
    steps = args.epochs * len(data)
    k = 0
    for e in range(args.epochs):
        order = rs.permutation(len(data))
        tot = np.zeros(2)
        for j in order:
            lr = args.lr * 0.5 * (1.0 + np.cos(np.pi * k / steps))
            tot += net.train_step(data[j][0], data[j][1], lr=lr, pos_weight=pos, lam=args.lam)
            k += 1
        log("epoch %d: bce %.4f size %.4f (%.1f s)" % (e, tot[0] / len(data), tot[1] / len(data),
                                                       time.perf_counter() - t0))
    if args.logit_shift:
        p = net.params()
        p[LOGIT_BIAS] -= args.logit_shift
        net.set_params(p)
    if args.out:
        np.save(args.out, net.params())
        log("saved %s" % args.out)
    del data

    summary = []
    es = (args.eval_size,) * 3
    if args.eval:  # size the context's and the network's buffers before timing anything
        f = make_field("fbm", es, 89999, dev)
        eps, rng = fields.eps_from_rel(f, 1e-2)
        g = fields.quantize(f, eps)
        dmtz_amd.correct_device(f, g.clone(), eps, rng, flags, ctx=ctx)
        L.correct_learned(net, f, g.clone(), eps, rng, flags)
    for i in range(args.eval):
        kind = "fbm" if i % 2 == 0 else "gm"
        rel = [1e-2, 3e-3][(i // 2) % 2]
        f = make_field(kind, es, 90000 + i, dev)
        eps, rng = fields.eps_from_rel(f, rel)
        g = fields.quantize(f, eps)
        g1, g2 = g.clone(), g.clone()
        torch.cuda.synchronize()
        a = time.perf_counter()
        plain = dmtz_amd.correct_device(f, g1, eps, rng, flags, ctx=ctx)
        b = time.perf_counter()
        warm, edits = L.correct_learned(net, f, g2, eps, rng, flags)
        c = time.perf_counter()
        ok = plain.converged and warm.converged and (g2.double() - f.double()).abs().max().item() <= eps
        n = g.numel()
        row = dict(field=kind, size=args.eval_size, rel_eps=rel, plain_sweeps=plain.sweeps,
                   plain_ms=round((b - a) * 1e3, 2), warm_sweeps=warm.sweeps, warm_ms=round((c - b) * 1e3, 2),
                   warm_edits=edits, speedup=round((b - a) / (c - b), 3), converged=bool(ok),
                   plain_edit_ratio=round((g1 != g).sum().item() / n, 4),
                   warm_edit_ratio=round((g2 != g).sum().item() / n, 4))
        print(json.dumps(row), flush=True)
        summary.append(row)
    if summary:
        tp = sum(r["plain_ms"] for r in summary)
        tw = sum(r["warm_ms"] for r in summary)
        print(json.dumps(dict(summary="eval", fields=len(summary), plain_ms=round(tp, 2), warm_ms=round(tw, 2),
                              speedup=round(tp / tw, 3))), flush=True)
    net.close()


if __name__ == "__main__":
    main()
