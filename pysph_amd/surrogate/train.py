"""Training of the edge-MLP surrogate on rollouts of the plain GPU solver.

    python -m pysph_amd.surrogate.train --out weights.npz [--rollouts 2000]

Rollouts: `bases` 2D dam breaks (column width in [0.3, 1.0] m, height in
[0.5, 2.0] m, 20k-100k fluid particles, drawn from the seed) are advanced by
the plain fused WCSPH solver.  The surrogate steps positions by leapfrog with
the step DT = dt_factor x dt: its velocity state is the mean velocity of the
last step, u_{n-1/2} = (x_n - x_{n-1}) / DT, and one step is
u_{n+1/2} = u_{n-1/2} + DT a_n, x_{n+1} = x_n + DT u_{n+1/2}.  Every
`spacing` plain steps one sample is recorded from three plain positions DT
apart: the state (x_n, u_{n-1/2}) and the acceleration that carries the
leapfrog step onto the plain solver's next positions,
a_n = (x_{n+1} - 2 x_n + x_{n-1}) / DT^2, in units of g.  Matching positions
rather than velocities carries everything the plain step moves particles by
(the RK2 path, the XSPH correction) into the target.  The network is
trained with Adam on the GPU (the library's own forward and backward
kernels).
"""
import argparse
import json
import math
import time

import numpy as np
import torch

from ..solver import generators as gen
from .graph import ParticleGraph, combined_array
from .model import EdgeMLP
from .solver import STATE, get_user, plain_solver


def dam_scene(rng, n_range=(20000, 100000), hdx=1.3):
    width = float(rng.uniform(0.3, 1.0))
    height = float(rng.uniform(0.5, 2.0))
    n = int(round(math.exp(rng.uniform(math.log(n_range[0]), math.log(n_range[1])))))
    dx = math.sqrt(width * height / n)
    f, b, info = gen.dam_break_2d(width=width, height=height, dx=dx, hdx=hdx,
                                  precision="single")
    info.update(width=width, height=height)
    return f, b, info


def _positions(pa):
    return get_user(pa, "x").double(), get_user(pa, "y").double()


def generate(n_rollouts, bases, seed, dt_factor=4, spacing=100, log=print, n_range=None,
             warmup=1000):
    """Returns a list of bases: dict(graph, info, samples=[(state, target)]).
    Each base first runs `warmup` plain steps (the summation-density free
    surface starts far from equilibrium).  Samples follow the leapfrog
    convention of the module docstring."""
    rng = np.random.Generator(np.random.PCG64(seed))
    per = max(1, n_rollouts // bases)
    out = []
    for bi in range(bases):
        f, b, info = dam_scene(rng, n_range or (20000, 100000))
        graph = ParticleGraph(combined_array(f, b), info, domain=info["domain"])
        s = plain_solver(f, b, info)
        nf, nb = f.get_number_of_particles(), b.get_number_of_particles()
        bstate = {k: get_user(graph.pa, k)[nf:] for k in STATE}
        DT = dt_factor * info["dt"]
        samples = []
        t0 = time.time()
        for _ in range(warmup):
            s.integrator.integrate(info["dt"])
        for r in range(per):
            for _ in range(spacing - 2 * dt_factor):
                s.integrator.integrate(info["dt"])
            xm, ym = _positions(f)
            for _ in range(dt_factor):
                s.integrator.integrate(info["dt"])
            x0, y0 = _positions(f)
            for _ in range(dt_factor):
                s.integrator.integrate(info["dt"])
            xp, yp = _positions(f)
            fl = {"x": x0, "y": y0, "u": (x0 - xm) / DT, "v": (y0 - ym) / DT}
            st = {k: torch.cat([fl[k].float(), bstate[k]]) for k in STATE}
            scale = 1.0 / (DT * DT * 9.81)
            tgt = torch.zeros((nf + nb, 2), dtype=torch.float32, device=x0.device)
            tgt[:nf, 0] = ((xp - 2.0 * x0 + xm) * scale).float()
            tgt[:nf, 1] = ((yp - 2.0 * y0 + ym) * scale).float()
            samples.append((st, tgt))
        torch.cuda.synchronize()
        log("base %d: %d fluid + %d boundary particles, W=%.2f H=%.2f, %d rollouts in %.1f s"
            % (bi, f.get_number_of_particles(), nb, info["width"], info["height"], per,
               time.time() - t0))
        out.append({"graph": graph, "info": info, "samples": samples})
    return out


def train(model, data, epochs, lr=1e-3, lr_min=1e-5, log=print, val=None):
    stream = torch.cuda.current_stream().cuda_stream
    order = [(bi, si) for bi, d in enumerate(data) for si in range(len(d["samples"]))]
    total = epochs * len(order)
    rng = np.random.Generator(np.random.PCG64(1234))
    step = 0
    hist = []
    t0 = time.time()
    for ep in range(epochs):
        perm = rng.permutation(len(order))
        acc = []
        for k in perm:
            bi, si = order[k]
            d = data[bi]
            st, tgt = d["samples"][si]
            g = d["graph"]
            g.set_state(st)
            g.build()
            model.zero_grad(stream)
            loss = model.loss_backward(g, g.user_rows_to_slot(tgt))
            frac = step / max(total - 1, 1)
            model.adam(stream, lr=lr_min + 0.5 * (lr - lr_min) * (1 + math.cos(math.pi * frac)))
            acc.append(loss)
            step += 1
            if step % 200 == 0:
                m = float(torch.stack(acc[-200:]).mean().item())
                log("epoch %d step %d/%d loss %.4e (%.0f s)" % (ep, step, total, m,
                                                               time.time() - t0))
        m = float(torch.stack(acc).mean().item())
        entry = {"epoch": ep, "train_loss": m}
        if val is not None:
            entry["val_loss"] = evaluate(model, val)
        hist.append(entry)
        log(json.dumps(entry))
    return hist


def evaluate(model, data):
    """Mean loss and the relative error ||a_pred - a|| / ||a|| over fluid."""
    num, den, ls = 0.0, 0.0, []
    for d in data:
        g = d["graph"]
        for st, tgt in d["samples"]:
            g.set_state(st)
            g.build()
            y = g.slot_rows_to_user(model.forward(g, infer=True)[:g.n])
            nf = g.n_fluid
            e = (y[:nf] - tgt[:nf])
            num += float((e * e).sum().item())
            den += float((tgt[:nf] * tgt[:nf]).sum().item())
            ls.append(float((e * e).sum().item()) / nf)
    return {"loss": float(np.mean(ls)), "rel_err": math.sqrt(num / max(den, 1e-30))}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--rollouts", type=int, default=2000)
    ap.add_argument("--bases", type=int, default=20)
    ap.add_argument("--spacing", type=int, default=100)
    ap.add_argument("--epochs", type=int, default=6)
    ap.add_argument("--lr", type=float, default=3e-4)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--val-bases", type=int, default=2)
    ap.add_argument("--n-min", type=int, default=20000)
    ap.add_argument("--n-max", type=int, default=100000)
    ap.add_argument("--eval-n", type=int, default=30000)
    ap.add_argument("--eval-steps", type=int, default=100)
    a = ap.parse_args(argv)
    log = lambda s: print(s, flush=True)  # noqa: E731
    t0 = time.time()
    data = generate(a.rollouts, a.bases, a.seed, spacing=a.spacing, log=log,
                    n_range=(a.n_min, a.n_max))
    val = generate(a.val_bases * 20, a.val_bases, a.seed + 1000, spacing=a.spacing * 5, log=log,
                   n_range=(a.n_min, a.n_max))
    log("data: %.1f s" % (time.time() - t0))
    model = EdgeMLP(seed=a.seed)
    log("initial val: %s" % json.dumps(evaluate(model, val)))
    hist = train(model, data, a.epochs, lr=a.lr, log=log, val=val)
    model.save(a.out, history=json.dumps(hist), rollouts=a.rollouts, bases=a.bases,
               seed=a.seed, dt_factor=4, epochs=a.epochs, lr=a.lr, target="leapfrog")
    log("saved %s (%.1f s total)" % (a.out, time.time() - t0))
    from .evaluate import compare
    baseline = EdgeMLP(seed=a.seed)
    for es in (a.seed + 5000, a.seed + 5001):
        r = compare(model, es, nsteps=a.eval_steps, n=a.eval_n, warm=2000, baseline=baseline)
        log(json.dumps(r))


if __name__ == "__main__":
    main()
