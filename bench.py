"""Benchmark of the WCSPH particle step on one MI355X.

Default workload: the 3D dam break (81x81x162 fluid lattice = 1,062,882
fluid particles, dx = 1/135 m, in a 216x81x162-cell tank with 3 layers of
fixed boundary particles, h = 1.2 dx, RK2, dt = 0.25 h / c0, FP32 storage).
Prints ONE JSON line: pair-interactions/s (whole job), steps/s, the roofline
of the dominant kernel measured with HIP events on the launch stream, and
the CPU baseline (the FP64 oracle timed on a bounded sample on this host).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--scale", type=float, default=1.0,
                    help="shrink the lattice (for quick checks); 1.0 = the named workload")
    ap.add_argument("--pair-prec", choices=["f64", "native"], default="native")
    ap.add_argument("--profile-steps", type=int, default=10)
    ap.add_argument("--cells-per-radius", type=int, default=2,
                    help="cell size = 2h / k (neighbour sets are unchanged)")
    ap.add_argument("--no-packed", action="store_true",
                    help="use the SoA kernels instead of the packed fast path")
    ap.add_argument("--no-list", action="store_true",
                    help="every pass searches the cells instead of reading neighbour lists")
    ap.add_argument("--reuse", type=float, default=None, metavar="SKIN",
                    help="reuse stage-1 neighbour lists in stage 2 with this skin")
    return ap.parse_args()


def build_case(scale=1.0, precision="single", cells_per_radius=2, packed=True, nlist=True,
               reuse=None):
    import pysph_amd.base.api as base
    import pysph_amd.solver.api as solver
    from pysph_amd.solver import generators as gen
    col = tuple(max(4, int(round(c * scale))) for c in (81, 81, 162))
    tank = (max(col[0] + 4, int(round(216 * scale))), col[1], col[2])
    f, b, info = gen.dam_break_3d(dx=1.0 / 135, column=col, tank=tank, seed=0,
                                  precision=precision)
    particles = base.Particles([f, b], update_particles=False)
    lo, hi = info["domain"]
    particles.cell_manager.set_fixed_domain(lo, hi)
    particles.cell_manager.check_jumps = "deferred"
    s = solver.wcsph_solver(3, integrator="rk2", co=info["co"], ro=info["rho0"], gamma=7.0,
                            alpha=info["alpha"], beta=info["beta"], eta=info["eta"],
                            gravity=info["g"], xsph_eps=info["eps"])
    s.setup_integrator(particles)
    s.set_time_step(info["dt"])
    plan = s.integrator.plan
    if plan is not None:
        if cells_per_radius != 1:
            plan.set_cells_per_radius(cells_per_radius)
        plan.use_packed = plan.use_packed and packed
        plan.use_list = plan.use_list and plan.use_packed and nlist
        if reuse is not None:
            plan.list_reuse = True
            plan.list_skin = reuse
    return s, particles, f, b, info, col, tank


def pair_counts(particles, f, b):
    """Accepted pairs (self included) of the density pass (all arrays from
    all arrays) and of the momentum pass (fluid from fluid and boundary)."""
    import torch
    from pysph_amd import _abi
    cm = particles.cell_manager
    cm.update()
    counts = {}
    for dst in (f, b):
        for src in (f, b):
            n = dst.get_number_of_particles()
            c = torch.empty(n, dtype=torch.int64, device=dst.device)
            _abi.call("sph_neighbor_counts", torch.cuda.current_stream().cuda_stream, cm.grid,
                      dst.view(), src.view(), 2.0, c.data_ptr())
            counts[(dst.name, src.name)] = int(c.sum().item())
    p_rho = sum(counts.values())
    p_m = counts[("fluid", "fluid")] + counts[("fluid", "boundary")]
    p_x = counts[("fluid", "fluid")]
    return p_rho, p_m, p_x


def kernel_bytes(nf, nb):
    """Algorithmic (compulsory) bytes per launch, FP32: every field a
    launch reads is read once, every field it writes is written once."""
    return {
        # density + Tait on one array: x,y,z,h,m of its particles and of the
        # other array's particles (sources), rho,p,cs written
        "density_fluid": 20 * nf + 20 * nb + 12 * nf,
        "density_boundary": 20 * nb + 20 * nf + 12 * nb,
        # stage 1: fluid x..w,h,m,rho,p,cs (44) read, k1 (24) + X1 (24) written;
        # boundary x,y,z,u,v,w,h,m,rho,p,cs (44) read
        "stage1": 92 * nf + 44 * nb,
        # stage 2: X1 + h,m,rho,p,cs (44) + X0 (24) + k1 (24) read, X (24) written
        "stage2": 116 * nf + 44 * nb,
    }


def time_kernels(s, nsteps):
    """Per-launch durations (ms) of the plan's kernels over nsteps steps,
    with events recorded on the launch stream around each launch."""
    import torch
    plan = s.integrator.plan
    rec = {}
    orig_density, orig_stage = plan._density, plan._stage

    def timed(name, fn, *a, **k):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        fn(*a, **k)
        e1.record()
        rec.setdefault(name, []).append((e0, e1))

    def density(disp=None, override=None, **kw):
        st = "2" if disp is not None else "1"
        timed("density_s" + st, orig_density, disp=disp, override=override, **kw)

    def stage(stage, dt, b, override=None, disp_in=None, disp_out=None):
        timed("stage%d" % stage, orig_stage, stage, dt, b, override=override, disp_in=disp_in,
              disp_out=disp_out)

    plan._density, plan._stage = density, stage
    try:
        for _ in range(nsteps):
            s.integrator.integrate(s.dt)
        torch.cuda.synchronize()
    finally:
        plan._density, plan._stage = orig_density, orig_stage
    return {k: float(np.mean([a.elapsed_time(b) for a, b in v])) for k, v in rec.items()}


def cpu_baseline(info):
    """One RK2 step of the FP64 oracle (single thread, the reference's
    structure: hash-map cells, per-(src,dst) neighbour caches, one kernel
    call per pair) on a 1/8-volume dam break at the same dx."""
    from oracle import oracle as O
    from pysph_amd.solver import generators as gen
    col = (40, 40, 81)
    tank = (108, 40, 81)
    f, b, inf = gen.dam_break_3d(dx=1.0 / 135, column=col, tank=tank, seed=0,
                                 precision="single")
    from tests.util import to_oracle
    rf, rb = to_oracle(f), to_oracle(b)
    sch = O.scheme(3, integrator="rk2", dt=info["dt"], co=info["co"], ro=info["rho0"],
                   gamma=7.0, alpha=info["alpha"], beta=info["beta"], eta=info["eta"],
                   gravity=info["g"], xsph_eps=info["eps"], cell_size=2.0 * info["h"])
    # pair counts of the sample (for pair-interactions/s)
    cs = 2.0 * info["h"]
    npairs = 0
    for dst in (rf, rb):
        for src in (rf, rb):
            off, _ = O.nnps_csr(dst, src if src is not dst else None, cell_size=cs)
            npairs += int(off[-1])
    off, _ = O.nnps_csr(rf, None, cell_size=cs)
    off2, _ = O.nnps_csr(rf, rb, cell_size=cs)
    p_m = int(off[-1] + off2[-1])
    O.set_threads(1)
    t0 = time.perf_counter()
    O.scheme_step(sch, [rf, rb], 1)
    t = time.perf_counter() - t0
    pairs_per_step = 2 * (npairs + p_m)
    return {"value": pairs_per_step / t, "unit": "pair-interactions/s", "cores": 1,
            "kind": "port",
            "sample": "one RK2 step of a 40x40x81-lattice dam break (%d fluid + %d boundary "
                      "particles, same dx/h/c0/dt) with the FP64 oracle, single thread, "
                      "%.1f s" % (len(rf["x"]), len(rb["x"]), t),
            "seconds_per_step": t}


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: this build runs on exactly one GPU (the particle step of one "
              "rank's share); --gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)
    import torch
    from pysph_amd.sph import _ops
    from pysph_amd import _abi
    _ops.DEFAULT_PAIR_PREC = _abi.SPH_PAIR_F64 if args.pair_prec == "f64" else _abi.SPH_PAIR_NATIVE
    s, particles, f, b, info, col, tank = build_case(args.scale,
                                                     cells_per_radius=args.cells_per_radius,
                                                     packed=not args.no_packed,
                                                     nlist=not args.no_list, reuse=args.reuse)
    nf, nb = f.get_number_of_particles(), b.get_number_of_particles()
    assert s.integrator.plan is not None, "fused plan not selected"
    p_rho0, p_m0, _ = pair_counts(particles, f, b)
    for _ in range(args.warmup):
        s.integrator.integrate(s.dt)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(args.steps):
        s.integrator.integrate(s.dt)
    e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = e0.elapsed_time(e1)
    t = wall
    particles.cell_manager.verify_jumps()
    p_rho1, p_m1, _ = pair_counts(particles, f, b)
    p_rho, p_m = 0.5 * (p_rho0 + p_rho1), 0.5 * (p_m0 + p_m1)
    pairs_per_step = 2 * (p_rho + p_m)
    value = pairs_per_step * args.steps / t
    kt = time_kernels(s, args.profile_steps)
    kb = kernel_bytes(nf, nb)
    dur = {"density_fluid": None, "stage1": kt.get("stage1"), "stage2": kt.get("stage2")}
    # dominant kernel: the longer of the two stage kernels
    dom = max(("stage1", "stage2"), key=lambda k: kt.get(k, 0.0))
    dom_ms = kt[dom]
    achieved = kb[dom] / (dom_ms * 1e-3) / 1e9
    flops_dom = 80.0 * p_m + 0.0
    out = {
        "metric": "pair-interactions/s",
        "value": value,
        "unit": "pair-interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * t / args.steps,
        "steps_per_s": args.steps / t,
        "gpu_ms_per_step": gpu_ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "3D dam break, RK2, FP32 storage (pair arithmetic %s)"
                               % args.pair_prec,
                   "cells_per_radius": args.cells_per_radius,
                   "packed": not args.no_packed, "neighbour_list": not args.no_list,
                   "list_reuse_skin": args.reuse,
                   "fluid_lattice": list(col), "tank_cells": list(tank),
                   "n_fluid": nf, "n_boundary": nb, "dx": info["dx"], "h": info["h"],
                   "dt": info["dt"], "c0": info["co"], "pairs_per_step": pairs_per_step,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": "momentum_kernel (%s)" % dom,
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "bytes_per_launch": kb[dom], "avg_launch_ms": dom_ms},
        "valu": {"kernel": dom, "flops_per_launch": flops_dom,
                 "achieved_tflops": flops_dom / (dom_ms * 1e-3) / 1e12,
                 "peak_tflops": FP32_PEAK_TFLOPS,
                 "frac": flops_dom / (dom_ms * 1e-3) / 1e12 / FP32_PEAK_TFLOPS},
        "kernel_ms": kt,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(info)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
