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
    ap.add_argument("--workload", default="dam3d",
                    choices=["dam3d", "cube-uniform", "cube-clustered", "batch2d",
                             "surrogate2d", "c1"],
                    help="dam3d: the 3D dam break step (default); cube-*: neighbour search + "
                         "summation density on 2^24 particles in the unit cube; batch2d: 512 "
                         "2D scenes (drops and dam breaks) stepped together; c1: the FP64 2D "
                         "elliptical drop (5025 particles) through the default solver objects")
    ap.add_argument("--c1-integrator", choices=["rk2", "euler"], default="rk2",
                    help="c1: RK2 (stable over the 100 steps) or the specified Euler, whose "
                         "free surface trips the jump tolerance near step 35 at this CFL")
    ap.add_argument("--scenes", type=int, default=512)
    ap.add_argument("--surrogate-n", type=int, default=60000,
                    help="surrogate2d: fluid particles of the 2D dam break")
    ap.add_argument("--no-graph", action="store_true",
                    help="dam3d, batch2d: launch every step eagerly (no HIP graph replay)")
    ap.add_argument("--no-autotune", action="store_true", help="batch2d: default schedule")
    ap.add_argument("--cube-n", type=int, default=1 << 24)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--scale", type=float, default=1.0,
                    help="shrink the lattice (for quick checks); 1.0 = the named workload")
    ap.add_argument("--pair-prec", choices=["f64", "native"], default="native")
    ap.add_argument("--profile-steps", type=int, default=10)
    ap.add_argument("--cells-per-radius", type=int, default=None,
                    help="bins per reference cell of 2h (neighbour sets are unchanged); "
                         "default 2, with --dropin the solver's default")
    ap.add_argument("--dropin", action="store_true",
                    help="dam3d: the solver as a reference script drives it -- grid from "
                         "the particle bounds, per-binning jump check, the solver's "
                         "default binning, steps timed through Solver.solve()")
    ap.add_argument("--no-packed", action="store_true",
                    help="use the SoA kernels instead of the packed fast path")
    ap.add_argument("--list", action="store_true",
                    help="the density pass records neighbour lists for the momentum pass")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only copy 256 MiB device-to-device 3 times (calibrates the "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE counters) and exit")
    ap.add_argument("--reuse", type=float, default=None, metavar="SKIN",
                    help="reuse stage-1 neighbour lists in stage 2 with this skin")
    ap.add_argument("--lds-rows", action="store_true",
                    help="stage the density scan's stencil rows in LDS (opt-in kernel variant: "
                         "the same sums)")
    ap.add_argument("--cache-static", action="store_true",
                    help="keep the boundary-from-boundary density sums between calls "
                         "(same bits; those pairs are then evaluated once, not per stage)")
    return ap.parse_args()


def build_case(scale=1.0, precision="single", cells_per_radius=2, packed=True, nlist=False,
               reuse=None, cache_static=False, dropin=False):
    import pysph_amd.base.api as base
    import pysph_amd.solver.api as solver
    from pysph_amd.solver import generators as gen
    col = tuple(max(4, int(round(c * scale))) for c in (81, 81, 162))
    tank = (max(col[0] + 4, int(round(216 * scale))), col[1], col[2])
    f, b, info = gen.dam_break_3d(dx=1.0 / 135, column=col, tank=tank, seed=0,
                                  precision=precision)
    particles = base.Particles([f, b], update_particles=False)
    if not dropin:
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
        if cells_per_radius is not None:
            plan.set_cells_per_radius(cells_per_radius)
        plan.use_packed = plan.use_packed and packed
        if nlist or reuse is not None:
            plan.enable_lists(True, skin=reuse, reuse=reuse is not None)
        plan.cache_static = bool(cache_static)
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
    pair_counts.boundary_boundary = counts[("boundary", "boundary")]
    # density pairs of the fluid destinations (= p_m: same source arrays)
    return p_rho, p_m, p_m


def config_key(args):
    """Kernel configuration the traffic figures belong to."""
    return "cpr%d-%s-%s%s" % (args.cells_per_radius, "packed" if not args.no_packed else "soa",
                              "list" if args.list else "nolist",
                              "" if args.reuse is None else "-reuse%g" % args.reuse) + (
                                  "-cache" if getattr(args, "cache_static", False) else "")


def kernel_names(plan):
    """Kernel symbols behind each pass (as rocprofv3 lists them)."""
    if not plan.use_packed:
        n = {"density_fluid": "sph::density_kernel<float, float, true>",
             "stage1": "sph::momentum_kernel<float, float, 1>",
             "stage2": "sph::momentum_kernel<float, float, 2>"}
    else:
        d = plan.dim
        if plan.use_list:
            n = {"density_fluid": "sph::density_scan<float, true, true, %d>" % d,
                 "stage1": "sph::momentum_list<float, 1>",
                 "stage2": "sph::momentum_list<float, 2>"}
        else:
            n = {"density_fluid": "sph::density_scan<float, true, false, %d>" % d,
                 "stage1": "sph::momentum_scan<float, 1, %d>" % d,
                 "stage2": "sph::momentum_scan<float, 2, %d>" % d}
    # the boundary's density pass runs the density kernel without lists
    n["density_boundary"] = n["density_fluid"].replace("true, true", "true, false")
    # both density passes of a stage in one launch
    n["density"] = "sph::density_scan2<float, true, %d>" % plan.dim
    return n


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
        # the fluid's and the boundary's density passes in one launch: the
        # sources' x,y,z,h,m read once, rho,p,cs of both arrays written
        "density": 32 * nf + 32 * nb,
    }


class LaunchTimer:
    """Stream events recorded immediately around the density and stage
    entry points (on torch's current stream, the launch stream), keyed by
    pass and destination array."""

    def __init__(self, nf):
        import torch
        self.torch = torch
        self.nf = nf
        self.rec = {}
        self.cur = None

    def key(self, name, args):
        if name == "sph_density_multi":
            return "density"
        if name in ("sph_density", "sph_density_cached"):
            return "density_fluid" if args[3][0].n == self.nf else "density_boundary"
        if name == "sph_wcsph_stage":
            return "stage%d" % args[1]
        return None

    def before(self, name, args):
        k = self.key(name, args)
        if k is not None:
            e0 = self.torch.cuda.Event(enable_timing=True)
            e0.record()
            self.cur = (k, e0)

    def after(self, name, args):
        if self.cur is not None:
            e1 = self.torch.cuda.Event(enable_timing=True)
            e1.record()
            self.rec.setdefault(self.cur[0], []).append((self.cur[1], e1))
            self.cur = None

    def durations(self):
        return {k: [a.elapsed_time(b) for a, b in v] for k, v in self.rec.items()}


def time_kernels(s, nf, nsteps):
    """Launch durations (ms) of the plan's passes over nsteps steps."""
    import torch
    from pysph_amd import _abi
    timer = LaunchTimer(nf)
    plan = s.integrator.plan
    # launches are traced one by one: the steps run eagerly here
    use_graph = plan.use_graph
    plan.use_graph = False
    try:
        # a few untimed steps first, without synchronising, so that the
        # launch queue is ahead of the GPU and no event waits on the host
        for _ in range(3):
            s.integrator.integrate(s.dt)
        _abi.tracer = timer
        try:
            for _ in range(nsteps):
                s.integrator.integrate(s.dt)
            torch.cuda.synchronize()
        finally:
            _abi.tracer = None
    finally:
        plan.use_graph = use_graph
    return timer.durations()


def _ev_ms(fn, reps):
    """Average duration (ms) of `reps` back-to-back calls of fn between two
    HIP events on the current (launch) stream, after one untimed call."""
    import torch
    fn()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def isolated_launch_ms(s, reps=20):
    """Average launch duration (ms) of every hot pass, each launched `reps`
    times back to back on the launch stream after one full step: the
    density passes of the fluid and of the boundary (idempotent: rho, p, cs
    from positions), the stage-1 kernel (writes X1 and k1 from X) and the
    stage-2 kernel reading X1, with its final state written to scratch
    arrays instead of in place, so that the repeated launches read the same
    inputs."""
    import torch
    from pysph_amd import _abi
    from pysph_amd._device import ptr
    from pysph_amd.sph import _ops
    plan = s.integrator.plan
    s.integrator.integrate(s.dt)
    b = plan._buffers()
    ov = plan._override()
    out = {}
    if plan._fuse_density():
        # the step's launch: both arrays' density passes in one
        out["density"] = _ev_ms(lambda: plan._density(override=ov), reps)
    else:
        for c in plan.rho_calcs:
            pa = c.dest
            key = "density_fluid" if pa is plan.fluid else "density_boundary"

            def dens(c=c, pa=pa):
                _ops.density(plan.cm, plan.prm, pa, c.sources, pa.get_carray("rho"),
                             pa.get_carray("p"), pa.get_carray("cs"), override=ov)
            out[key] = _ev_ms(dens, reps)
    stream = torch.cuda.current_stream().cuda_stream
    _abi.call("sph_fill_i32", stream, b["disp"].data_ptr(), 1, 0)
    out["stage1"] = _ev_ms(lambda: plan._stage(1, s.dt, b, override=ov, disp_out=b["disp"]),
                           reps)
    s1 = plan._s1_fields(b)
    plan._pack(plan.fluid, fields=s1, force=True)
    ov2 = plan._override(s1)
    pa = plan.fluid
    scratch = [torch.empty_like(t) for t in plan._state0()]
    vd = _ops.views([pa], ov2)
    vs = _ops.views(plan.srcs, ov2)
    s0 = _abi.ptr_array([ptr(t) for t in plan._state0()])
    k1 = _abi.ptr_array([ptr(t) for t in b["k1"]])
    o2 = _abi.ptr_array([ptr(t) for t in scratch])
    rmax = _ops.search_radius(plan.cm, plan.radius_scale)

    def stage2():
        _abi.call("sph_wcsph_stage", stream, 2, plan.cm.grid, plan.prm, vd, len(plan.srcs), vs,
                  plan.flags, rmax, ptr(b["disp"]), float(s.dt), s0, k1, o2, None, None)
    out["stage2"] = _ev_ms(stage2, reps)
    return out


def source_digest():
    """SHA-256 of the kernel sources (keys the committed traffic figures)."""
    import hashlib
    d = os.path.join(os.path.dirname(os.path.abspath(__file__)), "pysph_amd", "csrc")
    h = hashlib.sha256()
    for fn in sorted(os.listdir(d)):
        if fn.endswith((".hip", ".h")):
            with open(os.path.join(d, fn), "rb") as fh:
                h.update(fn.encode() + b"\0" + fh.read())
    return h.hexdigest()


def measured_traffic(config_key):
    """HBM bytes per launch from the rocprofv3 FETCH_SIZE / WRITE_SIZE
    summary under profiles/, when it was taken on the current sources."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        "dam3d_hbm_traffic.md")
    try:
        with open(path) as fh:
            for line in fh:
                if line.startswith("traffic-json:"):
                    rec = json.loads(line[len("traffic-json:"):])
                    if rec.get("source_sha256") != source_digest():
                        return None, "stale (kernel sources changed since the counter runs)"
                    t = rec.get("configs", {}).get(config_key)
                    if t is None:
                        return None, "no counter run for this configuration"
                    return t, path
    except OSError:
        pass
    return None, "no counter summary"


def measured_counters(config_key):
    """VALU instructions per wave and waves per launch from the rocprofv3
    SQ counter summary under profiles/, when it was taken on the current
    sources."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        "dam3d_shader_counters.md")
    try:
        with open(path) as fh:
            for line in fh:
                if line.startswith("counters-json:"):
                    rec = json.loads(line[len("counters-json:"):])
                    if rec.get("source_sha256") != source_digest():
                        return None, "stale (kernel sources changed since the counter runs)"
                    t = rec.get("configs", {}).get(config_key)
                    if t is None:
                        return None, "no counter run for this configuration"
                    return t, path
    except OSError:
        pass
    return None, "no counter summary"


def valu_issue_busy(valu_per_wave, waves, launch_ms):
    """Fraction of the GPU's VALU issue cycles a launch's instructions fill:
    a wave64 VALU instruction occupies one of MI355X's 1024 SIMDs (256 CUs x
    4) for 4 cycles at 2.4 GHz."""
    return valu_per_wave * waves * 4.0 / (1024 * 2.4e9 * launch_ms * 1e-3)


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
    pairs_per_step = 2 * (npairs + p_m)
    out = {}
    for threads, steps in ((1, 6), (len(os.sched_getaffinity(0)), 6)):
        O.set_threads(threads)
        t0 = time.perf_counter()
        O.scheme_step(sch, [rf, rb], steps)
        t = (time.perf_counter() - t0) / steps
        out[threads] = {
            "value": pairs_per_step / t, "unit": "pair-interactions/s", "cores": threads,
            "kind": "port",
            "sample": "%d RK2 steps of a 40x40x81-lattice dam break (%d fluid + %d boundary "
                      "particles, same dx/h/c0/dt) with the FP64 oracle, %d thread(s), "
                      "%.2f s per step" % (steps, len(rf["x"]), len(rb["x"]), threads, t),
            "seconds_per_step": t}
    O.set_threads(1)
    base = out[1]
    base["all_cores"] = out[max(out)]
    return base


def cube_case(kind, n, cells_per_radius=2, seed=0):
    """Unit-cube set of the neighbour/density microbenchmark: FP32, m = 1/n,
    h for 50 neighbours on average, binned on the fixed domain [0,1]^3."""
    import pysph_amd.base.api as base
    import pysph_amd.sph.api as sph
    from pysph_amd.solver import generators as gen
    pa = gen.cube_array(n, kind=kind, seed=seed, precision="single")
    particles = base.Particles([pa], update_particles=False)
    cm = particles.cell_manager
    cm.set_fixed_domain((0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    cm.check_jumps = False
    if cells_per_radius != 1:
        cm.set_cells_per_radius(cells_per_radius)
    calc = sph.SPHCalc(particles, [pa], pa, base.CubicSplineKernel(dim=3),
                       [sph.SPHRho.withargs().get_func(pa, pa)], updates=["rho"])
    return pa, particles, calc


def cube_check(pa, particles, nsample=4096, seed=1):
    """Neighbour sets of nsample random destinations (queried against the
    whole set on the GPU) compared with the FP64 oracle, and the CPU
    baseline: the oracle's summation density over all particles of the
    sub-cube [0, 1/4)^3, single thread."""
    import pysph_amd.base.api as base
    from pysph_amd.base import nnps
    from oracle import oracle as O
    cm = particles.cell_manager
    cm.update()
    x, y, z, h = (pa.get(f).astype(np.float64) for f in ("x", "y", "z", "h"))
    rng = np.random.Generator(np.random.PCG64(seed))
    pick = np.sort(rng.choice(len(x), nsample, replace=False))
    probe = base.get_particle_array(cl_precision="single", name="probe", x=x[pick],
                                    y=y[pick], z=z[pick], h=h[pick])
    cm.bin_array(probe, cm.grid)
    off_g, nb_g = nnps.neighbor_csr(cm, pa, probe, 2.0)
    src_o = O.make_array(x=x, y=y, z=z, h=h, m=pa.get("m").astype(np.float64))
    dst_o = O.make_array(x=x[pick], y=y[pick], z=z[pick], h=h[pick],
                         m=np.zeros(nsample))
    off_o, nb_o = O.nnps_csr(dst_o, src_o, cell_size=2.0 * float(h.max()))
    exact = bool(np.array_equal(off_g, off_o)) and all(
        np.array_equal(np.sort(nb_g[off_g[i]:off_g[i + 1]]), np.sort(nb_o[off_o[i]:off_o[i + 1]]))
        for i in range(nsample))
    sel = (x < 0.25) & (y < 0.25) & (z < 0.25)
    sub = O.make_array(x=x[sel], y=y[sel], z=z[sel], h=h[sel],
                       m=pa.get("m").astype(np.float64)[sel])
    off_s, _ = O.nnps_csr(sub, None, cell_size=2.0 * float(h.max()))
    O.set_threads(1)
    # repeat the sample until about 10 s of CPU work
    reps, total = 0, 0.0
    while total < 10.0 and reps < 20:
        t0 = time.perf_counter()
        O.sphrho(3, sub, [sub], cell_size=2.0 * float(h.max()))
        total += time.perf_counter() - t0
        reps += 1
    t = total / reps
    return ({"destinations": nsample, "bit_exact": exact,
             "pairs_checked": int(off_o[-1])},
            {"value": float(off_s[-1]) / t, "unit": "pair-interactions/s", "cores": 1,
             "kind": "port",
             "sample": "summation density (cell build, neighbour search, sums) of the %d "
                       "particles in [0,1/4)^3 with the FP64 oracle, single thread, %.2f s per "
                       "evaluation, %d evaluations" % (int(sel.sum()), t, reps)})


def main_cube(args):
    import torch
    from pysph_amd import _abi
    kind = args.workload.split("-", 1)[1]
    pa, particles, calc = cube_case(kind, args.cube_n, args.cells_per_radius)
    n = pa.get_number_of_particles()
    cm = particles.cell_manager

    def step():
        pa.set_dirty(True)
        calc.sph("rho")
    step()
    torch.cuda.synchronize()
    cnt = torch.empty(n, dtype=torch.int64, device=pa.device)
    _abi.call("sph_neighbor_counts", torch.cuda.current_stream().cuda_stream, cm.grid,
              pa.view(), pa.view(), 2.0, cnt.data_ptr())
    pairs = int(cnt.sum().item())
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record()
    for _ in range(args.steps):
        step()
    e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    # the density launch alone, 20 back-to-back (idempotent: rho from positions)
    from pysph_amd.sph import _ops
    prm = _ops.make_params(3, radius_scale=2.0, lds_rows=args.lds_rows)
    ov = _ops.packed_override([pa])
    rho = pa.get_carray("rho")

    def launch():
        _ops.density(cm, prm, pa, [pa], rho, override=ov)
    launch()
    torch.cuda.synchronize()
    e2 = torch.cuda.Event(enable_timing=True)
    e3 = torch.cuda.Event(enable_timing=True)
    e2.record()
    for _ in range(20):
        launch()
    e3.record()
    torch.cuda.synchronize()
    dens_ms = e2.elapsed_time(e3) / 20
    # algorithmic bytes: density launch reads x,y,z,h,m and writes rho (24 N);
    # a step adds the binning (16 N: key and permutation written and read)
    dens_bytes = 24.0 * n
    achieved = dens_bytes / (dens_ms * 1e-3) / 1e9
    out = {
        "metric": "pair-interactions/s",
        "value": pairs * args.steps / wall,
        "unit": "pair-interactions/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "gpu_ms_per_step": e0.elapsed_time(e1) / args.steps,
        "particles_per_s": n * args.steps / wall,
        "step_hbm_gbs": 40.0 * n * args.steps / wall / 1e9,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "unit-cube neighbour search + summation density, %s" % kind,
                   "n": n, "h": float(pa.get_carray("h")[0].item()),
                   "pairs_per_step": pairs, "mean_neighbours": pairs / n,
                   "cells_per_radius": args.cells_per_radius, "lds_rows": args.lds_rows,
                   "step": "re-bin (fixed domain) + pack + density", "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": "sph::density_scan<float, false, false>",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "bytes_per_launch": dens_bytes, "avg_launch_ms": dens_ms,
                     "launches_timed": 20},
    }
    if not args.no_cpu:
        out["nbr_check"], out["cpu_baseline"] = cube_check(pa, particles)
    print(json.dumps(out), flush=True)


def batch_pairs(sc, cells_per_radius=2):
    """Pairs per RK2 step of a list of (fluid, boundary, info) scenes: the
    density pass of every array from all arrays and the momentum pass of the
    fluid, twice (both stages), counted on the GPU per scene."""
    import torch
    import pysph_amd.base.api as base
    from pysph_amd import _abi
    total = 0
    stream = torch.cuda.current_stream().cuda_stream
    for f, b, info in sc:
        arrays = [f] + ([b] if b is not None else [])
        cm = base.CellManager(arrays_to_bin=arrays)
        cm.set_fixed_domain(*info["domain"])
        cm.check_jumps = False
        cm.set_cells_per_radius(cells_per_radius)
        cm.update()
        p = 0
        for dst in arrays:
            for src in arrays:
                n = dst.get_number_of_particles()
                c = torch.empty(max(n, 1), dtype=torch.int64, device=dst.device)
                _abi.call("sph_neighbor_counts", stream, cm.grid, dst.view(), src.view(), 2.0,
                          c.data_ptr())
                k = int(c[:n].sum().item())
                p += k
                if dst is f:
                    p += k  # the momentum pass visits the same pairs
        total += 2 * p
    return total


def main_batch(args):
    import torch
    from pysph_amd.batch.engine import BatchedWCSPH, scene_from_arrays
    from pysph_amd.solver import generators as gen
    # the batched configuration is quoted over 200 steps
    if args.steps == 1000:
        args.steps = 200
    sc = gen.batch_scenes(args.scenes, seed=0)
    nsc = len(sc)
    pairs = batch_pairs(gen.batch_scenes(args.scenes, seed=0), args.cells_per_radius)
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        eng = BatchedWCSPH([scene_from_arrays(f, b, i) for f, b, i in sc],
                           cells_per_radius=args.cells_per_radius)
        tune = None if args.no_autotune else eng.autotune(reps=3)
        if not args.no_graph:
            eng.capture()
        for _ in range(args.warmup):
            eng.step()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        e0.record()
        for _ in range(args.steps):
            eng.step()
        e1.record()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        dens_ms = eng.time_phase("density", eng.sched["density"], reps=20)
        stage_ms = eng.time_phase("stage", eng.sched["stage"], reps=20)
    nf, nb = eng.nf, eng.nb
    # algorithmic bytes (2D FP32): density on fluid destinations reads x,y,h,m
    # of both sets and writes rho,p,cs of the fluid; stage 1 reads
    # x,y,u,v,h,m,rho,p,cs of both sets and writes k1 (4) and X1 (4) of the fluid
    kb = {"density": 28.0 * nf + 16.0 * nb, "stage": 68.0 * nf + 36.0 * nb}
    per_step = {"density": 2 * dens_ms, "stage": 2 * stage_ms}
    dom = max(per_step, key=per_step.get)
    dom_ms = dens_ms if dom == "density" else stage_ms
    achieved = kb[dom] / (dom_ms * 1e-3) / 1e9
    kinds = [i["kind"] for _, _, i in sc]
    out = {
        "metric": "pair-interactions/s",
        "value": pairs * args.steps / wall,
        "unit": "pair-interactions/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "gpu_ms_per_step": e0.elapsed_time(e1) / args.steps,
        "scene_steps_per_s": nsc * args.steps / wall,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "batched 2D scenes: %d elliptical drops + %d dam breaks, RK2"
                               % (kinds.count("drop"), kinds.count("dam")),
                   "scenes": nsc, "n_fluid": nf, "n_boundary": nb, "pairs_per_step": pairs,
                   "graph": not args.no_graph, "cells_per_radius": args.cells_per_radius,
                   "schedules": eng.sched, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm",
                     "kernel": "sph::batch_density_kernel" if dom == "density"
                     else "sph::batch_stage_kernel<1>",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "bytes_per_launch": kb[dom], "avg_launch_ms": dom_ms, "launches_timed": 20},
        "phase_ms": {"density_fluid": dens_ms, "stage1": stage_ms},
    }
    if tune is not None:
        # the plain grid (mode 0) against the best persistent schedules
        # (mode 1 static segments, mode 2 per-XCD work counters)
        out["autotune"] = {ph: {"best": r[0], "ms": r[1],
                                "default_ms": next(t for c, t in r[2] if c["mode"] == 0),
                                "best_persistent": min(((c, t) for c, t in r[2] if c["mode"]),
                                                       key=lambda ct: ct[1], default=None)}
                           for ph, r in tune.items()}
    eng.close()
    if not args.no_cpu:
        out["cpu_baseline"] = batch_cpu_baseline(sc)
    print(json.dumps(out), flush=True)


def batch_cpu_baseline(sc):
    """RK2 steps of the FP64 oracle over the batch's smallest scenes in turn
    (single thread) until about 10 s of CPU work."""
    from oracle import oracle as O
    from tests.util import to_oracle
    from pysph_amd.solver import generators as gen
    order = sorted(range(len(sc)), key=lambda k: sc[k][0].get_number_of_particles())
    fresh = gen.batch_scenes(len(sc), seed=0)
    O.set_threads(1)
    total_t, total_pairs, used = 0.0, 0, 0
    for k in order:
        f, b, info = fresh[k]
        arrs = [to_oracle(f, 0)] + ([to_oracle(b, 1)] if b is not None else [])
        cs = 2.0 * info["h"]
        p = 0
        for dst in arrs:
            for src in arrs:
                off, _ = O.nnps_csr(dst, src if src is not dst else None, cell_size=cs)
                p += int(off[-1]) * (2 if dst is arrs[0] else 1)
        sch = O.scheme(2, integrator="rk2", dt=info["dt"], co=info["co"], ro=info["rho0"],
                       gamma=7.0, alpha=info["alpha"], beta=info["beta"], eta=info["eta"],
                       gravity=info["g"], xsph_eps=info["eps"] or None, cell_size=cs)
        t0 = time.perf_counter()
        O.scheme_step(sch, arrs, 1)
        total_t += time.perf_counter() - t0
        total_pairs += 2 * p
        used += 1
        if total_t > 10.0:
            break
    return {"value": total_pairs / total_t, "unit": "pair-interactions/s", "cores": 1,
            "kind": "port",
            "sample": "one RK2 step of each of the %d smallest scenes with the FP64 oracle, "
                      "single thread, %.1f s" % (used, total_t)}


def main_c1(args):
    """The 2D elliptical drop of the reference's example (5025 particles on
    a dx = 0.025 lattice in the unit circle, h = 1.3 dx, m = dx^2, rho0 = 1,
    u = -100 x, v = 100 y, Tait c0 = 1400, gamma = 7, alpha = 0.01),
    FP64 storage, dt = 1e-5, through the default Particles / Solver objects
    (the drop-in path: bounds and the jump count read back without host
    waits, steps replayed from HIP graphs after the first ones).  steps/s
    over `steps` steps after `warmup` steps."""
    import torch
    import pysph_amd.base.api as base
    import pysph_amd.solver.api as solver
    from pysph_amd.solver import generators as gen
    from pysph_amd import _abi
    nsteps = args.steps if args.steps != 1000 else 100
    dt = 1e-5
    pa = gen.elliptical_drop(dx=0.025, precision="double")
    n = pa.get_number_of_particles()
    particles = base.Particles([pa])
    s = solver.wcsph_solver(2, integrator=args.c1_integrator, co=1400.0, ro=1.0, gamma=7.0,
                            alpha=0.01, beta=0.0, eta=0.1, with_solid=False)
    s.setup_integrator(particles)
    s.set_time_step(dt)
    assert s.integrator.plan is not None, "fused plan not selected"
    cm = particles.cell_manager
    cm.update()
    cnt = torch.empty(n, dtype=torch.int64, device=pa.device)
    _abi.call("sph_neighbor_counts", torch.cuda.current_stream().cuda_stream, cm.grid,
              pa.view(), pa.view(), 2.0, cnt.data_ptr())
    pairs0 = int(cnt.sum().item())
    for _ in range(args.warmup):
        s.integrator.integrate(dt)
    c0 = s.conservation()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record()
    for _ in range(nsteps):
        s.integrator.integrate(dt)
    e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    plan = s.integrator.plan
    graph_info = {"replays": plan.graph_replays, "captures": plan.graph_captures,
                  "error": plan.graph_error}
    c1 = s.conservation()
    cm.update()
    _abi.call("sph_neighbor_counts", torch.cuda.current_stream().cuda_stream, cm.grid,
              pa.view(), pa.view(), 2.0, cnt.data_ptr())
    pairs1 = int(cnt.sum().item())
    stages = 2 if args.c1_integrator == "rk2" else 1
    # density and momentum pass over the same pairs in every stage
    pairs_per_step = stages * 2 * 0.5 * (pairs0 + pairs1)
    kt = time_kernels(s, n, args.profile_steps)
    # algorithmic bytes per launch, FP64 2D, one array: density reads
    # x,y,h,m and writes rho,p,cs (56 B); stage 1 reads x,y,u,v,h,m,rho,p,cs
    # and writes k1 (a) and X1 (x,y,u,v) (120 B); stage 2 also reads X0 and
    # k1 and writes X (152 B); the Euler stage writes X (104 B)
    kb = {"density_fluid": 56.0 * n, "stage1": 120.0 * n, "stage2": 152.0 * n,
          "stage0": 104.0 * n}
    per = {k: float(np.mean(v)) for k, v in kt.items()}
    dom = max(per, key=lambda k: per[k] * len(kt[k]))
    achieved = kb.get(dom, 0.0) / (per[dom] * 1e-3) / 1e9
    names = {"density_fluid": "sph::density_kernel<double, double, true>",
             "stage1": "sph::momentum_kernel<double, double, 1>",
             "stage2": "sph::momentum_kernel<double, double, 2>",
             "stage0": "sph::momentum_kernel<double, double, 3>"}
    out = {
        "metric": "steps/s", "value": nsteps / wall, "unit": "steps/s",
        "n_gpus": 1, "steps": nsteps, "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / nsteps, "gpu_ms_per_step": e0.elapsed_time(e1) / nsteps,
        "pair_interactions_per_s": pairs_per_step * nsteps / wall,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "2D elliptical drop (C1), FP64, %s, dt = 1e-5, default "
                               "Particles/Solver objects" % args.c1_integrator.upper(),
                   "n": n, "h": 1.3 * 0.025, "c0": 1400.0, "alpha": 0.01,
                   "pairs_per_step": pairs_per_step, "hip_graph": graph_info,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": names.get(dom, dom), "pass": dom,
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "bytes_per_launch": kb.get(dom), "avg_launch_ms": per[dom],
                     "launches_timed": len(kt[dom])},
        "kernel_ms_events_in_loop": per,
        "conservation": conservation_report(c0, c1, nsteps),
    }
    if not args.no_cpu:
        out["cpu_baseline"] = c1_cpu_baseline(args.c1_integrator, nsteps)
    print(json.dumps(out), flush=True)


def c1_cpu_baseline(integrator, nsteps):
    """The same drop stepped by the FP64 oracle, single thread (the
    reference's structure) and on all host cores."""
    from oracle import oracle as O
    from tests.util import to_oracle
    from pysph_amd.solver import generators as gen
    res = {}
    for threads in (1, len(os.sched_getaffinity(0))):
        ref = to_oracle(gen.elliptical_drop(dx=0.025, precision="double"))
        sch = O.scheme(2, integrator=integrator, dt=1e-5, co=1400.0, ro=1.0, gamma=7.0,
                       alpha=0.01, beta=0.0, eta=0.1, cell_size=2.0 * 1.3 * 0.025)
        O.set_threads(threads)
        t0 = time.perf_counter()
        O.scheme_step(sch, [ref], nsteps)
        t = time.perf_counter() - t0
        res[threads] = {"value": nsteps / t, "unit": "steps/s", "cores": threads,
                        "kind": "port",
                        "sample": "%d %s steps of the same 5025-particle drop with the FP64 "
                                  "oracle, %d thread(s), %.2f s"
                                  % (nsteps, integrator.upper(), threads, t)}
    O.set_threads(1)
    base = res[1]
    base["all_cores"] = res[max(res)]
    return base


def conservation_report(c0, c1, nsteps):
    """FP64 sums (mass, momentum, kinetic, potential, Tait elastic, total
    energy) at the start and the end of a run, and the drift of the total
    energy relative to its start value, per 100 steps."""
    e0, e1 = float(c0[7]), float(c1[7])
    return {"mass": float(c1[0]), "mass_rel_change": float((c1[0] - c0[0]) / c0[0]),
            "momentum": [float(v) for v in c1[1:4]],
            "kinetic": float(c1[4]), "potential": float(c1[5]), "elastic": float(c1[6]),
            "total_start": e0, "total_end": e1, "steps": int(nsteps),
            "energy_drift_per_100_steps": (e1 - e0) / abs(e0) * 100.0 / max(nsteps, 1)
            if e0 != 0 else None}


SURROGATE_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "pysph_amd",
                                 "surrogate", "weights", "dam2d_edge_mlp.npz")


def main_surrogate(args):
    """The trained edge-MLP surrogate (steps of 4 dt, checked against the
    plain step every 10 steps) against the plain GPU solver on a 2D dam
    break after 1000 plain steps of settling."""
    import torch
    from pysph_amd.surrogate.evaluate import compare
    from pysph_amd.surrogate.graph import ParticleGraph, combined_array
    from pysph_amd.surrogate.model import EdgeMLP
    from pysph_amd.surrogate.train import dam_scene
    model = EdgeMLP.load(SURROGATE_WEIGHTS)
    untrained = EdgeMLP()
    steps = args.steps if args.steps != 1000 else 100
    r = compare(model, 7000, nsteps=steps, n=args.surrogate_n, warm=1000, baseline=untrained)
    # forward-pass launch timing on a graph of the same scene
    rng = np.random.Generator(np.random.PCG64(7000))
    f, b, info = dam_scene(rng, (args.surrogate_n, args.surrogate_n))
    g = ParticleGraph(combined_array(f, b), info, domain=info["domain"]).build()
    # the surrogate's stepping runs inference passes
    model.forward(g, infer=True)
    model.timing(1)
    for _ in range(10):
        model.forward(g, infer=True)
    ms, calls = model.timing(0)
    per = {k: v / calls for k, v in ms.items()}
    n, ne = g.n, g.nedges_exact()
    gemm_flops = 2.0 * n * (16 * 128 + 2 * (128 * 256 + 256 * 128))
    # edge aggregation, per round: P, Q rows (2 x 512 B) and the BF16
    # aggregate (256 B) per node, the 16-float feature row per edge
    edge_bytes = 2 * (n * (1024 + 256) + ne * 64)
    if per["gemm"] >= per["edges"]:
        roof = {"bound": "mfma", "kernel": "sph::mlp::gemm_nt_lds_kernel / gemm_nt_kernel "
                "(5 launches)",
                "achieved": gemm_flops / (per["gemm"] * 1e-3) / 1e12, "peak": 2500.0,
                "unit": "TFLOP/s", "flops_per_forward": gemm_flops, "avg_ms": per["gemm"]}
    else:
        # graphs under 2^21 nodes take the matrix-core aggregation
        kname = "edge_fwd_mfma_kernel" if n < (1 << 21) else "edge_fwd_kernel"
        roof = {"bound": "hbm", "kernel": "sph::mlp::%s (2 launches)" % kname,
                "achieved": edge_bytes / (per["edges"] * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "bytes_per_forward": edge_bytes, "avg_ms": per["edges"]}
    roof["frac"] = roof["achieved"] / roof["peak"]
    roof["traffic"] = None
    out = {
        "metric": "simulated-time/s",
        "value": r["surrogate_sim_time_per_s"],
        "unit": "s/s",
        "n_gpus": 1, "steps": steps, "warmup": 1000,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16 network (FP32 accumulate), f32 particles", "data": "synthetic",
        "config": {"workload": "2D dam break, edge-MLP surrogate at 4 dt vs plain RK2 at dt",
                   "n_fluid": r["n_fluid"], "n_boundary": r["n_boundary"], "dt": r["dt"],
                   "graph_nodes": n, "graph_edges": ne, "check_every": 10,
                   "density_threshold": 0.01, "parallelism": "single GPU"},
        "plain_sim_time_per_s": r["plain_sim_time_per_s"],
        "speedup_vs_plain": r["surrogate_sim_time_per_s"] / r["plain_sim_time_per_s"],
        "surrogate_density_error": r["surrogate_density_error"],
        "surrogate_position_error_dx": r["surrogate_position_error"],
        "fallbacks": r["fallbacks"], "checks": r["checks"],
        "untrained_density_error": r.get("untrained_density_error"),
        "plain_4dt_density_error": r["big_dt_density_error"],
        "plain_4dt_position_error_dx": r["big_dt_position_error"],
        "forward_ms": per,
        "roofline": roof,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = surrogate_cpu_baseline(info)
    print(json.dumps(out), flush=True)


def surrogate_cpu_baseline(info):
    """Plain RK2 steps of the FP64 oracle on the same 2D dam break, single
    thread, until about 10 s of CPU work (simulated-time/s)."""
    from oracle import oracle as O
    from tests.util import to_oracle
    from pysph_amd.surrogate.train import dam_scene
    rng = np.random.Generator(np.random.PCG64(7000))
    n = int(info.get("n_fluid", 60000))
    f, b, inf = dam_scene(rng, (n, n))
    arrs = [to_oracle(f, 0), to_oracle(b, 1)]
    sch = O.scheme(2, integrator="rk2", dt=inf["dt"], co=inf["co"], ro=inf["rho0"], gamma=7.0,
                   alpha=inf["alpha"], beta=inf["beta"], eta=inf["eta"], gravity=inf["g"],
                   xsph_eps=inf["eps"], cell_size=2.0 * inf["h"])
    O.set_threads(1)
    steps, total = 0, 0.0
    while total < 10.0 and steps < 50:
        t0 = time.perf_counter()
        O.scheme_step(sch, arrs, 1)
        total += time.perf_counter() - t0
        steps += 1
    return {"value": steps * inf["dt"] / total, "unit": "s/s", "cores": 1, "kind": "port",
            "sample": "%d RK2 steps of the same dam break (%d fluid + %d boundary particles) "
                      "with the FP64 oracle, single thread, %.1f s"
                      % (steps, f.get_number_of_particles(), b.get_number_of_particles(), total)}


def main():
    args = parse()
    if args.cells_per_radius is None and not args.dropin:
        args.cells_per_radius = 2
    if args.gpus != 1:
        print("bench.py: this build runs on exactly one GPU (the particle step of one "
              "rank's share); --gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)
    import torch
    if args.copy_probe:
        a = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": a.numel() * 4, "copies": 3}), flush=True)
        return
    if args.workload == "batch2d":
        return main_batch(args)
    if args.workload == "surrogate2d":
        return main_surrogate(args)
    if args.workload == "c1":
        return main_c1(args)
    if args.workload != "dam3d":
        return main_cube(args)
    from pysph_amd.sph import _ops
    from pysph_amd import _abi
    _ops.DEFAULT_PAIR_PREC = _abi.SPH_PAIR_F64 if args.pair_prec == "f64" else _abi.SPH_PAIR_NATIVE
    s, particles, f, b, info, col, tank = build_case(args.scale,
                                                     cells_per_radius=args.cells_per_radius,
                                                     packed=not args.no_packed,
                                                     nlist=args.list, reuse=args.reuse,
                                                     cache_static=args.cache_static,
                                                     dropin=args.dropin)
    if args.cells_per_radius is None:
        args.cells_per_radius = s.integrator.plan.cells_per_radius
    s.integrator.plan.use_graph = not args.no_graph
    if args.lds_rows:
        s.integrator.plan.set_lds_rows(True)
    nf, nb = f.get_number_of_particles(), b.get_number_of_particles()
    assert s.integrator.plan is not None, "fused plan not selected"
    p_rho0, p_m0, p_rho_fluid0 = pair_counts(particles, f, b)
    for _ in range(args.warmup):
        s.integrator.integrate(s.dt)
    # FP64 sums at the start of the timed steps (after the warm-up, whose
    # first density evaluation replaces the generator's rho = rho0)
    cons0 = s.conservation()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    if args.dropin:
        s.set_print_freq(0)
        s.set_final_time(s.t + (args.steps - 0.5) * s.dt)
        s.solve()
        assert s.count == args.steps, (s.count, args.steps)
    else:
        for _ in range(args.steps):
            s.integrator.integrate(s.dt)
    e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = e0.elapsed_time(e1)
    t = wall
    particles.cell_manager.verify_jumps()
    plan = s.integrator.plan
    graph_info = {"replays": plan.graph_replays, "captures": plan.graph_captures,
                  "error": plan.graph_error}
    cons1 = s.conservation()
    p_rho1, p_m1, p_rho_fluid1 = pair_counts(particles, f, b)
    p_rho, p_m = 0.5 * (p_rho0 + p_rho1), 0.5 * (p_m0 + p_m1)
    pairs_per_step = 2 * (p_rho + p_m)
    value = pairs_per_step * args.steps / t
    kt = time_kernels(s, nf, args.profile_steps)
    kb = kernel_bytes(nf, nb)
    plan = s.integrator.plan
    names = kernel_names(plan)
    iso = isolated_launch_ms(s, reps=20)
    cfg_key = config_key(args)
    traffic, traffic_src = measured_traffic(cfg_key)
    counters, counters_src = measured_counters(cfg_key)
    # algorithmic flops: 20 per density pair, 80 per momentum pair
    p_rho_f = 0.5 * (p_rho_fluid0 + p_rho_fluid1)
    p_rho_b = p_rho - p_rho_f
    flops = {"density_fluid": 20.0 * p_rho_f, "density_boundary": 20.0 * p_rho_b,
             "density": 20.0 * p_rho, "stage1": 80.0 * p_m, "stage2": 80.0 * p_m}
    # passes grouped by the kernel that runs them, with launches per step
    groups = {}
    for key, per_step in (("density", 2), ("density_fluid", 2), ("density_boundary", 2),
                          ("stage1", 1), ("stage2", 1)):
        if key not in iso:
            continue
        groups.setdefault(names[key], []).append((key, per_step))
    kernels = []
    for kname, passes in groups.items():
        launches = sum(n for _, n in passes)
        ms_step = sum(n * iso[k] for k, n in passes)
        bytes_step = sum(n * float(kb[k]) for k, n in passes)
        flops_step = sum(n * flops[k] for k, n in passes)
        ach = bytes_step / (ms_step * 1e-3) / 1e9
        tr = None
        if traffic and all(traffic.get(k) is not None for k, _ in passes):
            tr = sum(n * traffic[k] for k, n in passes) / launches
        kernels.append({
            "kernel": kname, "passes": [k for k, _ in passes], "launches_per_step": launches,
            "ms_per_step": ms_step, "avg_launch_ms": ms_step / launches,
            "bytes_per_launch": bytes_step / launches, "achieved": ach, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": ach / HBM_PEAK_GBS, "traffic": tr,
            "traffic_over_algorithmic": (tr / (bytes_step / launches)) if tr else None,
            "valu_tflops": flops_step / (ms_step * 1e-3) / 1e12,
            "valu_frac": flops_step / (ms_step * 1e-3) / 1e12 / FP32_PEAK_TFLOPS,
            "valu_issue_busy": (valu_issue_busy(counters[passes[0][0]]["valu_per_wave"],
                                                counters[passes[0][0]]["waves"],
                                                ms_step / launches)
                                if counters and passes[0][0] in counters else None),
            # the texture-data unit (vector memory loads returned to the
            # lanes): busy cycles per CU over the launch's cycles, from the
            # committed counter summary
            "td_busy": (counters[passes[0][0]].get("td_busy")
                        if counters and passes[0][0] in counters else None)})
    kernels.sort(key=lambda k: -k["ms_per_step"])
    dom = kernels[0]
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
                   "packed": not args.no_packed, "neighbour_list": args.list,
                   "list_reuse_skin": args.reuse,
                   "fluid_lattice": list(col), "tank_cells": list(tank),
                   "n_fluid": nf, "n_boundary": nb, "dx": info["dx"], "h": info["h"],
                   "dt": info["dt"], "c0": info["co"], "pairs_per_step": pairs_per_step,
                   "cache_static": args.cache_static, "lds_rows": args.lds_rows,
                   # with the static cache the boundary-from-boundary density
                   # pairs are evaluated once, not in every stage
                   "pairs_evaluated_per_step": pairs_per_step - (
                       2 * pair_counts.boundary_boundary if args.cache_static else 0),
                   "driver": ("Solver.solve (grid from particle bounds, per-binning jump "
                              "check)" if args.dropin else
                              "integrator steps (fixed domain, deferred jump check)"),
                   "hip_graph": graph_info,
                   "parallelism": "single GPU"},
        # the kernel with the most GPU time per step, all its launches of a
        # step: algorithmic bytes per launch over the average isolated launch
        # duration (HIP events around 20 back-to-back launches of each pass
        # on the launch stream)
        "roofline": {"bound": "hbm", "kernel": dom["kernel"], "passes": dom["passes"],
                     "achieved": dom["achieved"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": dom["frac"], "traffic": dom["traffic"],
                     "traffic_source": traffic_src,
                     "bytes_per_launch": dom["bytes_per_launch"],
                     "avg_launch_ms": dom["avg_launch_ms"],
                     # what bounds these kernels: the vector memory data
                     # path (TD busy, from the committed counter summary)
                     # with VALU issue close behind (its instruction count
                     # x 4 cycles over this run's launch time)
                     "td_busy": dom["td_busy"],
                     "valu_issue_busy": dom["valu_issue_busy"],
                     "valu_issue_source": counters_src,
                     "launches_per_step": dom["launches_per_step"], "launches_timed": 20},
        "roofline_kernels": kernels,
        "valu": {"kernel": dom["kernel"], "achieved_tflops": dom["valu_tflops"],
                 "peak_tflops": FP32_PEAK_TFLOPS, "frac": dom["valu_frac"],
                 # the kernels are built without packed FP32 instructions:
                 # one FP32 FMA per lane and cycle is half the published peak
                 "peak_tflops_unpacked": FP32_PEAK_TFLOPS / 2,
                 "frac_unpacked": dom["valu_frac"] * 2},
        "isolated_launch_ms": iso,
        # per-launch events inside the step loop (each event pair adds its
        # own cost to the launch it brackets)
        "kernel_ms_events_in_loop": {k: float(np.mean(v)) for k, v in kt.items()},
        "pass_ms_per_step_in_loop": {k: float(np.sum(v)) / args.profile_steps
                                     for k, v in kt.items()},
        "conservation": conservation_report(cons0, cons1, args.steps),
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(info)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
