"""Benchmark of the 2D sparc solver on one MI355X.

Workload (the configuration the headline metric is quoted on): production-size
2D forward + adjoint + kernel, FP64, one source, plain method -- nx=256 by
nz=300 on the adiabatic polytrope with 12-point vertical PMLs, steady
supergranule flow (peak 240 m/s), 4000 time steps forward and 4000 adjoint,
the windowed cross-correlation travel-time misfit of vz over receivers 5-60 Mm
from the source (30-minute window, reference traces from the same source
without flow, computed before the timed region), kernels for psi and c.  One
bench "step" = one full gradient (forward + misfit + adjoint + kernel
accumulation) with all inputs resident in HBM.

    python bench.py [--steps K] [--warmup W] [--gpus 1]

Prints one JSON line: value = Mpoint-steps/s (nx*nz*time-steps*(forward +
adjoint)*sources / wall-s / 1e6), plus the batched 16-source gradient, the
roofline of the dominant kernel (HIP events) and two CPU baselines on the host
(the C oracle on one core, and min(16, usable cores) oracle instances stepping
concurrently; bounded samples).  other_configurations holds
  * the plain forward at 128x150 (500 steps, Gaussian force source, no flow),
    with the oracle's whole run on one core beside it;
  * the batched 16-source forward at 128x150 (flow and sound-speed bump,
    Gaussian force sources, 4000 steps) beside 16 sequential single-source runs;
  * the 8-source autograd gradient at 256x300 (travel-time misfit, 4000 steps,
    checkpoints every 100 steps), with the reference's adjoint recipe and
    with the exact discrete adjoint;
  * the 3D (option 16) forward and checkpointed gradient at 256x256x300
    (3000 steps, checkpoints every 150 steps).
`--only-3d` prints just the two 3D lines.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "Mpoint-steps/s FP64 + s per fwd+adjoint gradient, batched vs plain, 1 GPU"
HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "colseg_traffic_256x300.md")


def measured_traffic(kernel: str, ms: int, nseg: int, tc: int, nsrc: int):
    """HBM bytes per launch of the column pass from the committed rocprofv3
    --pmc FETCH_SIZE / WRITE_SIZE summary, when one of its lines was taken
    with this kernel, batch size and launch shape (`traffic: <kernel><MS,0>
    S=.. TC=.. P=.. fetch_kib=.. write_kib=.. fetch_scale=.. write_scale=..`);
    None otherwise."""
    import re
    if not os.path.exists(TRAFFIC_FILE):
        return None
    for m in re.finditer(r"^traffic: (k_col\w+)<(\d+),0> S=(\d+) TC=(\d+) P=(\d+) fetch_kib=([\d.]+) "
                         r"write_kib=([\d.]+) fetch_scale=([\d.]+) write_scale=([\d.]+)",
                         open(TRAFFIC_FILE).read(), re.M):
        if (m.group(1), int(m.group(2)), int(m.group(3)), int(m.group(4)), int(m.group(5))) == \
                (kernel, ms, nsrc, tc, nseg):
            return (float(m.group(6)) * float(m.group(8)) + float(m.group(7)) * float(m.group(9))) * 1024.0
    return None


def make_config(nx=256, nz=300, npml=12):
    """The polytrope at nx x nz (200 Mm wide), the stable time step, 12-point
    vertical PMLs, vz at +0.2 Mm every 30 s, and the supergranule stream
    function (peak 240 m/s)."""
    from sparc_flows_amd import background as B
    from sparc_flows_amd.params import SparcParams

    bg = B.polytrope(nz)
    dt = B.stable_timestep(bg)
    p = SparcParams(nx=nx, nz=nz, xlength=200e8, timestep=dt, npmlbot=npml, npmltop=npml, npmlhor=9)
    psi = B.supergranule_psi(nx, p.xlength, bg)
    return p, bg, psi


def forward_rate(nx, nz, nsrc, nsteps, flow, dc, reps=1, sequential=False, cpu_full_run=False):
    """Forward-only Mpoint-steps/s of a batch with the Gaussian force source
    (sigma 0.5 Mm at z = -0.15 Mm, derivative of a Gaussian at 3 mHz): the
    plain-forward (128x150, one source at x = 0, no flow, 500 steps) and the
    batched multi-source (128x150, 16 sources from source_positions(16, 0),
    flow and sound-speed bump, 4000 steps) configurations.  `sequential` adds
    the same sources run one at a time on a single-source handle;
    `cpu_full_run` times the oracle over the whole run on one host core."""
    import torch
    from sparc_flows_amd import background as B
    from sparc_flows_amd.solver import Sparc
    from sparc_flows_amd.sources import gaussian_force

    p, bg, psi = make_config(nx, nz)
    c2d = B.sound_speed_bump(nx, p.xlength, bg) if dc else None
    psi = psi if flow else None
    xs = list(B.source_positions(nsrc, seed=0)) if nsrc > 1 else [0.0]
    maxtime = nsteps - 1
    vr, zp, ser = gaussian_force(p, bg, xs, maxtime)
    s = Sparc(p, bg, nsrc=nsrc, c2d=c2d, psi=psi)
    s.forward_profile(vr, zp, ser, maxtime)            # warm-up (and autotune)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        vz, _ = s.forward_profile(vr, zp, ser, maxtime)
    torch.cuda.synchronize()
    sec = (time.perf_counter() - t0) / reps
    out = {"Mpoint_steps_per_s": nx * nz * nsteps * nsrc / sec / 1e6, "s_per_run": sec,
           "grid": [nx, nz], "sources": nsrc, "time_steps": nsteps, "flow": flow, "dc": dc,
           "source": "Gaussian force, sigma 0.5 Mm, z -0.15 Mm, derivative of a Gaussian at 3 mHz",
           "finite": bool(torch.isfinite(vz).all()) and float(vz.abs().max()) > 0,
           "column_pass_shape": {k: s.query(k) for k in ("colseg_tc", "colseg_ms", "colseg_p")}}
    if sequential:
        s1 = Sparc(p, bg, nsrc=1, c2d=c2d, psi=psi)
        s1.forward_profile(vr[:1], zp, ser, maxtime)   # warm-up (and autotune)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(nsrc):
            s1.forward_profile(vr[i:i + 1], zp, ser, maxtime)
        torch.cuda.synchronize()
        seq = time.perf_counter() - t0
        out["sequential_single_source_s"] = seq
        out["speedup_vs_sequential"] = seq / sec
    if cpu_full_run:
        import oracle
        oracle.build()
        o = oracle.Oracle(p, bg, c2d=c2d, psi=psi)
        o.set_forward_profile(vr[0], zp, ser, maxtime)
        t0 = time.perf_counter()
        o.run(maxtime)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": nx * nz * nsteps / dt / 1e6, "unit": "Mpoint-steps/s", "cores": 1,
                               "kind": "port", "s_per_run": dt,
                               "sample": f"the whole {nsteps}-step run of source 0, C oracle, one host core"}
    return out


def forward3d_rate(nsteps=20, warm=2, nx=256, ny=256, nz=300, cpu_seconds=8.0):
    """3D forward (option 16) at 256x256x300 with the axisymmetric supergranule
    flow and the radially symmetric sound-speed bump: Mpoint-steps/s, the
    per-pass launch times (HIP events on the solver's stream) and the roofline
    of the dominant pass.  Algorithmic bytes per point of each pass: what it
    must read and write once (the Thomas sweep's scratch round trip excluded)."""
    import torch
    from sparc_flows_amd import background as B
    from sparc_flows_amd.solver3d import Params3D, Sparc3D

    bg = B.polytrope(nz)
    p = Params3D(nx=nx, ny=ny, nz=nz, timestep=B.stable_timestep(bg), npmlbot=12, npmltop=12)
    c3d = B.sound_speed_bump_3d(nx, ny, p.xlength, p.ylength, bg)
    s = Sparc3D(p, bg, c3d=c3d)
    s.set_flow(*B.supergranule_flow_3d(nx, ny, p.xlength, p.ylength, bg))
    s.set_forward_source(0.0, 0.0, warm + nsteps)
    s.reset()
    for t in range(warm):
        s.step(t)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for t in range(warm, warm + nsteps):
        s.step(t)
    torch.cuda.synchronize()
    sec = (time.perf_counter() - t0) / nsteps
    ms = s.profile(2)
    # launches per step: 5 x 3 ddx, 5 x 3 ddy, 5 Zp, 5 Zv, a filter every other step
    per_step = {"ddx": 15 * ms["ddx"], "ddy": 15 * ms["ddy"], "z_p": 5 * ms["z_p"], "z_v": 5 * ms["z_v"],
                "filter": 0.5 * ms["filter"]}
    # Zv: v_x, v_y, v_z, p, d v_i/dx, d v_i/dy, v0 (3), d p/dx, d p/dy, rho' g/rho0
    # and the six state fields read, the six next-stage fields written: 28
    # doubles; Zp: xi_z, d xi_x/dx, d xi_y/dy, c2 rho0, d p0/dz, d rho0/dz read,
    # p and rho' g/rho0 written: 8
    bpp = {"z_v": 28 * 8, "z_p": 8 * 8, "ddx": 16, "ddy": 16, "filter": 6 * 16}
    dom = max(per_step, key=per_step.get)
    npts = nx * ny * nz
    achieved = bpp[dom] * npts / (ms[dom] * 1e-3) / 1e9
    cpu = cpu_baseline_3d(p, bg) if cpu_seconds > 0 else None
    return {"Mpoint_steps_per_s": npts / sec / 1e6, "ms_per_step": sec * 1e3, "grid": [nx, ny, nz],
            "cpu_baseline": cpu,
            "time_steps_timed": nsteps, "flow": "axisymmetric supergranule 240 m/s",
            "dc": "radial bump 5%", "ms_per_launch": ms, "ms_per_step_by_pass": per_step,
            "roofline": {"kernel": dom, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                         "algorithmic_bytes_per_point": bpp[dom]}}


def cpu_baseline_3d(p, bg, seconds=8.0, nxy=64):
    """The 3D oracle (numpy, one process) stepping a 64x64 column subset of the
    3D workload's grid (same nz, PML and time step) for about `seconds`."""
    import dataclasses
    from oracle.oracle3d import Oracle3D, Params3D as OP

    q = dataclasses.replace(p, nx=nxy, ny=nxy, xlength=p.xlength * nxy / p.nx, ylength=p.ylength * nxy / p.ny)
    o = Oracle3D(OP(**dataclasses.asdict(q)), bg)
    o.set_forward_source(0.0, 0.0, 100000)
    n = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        o.step(n)
        n += 1
    dt = time.perf_counter() - t0
    return {"value": nxy * nxy * p.nz * n / dt / 1e6, "unit": "Mpoint-steps/s", "cores": 1, "kind": "port",
            "sample": f"{n} forward time steps of a {nxy}x{nxy}x{p.nz} grid (same dz, PML and time step), "
                      f"{dt:.1f} s, numpy oracle"}


def gradient3d_rate(maxtime=449, interval="auto", nx=256, ny=256, nz=300):
    """3D forward + adjoint + kernels of one source at 256x256x300 with
    in-HBM checkpoint-and-recompute (a forward run, a second run of all but
    the last `interval` steps, one adjoint run, one kernel accumulation per
    output pair); "auto" sizes the interval to half the free HBM."""
    import torch
    from sparc_flows_amd import background as B
    from sparc_flows_amd.solver3d import Gradient3D, Params3D

    bg = B.polytrope(nz)
    dt = B.stable_timestep(bg)
    p = Params3D(nx=nx, ny=ny, nz=nz, timestep=dt, npmlbot=12, npmltop=12)
    p.outputcad = p.cadforcing = 30.0
    c3d = B.sound_speed_bump_3d(nx, ny, p.xlength, p.ylength, bg)
    flow = B.supergranule_flow_3d(nx, ny, p.xlength, p.ylength, bg)
    g = Gradient3D(p, bg, c3d=c3d, flow=flow, interval=interval)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    K = g.run(0.0, 0.0, maxtime)
    torch.cuda.synchronize()
    sec = time.perf_counter() - t0
    finite = bool(all(torch.isfinite(K[k]).all() for k in ("kc", "kvx", "kvy", "kvz")))
    return {"s_per_gradient": sec, "time_steps": maxtime + 1, "checkpoint_interval": K["interval"],
            "grid": [nx, ny, nz], "output_pairs": maxtime // p.steps + 1, "finite": finite,
            "Mpoint_steps_per_s_fwd_adj": nx * ny * nz * (maxtime + 1) * 2 / sec / 1e6}


def host_cpu():
    """nproc, the cores this process may run on, and the CPU model."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return {"nproc": os.cpu_count(), "usable_cores": usable, "cpu_model": model}


def cpu_baseline(p, bg, psi, seconds=12.0, instances=1):
    """Time the C oracle on a bounded sample of the same forward run:
    `instances` independent oracles (one per source, like one `sparc` process
    per source) step concurrently on their own threads -- the C calls release
    the GIL -- from a common start for about `seconds`; the rate is all their
    point-steps over the wall time until the last one stops."""
    import threading

    import oracle
    from sparc_flows_amd import background as B

    oracle.build()
    xs = list(B.source_positions(instances, seed=0)) if instances > 1 else [0.0]
    objs = []
    for x in xs:
        o = oracle.Oracle(p, bg, psi=psi)
        o.set_forward_source(float(x), 100000)
        objs.append(o)
    steps = [0] * instances
    ends = [0.0] * instances
    gate = threading.Barrier(instances + 1)
    t_start = [0.0]

    def work(i):
        gate.wait()
        o, n, stop = objs[i], 0, t_start[0] + seconds
        while time.perf_counter() < stop:
            o.step(n)
            n += 1
        steps[i], ends[i] = n, time.perf_counter()

    th = [threading.Thread(target=work, args=(i,)) for i in range(instances)]
    for t in th:
        t.start()
    t_start[0] = time.perf_counter()
    gate.wait()
    for t in th:
        t.join()
    wall = max(ends) - t_start[0]
    total = sum(steps)
    out = {"value": p.nx * p.nz * total / wall / 1e6, "unit": "Mpoint-steps/s", "cores": instances,
           "kind": "port",
           "sample": (f"{total} forward time steps of the {p.nx}x{p.nz} workload over {instances} concurrent "
                      f"oracle instance(s) (one source each), {wall:.1f} s wall on {instances} host core(s)")}
    if instances > 1:
        out.update(host_cpu())
        out["steps_per_instance"] = steps
    return out


def autograd_rate(nsrc=8, nsteps=4000, interval=100, adjoint="sparc"):
    """Seconds per autograd gradient of 8 sources at 256x300 (flow and
    sound-speed bump): DifferentiableSolver with a full-state checkpoint every
    `interval` steps, the travel-time misfit against reference traces of the
    model without flow (computed before the timed region), L.backward()
    giving dL/dpsi and dL/dc on the grid; `adjoint` "sparc" (the reference's
    recipe) or "discrete" (the exact transpose of the solver).  The timed gradient is the second
    one (the first fixes the autotuned shapes and the allocator's pools)."""
    import torch
    from sparc_flows_amd import background as B
    from sparc_flows_amd.autograd import DifferentiableSolver, travel_time_misfit, window_weights
    from sparc_flows_amd.gradient import receiver_mask

    p, bg, psi = make_config()
    c2d = B.sound_speed_bump(p.nx, p.xlength, bg)
    xs = list(B.source_positions(nsrc, seed=0))
    maxtime = nsteps - 1
    ds = DifferentiableSolver(p, bg, xs, maxtime, checkpoint_interval=interval, adjoint=adjoint)
    mask = torch.as_tensor(receiver_mask(p, xs), device="cuda")
    cad = p.steps * p.timestep
    w = window_weights(p.nt_out(maxtime), cad, 1800.0, device="cuda")
    tau_obs = torch.zeros(nsrc, p.nx, dtype=torch.float64, device="cuda")
    psi_t, c_t = torch.as_tensor(psi), torch.as_tensor(c2d)
    with torch.no_grad():
        vref = ds.solve(torch.zeros_like(psi_t), c_t).clone()

    def gradient():
        ps = psi_t.clone().requires_grad_(True)
        cc = c_t.clone().requires_grad_(True)
        vz = ds.solve(ps, cc)
        L = travel_time_misfit(vz, vref, tau_obs, mask, w, cad)
        L.backward()
        return float(L.detach()), ps.grad, cc.grad

    gradient()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    L, gp, gc = gradient()
    torch.cuda.synchronize()
    sec = time.perf_counter() - t0
    finite = bool(np.isfinite(L) and torch.isfinite(gp).all() and torch.isfinite(gc).all()
                  and float(gp.abs().max()) > 0 and float(gc.abs().max()) > 0)
    return {"s_per_gradient": sec, "sources": nsrc, "grid": [p.nx, p.nz], "time_steps": nsteps,
            "checkpoint_interval": interval, "misfit": L, "finite": finite, "adjoint": adjoint,
            "Mpoint_steps_per_s_fwd_adj": p.nx * p.nz * nsteps * 2 * nsrc / sec / 1e6,
            "runs": ("forward with checkpoints, recompute of all but the last segment, adjoint"
                     if adjoint == "sparc" else
                     "forward, trajectory recompute from checkpoints, transposed step (exact discrete adjoint)")}


def reference_traces(p, bg, xs, maxtime):
    """vz of the same sources in the model without flow: the travel-time
    misfit's reference traces, [nsrc][nt][nx] in HBM."""
    from sparc_flows_amd.solver import Sparc
    s = Sparc(p, bg, nsrc=len(xs))
    vz, _ = s.forward(xs, maxtime)
    return vz.clone()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3, help="timed gradients")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nsteps", type=int, default=4000, help="time steps per run")
    ap.add_argument("--batch", type=int, default=16, help="sources of the batched gradient (0: skip)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the other_configurations lines (plain and batched forward at 128x150, the "
                         "8-source autograd gradient, the 3D forward and gradient)")
    ap.add_argument("--only-3d", action="store_true",
                    help="only run the 3D forward and gradient lines (256x256x300) and print them")
    ap.add_argument("--gradient3d-steps", type=int, default=3000,
                    help="time steps of the 3D checkpointed gradient line")
    ap.add_argument("--gradient3d-interval", type=int, default=150,
                    help="checkpoint interval of the 3D gradient line")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run a plain device-to-device copy of this many MiB ten times "
                         "(calibrates the FETCH_SIZE / WRITE_SIZE counters) and exit")
    ap.add_argument("--forward-grid", default="",
                    help="only time a one-source forward run (flow and sound-speed bump, --nsteps steps) "
                         "on this NXxNZ grid, e.g. 300x300 (an nx that is not a power of two runs the "
                         "circulant row pass), print it and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import torch
    from sparc_flows_amd.gradient import GradientPipeline

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)

    def log(msg):
        print(f"[bench] {msg}", file=sys.stderr, flush=True)

    if args.copy_probe_mib > 0:
        n = args.copy_probe_mib * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        for _ in range(10):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": n * 8, "launches": 10}))
        return
    if args.forward_grid:
        nx, nz = (int(v) for v in args.forward_grid.lower().split("x"))
        print(json.dumps({f"forward_{nx}x{nz}": forward_rate(nx, nz, 1, args.nsteps, True, True)}))
        return
    if args.only_3d:
        f3 = forward3d_rate()
        log("3D forward done")
        print(json.dumps({"forward3d_256x256x300": f3,
                          "gradient3d_256x256x300": gradient3d_rate(maxtime=args.gradient3d_steps - 1,
                                                                    interval=args.gradient3d_interval)}))
        return
    p, bg, psi = make_config()
    maxtime = args.nsteps - 1
    pipe = GradientPipeline(p, bg, nsrc=1, psi=psi, misfit="travel_time")
    xsrc = [0.0]
    vref = reference_traces(p, bg, xsrc, maxtime)
    log("reference traces done")

    def one():
        return pipe.run(xsrc, maxtime, vz_obs=vref)

    for i in range(args.warmup):
        one()
        torch.cuda.synchronize()
        log(f"warmup gradient {i + 1}/{args.warmup} done")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        out = one()
        torch.cuda.synchronize()
        log(f"timed gradient {i + 1}/{args.steps} done ({time.perf_counter() - t0:.2f} s)")
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    sec = (t1 - t0) / args.steps
    pts = p.nx * p.nz * (maxtime + 1) * 2
    value = pts / sec / 1e6
    finite = bool(torch.isfinite(out["kc"]).all() and torch.isfinite(out["kpsi"]).all()
                  and torch.isfinite(out["misfit"]).all())

    npts = p.nx * p.nz
    C = 2  # x-varying coefficient arrays read per point (v0x, v0z)
    bytes_per_pt = {"col": (14 + C) * 8, "col_last": (14 + C) * 8, "row": 9 * 8, "row_first": 7 * 8,
                    "row_last": 4 * 8}

    # batched gradient of args.batch sources vs the plain one
    batched = None
    if args.batch > 0:
        from sparc_flows_amd import background as B
        xs = list(B.source_positions(args.batch, seed=0))
        bref = reference_traces(p, bg, xs, maxtime)
        bpipe = GradientPipeline(p, bg, nsrc=args.batch, psi=psi, misfit="travel_time")
        bpipe.run(xs, maxtime, vz_obs=bref)
        torch.cuda.synchronize()
        log(f"batched warmup ({args.batch} sources) done")
        tb0 = time.perf_counter()
        bout = bpipe.run(xs, maxtime, vz_obs=bref)
        torch.cuda.synchronize()
        bsec = time.perf_counter() - tb0
        log(f"batched gradient {bsec:.2f} s")
        bfinite = bool(all(torch.isfinite(bout[k]).all() and float(bout[k].abs().max()) > 0
                           for k in ("kc", "kpsi")) and torch.isfinite(bout["misfit"]).all())
        # the batch's column pass against HBM: every launch moves args.batch
        # times the single-source bytes, more than the 256 MB MALL holds
        bms = bpipe.fwd.profile_kernels(200)
        bach = bytes_per_pt["col"] * npts * args.batch / (bms["col"] * 1e-3) / 1e9
        btraffic = measured_traffic("k_colsplit", bpipe.fwd.query("colseg_ms"), bpipe.fwd.query("colseg_p"),
                                    bpipe.fwd.query("colseg_tc"), args.batch)
        batched = {"sources": args.batch, "s_per_gradient": bsec, "finite": bfinite,
                   "Mpoint_steps_per_s": pts * args.batch / bsec / 1e6,
                   "speedup_vs_sequential_plain": (sec * args.batch) / bsec,
                   "roofline_col": {"achieved": bach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                    "frac": bach / HBM_PEAK_GBS, "ms_per_launch": bms, "traffic": btraffic,
                                    "algorithmic_bytes_per_launch": bytes_per_pt["col"] * npts * args.batch},
                   "shape": {k: bpipe.fwd.query(k) for k in ("colseg_tc", "colseg_ms", "colseg_p", "row_wide")}}
        del bpipe, bref

    # roofline of the dominant kernel: HIP events on the solver's stream
    ms = pipe.fwd.profile_kernels(200)
    per_step_ms = {"col": 4 * ms["col"], "col_last": ms["col_last"], "row": 4 * ms["row"],
                   "row_first": ms["row_first"], "row_last": ms["row_last"]}
    dom = max(per_step_ms, key=per_step_ms.get)
    achieved = bytes_per_pt[dom] * npts / (ms[dom] * 1e-3) / 1e9
    traffic = (measured_traffic("k_colsplit", pipe.fwd.query("colseg_ms"), pipe.fwd.query("colseg_p"),
                                pipe.fwd.query("colseg_tc"), 1) if dom == "col" else None)
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": dom,
                "ms_per_launch": ms, "algorithmic_bytes_per_point": bytes_per_pt[dom]}
    log("kernel profile done")

    others = None
    if not args.no_configs:
        others = {}
        others["plain_forward_128x150"] = forward_rate(128, 150, 1, 500, False, False, reps=3,
                                                       cpu_full_run=not args.no_cpu)
        log("plain forward 128x150 done")
        others["batched16_forward_128x150"] = forward_rate(128, 150, 16, 4000, True, True, sequential=True)
        log("batched forward 128x150 done")
        others["autograd8_gradient_256x300"] = autograd_rate(8, 4000, 100)
        log("autograd gradient done")
        others["autograd8_discrete_gradient_256x300"] = autograd_rate(8, 4000, 100, adjoint="discrete")
        log("discrete-adjoint gradient done")
        others["forward3d_256x256x300"] = forward3d_rate(cpu_seconds=0.0 if args.no_cpu else 8.0)
        log("3D forward done")
        others["gradient3d_256x256x300"] = gradient3d_rate(maxtime=args.gradient3d_steps - 1,
                                                           interval=args.gradient3d_interval)
        log("3D gradient done")
    cpu = cpu_batch = None
    if not args.no_cpu:
        cpu = cpu_baseline(p, bg, psi, args.cpu_seconds)
        ncpu = host_cpu()["usable_cores"] or 1
        cpu_batch = cpu_baseline(p, bg, psi, args.cpu_seconds, instances=min(16, ncpu))
        log("cpu baselines done")
    line = {
        "metric": METRIC, "value": value, "unit": "Mpoint-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": sec * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "2D forward+adjoint+kernel, one source, plain (256x300, 4000 steps)",
                   "nx": p.nx, "nz": p.nz, "time_steps": maxtime + 1, "sources": 1,
                   "timestep_s": p.timestep, "flow": "supergranule 240 m/s", "npml": p.npmlbot,
                   "misfit": "travel time, receivers 5-60 Mm, 30-minute window"},
        "s_per_gradient": sec,
        "finite": finite,
        "batched": batched,
        "roofline": roofline,
        "column_pass_shape": {k: pipe.fwd.query(k) for k in ("colseg_tc", "colseg_ms", "colseg_p", "colseg_hw",
                                                             "row_wide")},
        "cpu_baseline": cpu,
        "cpu_baseline_batched": cpu_batch,
        "other_configurations": others,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
