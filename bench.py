"""Benchmark of the 2D sparc solver on one MI355X.

Workload (the configuration the headline metric is quoted on): production-size
2D forward + adjoint + kernel, FP64, one source, plain method -- nx=256 by
nz=300 on the adiabatic polytrope with 12-point vertical PMLs, steady
supergranule flow (peak 240 m/s), 4000 time steps forward and 4000 adjoint,
windowed vz misfit over receivers 5-60 Mm from the source, kernels for psi and
c.  One bench "step" = one full gradient (forward + misfit + adjoint + kernel
accumulation) with all inputs resident in HBM.

    python bench.py [--steps K] [--warmup W] [--gpus 1]

Prints one JSON line: value = Mpoint-steps/s (nx*nz*time-steps*(forward +
adjoint)*sources / wall-s / 1e6), plus the batched 16-source gradient, the
roofline of the dominant kernel (HIP events) and the CPU baseline (the C
oracle on the host, one core, bounded sample).  other_configurations holds
the plain and batched 2D forward runs and the 3D (option 16) forward and
checkpointed gradient at 256x256x300; `--only-3d` prints just those two.
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


def measured_traffic(kernel: str, ms: int, nseg: int):
    """HBM bytes per launch of the column pass from the committed rocprofv3
    --pmc FETCH_SIZE / WRITE_SIZE summary, when it was taken with this kernel
    and segmentation (the line `traffic: <kernel><MS,0> P=.. fetch_kib=..
    write_kib=.. fetch_scale=.. write_scale=..`); None otherwise."""
    import re
    if not os.path.exists(TRAFFIC_FILE):
        return None
    m = re.search(r"^traffic: (k_col\w+)<(\d+),0> P=(\d+) fetch_kib=([\d.]+) write_kib=([\d.]+) "
                  r"fetch_scale=([\d.]+) write_scale=([\d.]+)", open(TRAFFIC_FILE).read(), re.M)
    if not m or m.group(1) != kernel or int(m.group(2)) != ms or int(m.group(3)) != nseg:
        return None
    return (float(m.group(4)) * float(m.group(6)) + float(m.group(5)) * float(m.group(7))) * 1024.0


def make_config(nx=256, nz=300, npml=12):
    from sparc_flows_amd import background as B
    from sparc_flows_amd.params import SparcParams

    bg = B.polytrope(nz)
    dt = B.stable_timestep(bg)
    p = SparcParams(nx=nx, nz=nz, xlength=200e8, timestep=dt, npmlbot=npml, npmltop=npml, npmlhor=9)
    psi = B.supergranule_psi(nx, p.xlength, bg)
    return p, bg, psi


def forward_rate(nx, nz, nsrc, nsteps, flow, dc, reps=1):
    """Forward-only Mpoint-steps/s of a batch (the plain-forward and batched
    multi-source configurations: 128x150 polytrope, 12-point PML)."""
    import torch
    from sparc_flows_amd import background as B
    from sparc_flows_amd.solver import Sparc

    p, bg, psi = make_config(nx, nz)
    c2d = B.sound_speed_bump(nx, p.xlength, bg) if dc else None
    s = Sparc(p, bg, nsrc=nsrc, c2d=c2d, psi=psi if flow else None)
    xs = list(B.source_positions(nsrc, seed=0)) if nsrc > 1 else [0.0]
    s.forward(xs, nsteps - 1)                          # warm-up (and autotune)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        s.forward(xs, nsteps - 1)
    torch.cuda.synchronize()
    sec = (time.perf_counter() - t0) / reps
    return {"Mpoint_steps_per_s": nx * nz * nsteps * nsrc / sec / 1e6, "s_per_run": sec,
            "grid": [nx, nz], "sources": nsrc, "time_steps": nsteps, "flow": flow, "dc": dc,
            "column_pass_shape": {k: s.query(k) for k in ("colseg_tc", "colseg_ms", "colseg_p")}}


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


def cpu_baseline(p, bg, psi, seconds=12.0):
    """Time the C oracle (one core) on a bounded sample of the same forward run."""
    import oracle

    oracle.build()
    o = oracle.Oracle(p, bg, psi=psi)
    o.set_forward_source(0.0, 100000)
    n = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        o.step(n)
        n += 1
    dt = time.perf_counter() - t0
    return {"value": p.nx * p.nz * n / dt / 1e6, "unit": "Mpoint-steps/s", "cores": 1, "kind": "port",
            "sample": f"{n} forward time steps of the {p.nx}x{p.nz} workload (one source), "
                      f"{dt:.1f} s on one host core"}


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
                    help="skip the plain-forward (128x150, 500 steps) and batched 16-source forward "
                         "(128x150 with flow and sound-speed bump, 4000 steps) lines")
    ap.add_argument("--only-3d", action="store_true",
                    help="only run the 3D forward and gradient lines (256x256x300) and print them")
    ap.add_argument("--gradient3d-steps", type=int, default=450,
                    help="time steps of the 3D checkpointed gradient line (configuration 5: 3000)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run a plain device-to-device copy of this many MiB ten times "
                         "(calibrates the FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import torch
    from sparc_flows_amd.gradient import GradientPipeline

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    if args.copy_probe_mib > 0:
        n = args.copy_probe_mib * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        for _ in range(10):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": n * 8, "launches": 10}))
        return
    if args.only_3d:
        print(json.dumps({"forward3d_256x256x300": forward3d_rate(),
                          "gradient3d_256x256x300": gradient3d_rate(maxtime=args.gradient3d_steps - 1)}))
        return
    p, bg, psi = make_config()
    maxtime = args.nsteps - 1
    pipe = GradientPipeline(p, bg, nsrc=1, psi=psi)
    xsrc = [0.0]

    def one():
        return pipe.run(xsrc, maxtime)

    def log(msg):
        print(f"[bench] {msg}", file=sys.stderr, flush=True)

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
    finite = bool(torch.isfinite(out["kc"]).all() and torch.isfinite(out["kpsi"]).all())

    npts = p.nx * p.nz
    C = 2  # x-varying coefficient arrays read per point (v0x, v0z)
    bytes_per_pt = {"col": (14 + C) * 8, "col_last": (14 + C) * 8, "row": 9 * 8, "row_first": 7 * 8,
                    "row_last": 4 * 8}

    # batched gradient of args.batch sources vs the plain one
    batched = None
    if args.batch > 0:
        from sparc_flows_amd import background as B
        xs = list(B.source_positions(args.batch, seed=0))
        bpipe = GradientPipeline(p, bg, nsrc=args.batch, psi=psi)
        bpipe.run(xs, maxtime)
        torch.cuda.synchronize()
        log(f"batched warmup ({args.batch} sources) done")
        tb0 = time.perf_counter()
        bpipe.run(xs, maxtime)
        torch.cuda.synchronize()
        bsec = time.perf_counter() - tb0
        log(f"batched gradient {bsec:.2f} s")
        # the batch's column pass against HBM: every launch moves args.batch
        # times the single-source bytes, more than the 256 MB MALL holds
        bms = bpipe.fwd.profile_kernels(200)
        bach = bytes_per_pt["col"] * npts * args.batch / (bms["col"] * 1e-3) / 1e9
        batched = {"sources": args.batch, "s_per_gradient": bsec,
                   "Mpoint_steps_per_s": pts * args.batch / bsec / 1e6,
                   "speedup_vs_sequential_plain": (sec * args.batch) / bsec,
                   "roofline_col": {"achieved": bach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                    "frac": bach / HBM_PEAK_GBS, "ms_per_launch": bms},
                   "shape": {k: bpipe.fwd.query(k) for k in ("colseg_tc", "colseg_ms", "row_wide")}}
        del bpipe

    # roofline of the dominant kernel: HIP events on the solver's stream
    ms = pipe.fwd.profile_kernels(200)
    per_step_ms = {"col": 4 * ms["col"], "col_last": ms["col_last"], "row": 4 * ms["row"],
                   "row_first": ms["row_first"], "row_last": ms["row_last"]}
    dom = max(per_step_ms, key=per_step_ms.get)
    achieved = bytes_per_pt[dom] * npts / (ms[dom] * 1e-3) / 1e9
    colk = "k_colsplit"
    traffic = (measured_traffic(colk, pipe.fwd.query("colseg_ms"), pipe.fwd.query("colseg_p"))
               if dom == "col" else None)
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": dom,
                "ms_per_launch": ms, "algorithmic_bytes_per_point": bytes_per_pt[dom]}

    log("kernel profile done")
    others = None
    if not args.no_configs:
        others = {"plain_forward_128x150": forward_rate(128, 150, 1, 500, False, False, reps=3),
                  "batched16_forward_128x150": forward_rate(128, 150, 16, 4000, True, True),
                  "forward3d_256x256x300": forward3d_rate(),
                  "gradient3d_256x256x300": gradient3d_rate(maxtime=args.gradient3d_steps - 1)}
        log("forward configurations done")
    cpu = None if args.no_cpu else cpu_baseline(p, bg, psi, args.cpu_seconds)
    log("cpu baseline done")
    line = {
        "metric": METRIC, "value": value, "unit": "Mpoint-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": sec * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "2D forward+adjoint+kernel, one source, plain (256x300, 4000 steps)",
                   "nx": p.nx, "nz": p.nz, "time_steps": maxtime + 1, "sources": 1,
                   "timestep_s": p.timestep, "flow": "supergranule 240 m/s", "npml": p.npmlbot},
        "s_per_gradient": sec,
        "finite": finite,
        "batched": batched,
        "roofline": roofline,
        "column_pass_shape": {k: pipe.fwd.query(k) for k in ("colseg_tc", "colseg_ms", "colseg_p", "colseg_hw",
                                                             "row_wide")},
        "cpu_baseline": cpu,
        "other_configurations": others,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
