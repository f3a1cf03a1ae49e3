#!/usr/bin/env python3
"""Benchmark of the gfx950 Rast-3DMHD time step (one JSON line on stdout).

Workload: the production hydro grid, 504 x 504 x 2048 global interior
(default geometry), as virtual 4 x 2 subdomains on one GPU, default physics
(Re=2, Pr=f4(0.05), theta=0.25, grav=0.625), ampt=f4(0.005) with the
reference's ran2 seed; W warm-up steps then K timed RK3 steps.

value  = interior cells * 3 stages * K / wall seconds of the timed region
         (state resident in HBM; the host only launches).
roofline: the fused stage kernel, timed with HIP events on the library's
         stream over the timed region; algorithmic bytes per launch =
         133.33 B per cell (the 400 B/cell/step of reading X, writing X' and
         Z, reading Z, ... over 5 FP64 fields, split over 3 stages); also
         against the measured bandwidth of a 4 GiB device-to-device copy
         (measured_copy_gbs, frac_of_copy).
cpu_baseline: the reference CPU build (oracle/_ref/3dmhd_nodump, MPI +
         OpenMP) on a bounded sample of the same workload, run before the GPU
         is touched; falls back to the C oracle when MPI or the build is absent.
secondary: after the headline workload, the default run also measures the
         metric's other workloads and adds them to its line: rotating MHD
         252x252x512 (20 steps), the MHD adjoint 128x128x256 (backward :
         forward wall-time ratio over 20 steps) and the 64-member ensemble
         against the same members stepped one after another
         (--no-secondary skips them).
"""
import argparse
import json
import os
import re
import shutil
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
METRIC = "FP64 cell-updates/s per RK3 stage, % HBM roofline; adjoint:forward time ratio"
HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL_STAGE = 400.0 / 3.0   # hydro: 10 passes x 5 fields x 8 B per step
BYTES_PER_CELL_STAGE_MHD = 640.0 / 3.0  # MHD: 10 passes x 8 fields x 8 B per step
MHD_FLAGS = ["--lmag", "1", "--rm", "50", "--beta", "10", "--ampb", "1", "--bfh", "2",
             "--bzp", "10", "--lrot", "1", "--ry", "0.1", "--ang", "0"]
CPU_CORES = 16                        # the host's CPU share per GPU on the MI355X boxes


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


def cpu_baseline(npx, npy, npey, npez, cores):
    """Time the reference CPU build on a bounded sample of the workload."""
    ref = os.path.join(ROOT, "oracle", "_ref", "3dmhd_nodump")
    mpirun = shutil.which("mpirun") or "/opt/conda/bin/mpirun"
    nsample = 256  # reduced z extent: ~15 GB of reference state instead of ~115 GB
    # steps 2..14 are timed from the reference's WALL log (0.1 s resolution):
    # about 12 s at this size, so the resolution is under 1%
    steps = 14
    if os.path.exists(ref) and os.path.exists(mpirun):
        np_ = npey * npez
        per = max(1, cores // np_)
        env = dict(os.environ, OMP_NUM_THREADS=str(per), OMP_PROC_BIND="true",
                   OMP_PLACES="cores")
        args = [ref, "--npx", str(npx), "--npy", str(npy), "--npz", str(nsample),
                "--npey", str(npey), "--ntotal", str(steps), "--nstep0", "1"]
        # each rank bound to its own `per` cores, its OpenMP threads one per core
        for bind in (["-bind-to", f"core:{per}"], []):
            cmd = [mpirun] + bind + ["-np", str(np_)] + args
            log(f"[bench] cpu baseline: {' '.join(cmd)} (OMP_NUM_THREADS={per})")
            try:
                p = subprocess.run(cmd, env=env, capture_output=True, text=True, timeout=400,
                                   cwd="/tmp")
            except (OSError, subprocess.TimeoutExpired) as e:
                log(f"[bench] reference run unavailable: {e}")
                break
            walls = {int(m.group(1)): float(m.group(2)) for m in
                     re.finditer(r"NIT (\d+) .*WALL ([0-9.]+)s", p.stderr)}
            if p.returncode == 0 and 1 in walls and steps in walls:
                per_step = (walls[steps] - walls[1]) / (steps - 1)
                cells = npx * npy * nsample
                return {"value": cells * 3 / per_step, "unit": "cell-updates/s",
                        "cores": np_ * per, "kind": "reference",
                        "sample": f"reference CPU build (MPI {npey}x{npez} ranks x {per} OpenMP "
                                  f"threads{', each rank bound to its own cores' if bind else ''}"
                                  f") on {cpu_model()}, {npx}x{npy}x{nsample} (z reduced from "
                                  f"2048), steps 2..{steps} ({walls[steps] - walls[1]:.1f} s) "
                                  f"timed from its own WALL log"}
            log(f"[bench] reference run failed (rc={p.returncode}): {p.stderr[-400:]}")
    # fallback: the C oracle (same sweeps, OpenMP) on a smaller sample
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    os.environ["OMP_NUM_THREADS"] = str(cores)
    import oracle
    nz = 64
    sim = oracle.Sim(["--npx", str(npx), "--npy", str(npy), "--npz", str(nz),
                      "--npey", str(npey), "--npez", str(npez)])
    sim.step(1)
    t0 = time.perf_counter()
    sim.step(2)
    dt = time.perf_counter() - t0
    return {"value": npx * npy * nz * 3 * 2 / dt, "unit": "cell-updates/s", "cores": cores,
            "kind": "port", "sample": f"C oracle (OpenMP), {npx}x{npy}x{nz}, 2 steps"}


def measure_copy_gbs(nbytes=4 << 30, reps=5):
    """Device-to-device copy bandwidth (bytes read + written per second) of
    the runtime's own copy kernel, timed with HIP events: the practical
    streaming ceiling the stage kernel's rate is also compared with."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) != 0:
        return None
    try:
        if hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) != 0:
            return None
        e0, e1 = ctypes.c_void_p(), ctypes.c_void_p()
        hip.hipEventCreate(ctypes.byref(e0))
        hip.hipEventCreate(ctypes.byref(e1))
        d2d = 3  # hipMemcpyDeviceToDevice
        hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), d2d)
        best = None
        for _ in range(reps):
            hip.hipEventRecord(e0, None)
            hip.hipMemcpyAsync(b, a, ctypes.c_size_t(nbytes), d2d, None)
            hip.hipEventRecord(e1, None)
            hip.hipEventSynchronize(e1)
            ms = ctypes.c_float()
            hip.hipEventElapsedTime(ctypes.byref(ms), e0, e1)
            best = ms.value if best is None else min(best, ms.value)
        hip.hipEventDestroy(e0)
        hip.hipEventDestroy(e1)
        hip.hipFree(b)
        return 2.0 * nbytes / (best / 1e3) / 1e9
    finally:
        hip.hipFree(a)


def load_package():
    import importlib.util
    path = os.path.join(ROOT, "rast-3dmhd_amd", "__init__.py")
    spec = importlib.util.spec_from_file_location(
        "rast3dmhd_amd", path, submodule_search_locations=[os.path.dirname(path)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["rast3dmhd_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


# Physics branches beyond the reference's time step, timed on the same grid.
PHYSICS = {
    "plain": [],
    "heat": ["--itcon", "2", "--tp", "1.5", "--zp", "20.0", "--hh", "4.0"],
    "density": ["--ngrid", "0", "--id", "1", "--dh", "0.4", "--dv", "0.3"],
    "relax": ["--ngrid", "0", "--itcon", "0", "--izcon", "1", "--rlax", "0.5", "--toff", "0.5",
              "--tstart", "-1"],
}
PHYSICS_NAMES = {"heat": "ITCON=2 embedded heat loss (TP 1.5, ZP 20, HH 4)",
                 "density": "density diffusion ID=1 (DH 0.4, DV 0.3, uniform grid)",
                 "relax": "flux relaxation RLAX 0.5, TOFF 0.5 (ITCON=0, IZCON=1, uniform grid)"}


def measure_adjoint(pkg, sizes, mhd, steps):
    """Forward and backward wall times of the discrete adjoint (J = final
    kinetic energy, checkpoints every 20 steps) after one untimed pair."""
    npx, npy, npz, npey, npez = sizes
    args = ["--npx", str(npx), "--npy", str(npy), "--npz", str(npz),
            "--npey", str(npey), "--npez", str(npez)] + (MHD_FLAGS if mhd else [])
    run = pkg.Run(args)
    adj = pkg.Adjoint(run.cfg(), run.metrics(), 20)
    adj.set_initial(run.state())
    run.close()
    log(f"[bench] adjoint: {steps} steps of {args}")
    # warm-up: kernel loading and the first touch of every checkpoint and
    # segment buffer
    adj.forward(steps)
    adj.backward()
    J = adj.forward(steps)
    adj.backward()
    fwd_ms, bwd_ms = adj.timing()
    adj.close()
    log(f"[bench] J = {J:.9e}; forward {fwd_ms:.1f} ms, backward {bwd_ms:.1f} ms")
    return fwd_ms, bwd_ms


def run_adjoint(a):
    """Adjoint:forward time ratio of the hand-written discrete adjoint."""
    pkg = load_package()
    cells = a.npx * a.npy * a.npz
    fwd_ms, bwd_ms = measure_adjoint(pkg, (a.npx, a.npy, a.npz, a.npey, a.npez), a.mhd,
                                     a.steps)
    out = {
        "metric": METRIC, "value": bwd_ms / fwd_ms, "unit": "adjoint:forward time ratio",
        "n_gpus": 1, "steps": a.steps, "warmup": 1, "ms_per_step": bwd_ms / a.steps,
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic: the reference's static polytrope with its seeded ran2 perturbation",
        "config": {"workload": f"{'MHD' if a.mhd else 'hydro'} adjoint {a.npx}x{a.npy}x{a.npz}, "
                               "checkpoints every 20 steps, J = final kinetic energy",
                   "global_cells": cells, "forward_ms": fwd_ms, "backward_ms": bwd_ms,
                   "forward_cell_updates_per_s": cells * 3 * a.steps / (fwd_ms / 1e3)},
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def run_checkpoint(a):
    """Cost of a checkpoint: an asynchronous device snapshot (X, Z and the
    scalars; include/r3d_hip.h) taken while the run keeps stepping, against
    the same steps without one and against a blocking download of the same
    fields (how the reference's dumps stop its GPU run)."""
    pkg = load_package()
    args = ["--npx", str(a.npx), "--npy", str(a.npy), "--npz", str(a.npz),
            "--npey", str(a.npey), "--npez", str(a.npez)] + (MHD_FLAGS if a.mhd else [])
    run = pkg.Run(args)
    cells = a.npx * a.npy * a.npz
    sim = run.make_sim()
    run.close()
    nf = 8 if a.mhd else 5
    snap_bytes = 2 * nf * (a.npx + 2) * (a.npy + 2) * (a.npz + 2) * 8
    log(f"[bench] checkpoint: {args}, snapshot of {snap_bytes / 1e9:.1f} GB")
    sim.step(a.warmup)
    sim.snapshot_begin()  # first use allocates the buffers
    sim.snapshot_rank(0, ["ru"])
    sim.sync()

    def timed(fn):
        t0 = time.perf_counter()
        fn()
        return time.perf_counter() - t0

    def steps():
        sim.step(a.steps)
        sim.sync()

    def snap_alone():
        sim.snapshot_begin()
        sim.snapshot_rank(0, ["ru"])

    t_plain = timed(steps)
    t_snap = timed(snap_alone)
    t0 = time.perf_counter()
    sim.snapshot_begin()
    sim.step(a.steps)
    sim.sync()
    t_steps_with = time.perf_counter() - t0
    sim.snapshot_rank(0, ["ru"])
    t_total_with = time.perf_counter() - t0
    names = pkg.PRIMARIES[:nf] + ["z" + n for n in pkg.PRIMARIES[:nf]]
    t_block = timed(lambda: sim.download_global(names))
    sim.close()
    log(f"[bench] {a.steps} steps {t_plain * 1e3:.1f} ms; snapshot alone {t_snap * 1e3:.1f} ms; "
        f"steps with a snapshot in flight {t_steps_with * 1e3:.1f} ms (snapshot done after "
        f"{t_total_with * 1e3:.1f} ms); blocking download {t_block * 1e3:.1f} ms")
    out = {
        "metric": METRIC, "value": t_steps_with / t_plain - 1.0,
        "unit": "step-time overhead of a checkpoint snapshot in flight",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": t_steps_with / a.steps * 1e3, "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic: the reference's static polytrope with its seeded ran2 perturbation",
        "config": {"workload": f"{'MHD' if a.mhd else 'hydro'} {a.npx}x{a.npy}x{a.npz} checkpoint",
                   "global_cells": cells, "snapshot_bytes": snap_bytes,
                   "steps_ms": t_plain * 1e3, "snapshot_alone_ms": t_snap * 1e3,
                   "steps_with_snapshot_ms": t_steps_with * 1e3,
                   "snapshot_complete_ms": t_total_with * 1e3,
                   "blocking_download_ms": t_block * 1e3},
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def measure_ensemble(pkg, nmem, steps, warmup):
    """The batched ensemble (autotuned tiles; graph replay, then the same
    steps as one persistent cooperative launch) and the same members stepped
    one after another through the plain path."""
    members = pkg.ensemble_members(nmem)
    runs = [pkg.Run(m) for m in members]
    cells = sum(r.npx * r.npy * r.npz for r in runs)
    log(f"[bench] ensemble of {len(runs)} members, {cells} cells")
    ens = pkg.Ensemble(runs)
    ms_tuned = ens.autotune(3)
    ty, kc, tw, nt = ens.tile_shape()
    log(f"[bench] autotuned tiles: ty={ty} kc={kc} tw={tw} ({nt} tiles/stage), "
        f"{ms_tuned:.3f} ms/step")
    ens.step(warmup)
    ens.sync()
    t0 = time.perf_counter()
    ens.step(steps)
    ens.sync()
    wall = time.perf_counter() - t0
    # the same steps as one persistent cooperative launch
    ens.set_persistent(True)
    ens.step(1)
    ens.sync()
    t0 = time.perf_counter()
    ens.step(steps)
    ens.sync()
    wall_p = time.perf_counter() - t0
    mem = ens.memory_bytes()
    ens.close()
    log(f"[bench] ensemble {wall / steps * 1e3:.3f} ms/step (graph replay), "
        f"{wall_p / steps * 1e3:.3f} ms/step (one persistent launch); sequential reference next")
    seq = 0.0
    for r in runs:
        sim = r.make_sim()
        sim.step(warmup)
        sim.sync()
        t0 = time.perf_counter()
        sim.step(steps)
        sim.sync()
        seq += time.perf_counter() - t0
        sim.close()
        r.close()
    return {"members": len(runs), "cells": cells, "wall": wall, "seq": seq, "tiles": (ty, kc, tw),
            "mem": mem, "wall_persistent": wall_p}


def run_ensemble(a):
    """Aggregate cell-updates/s of the batched ensemble (autotuned tiles, graph
    replay) next to the same members stepped one after another."""
    pkg = load_package()
    e = measure_ensemble(pkg, a.ensemble, a.steps, a.warmup)
    cells, wall, seq, (ty, kc, tw), mem = e["cells"], e["wall"], e["seq"], e["tiles"], e["mem"]
    value = cells * 3 * a.steps / wall
    seq_value = cells * 3 * a.steps / seq
    achieved = BYTES_PER_CELL_STAGE * cells * 3 * a.steps / wall / 1e9
    out = {
        "metric": METRIC, "value": value, "unit": "cell-updates/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": wall / a.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic: seeded per-member polytropes and ran2 perturbations",
        "config": {"workload": f"ensemble of {e['members']} hydro members (48..96)^2 x 96..192, "
                               "Re in [1,4], theta in [0.1,0.3], virtual 1x2 each",
                   "global_cells": cells, "tiles": [ty, kc, tw], "graph": True,
                   "device_bytes": mem, "sequential_value": seq_value,
                   "speedup_vs_sequential": value / seq_value,
                   "persistent_value": cells * 3 * a.steps / e["wall_persistent"],
                   "persistent_speedup_vs_sequential": seq / e["wall_persistent"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "note": "whole-step rate (all kernels of the step), not one kernel"},
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    # grid and virtual ranks; unset sizes take the mode's workload: hydro
    # 504x504x2048 (4x2), --mhd 252x252x512 (2x4), --adjoint 128x128x256 (2x4)
    ap.add_argument("--npx", type=int, default=None)
    ap.add_argument("--npy", type=int, default=None)
    ap.add_argument("--npz", type=int, default=None)
    ap.add_argument("--npey", type=int, default=None)
    ap.add_argument("--npez", type=int, default=None)
    ap.add_argument("--mhd", action="store_true",
                    help="rotating MHD with the magnetic layer (Rm=50, beta=10, ampb=1, bfh=2, "
                         "bzp=10, ry=0.1); default size 252x252x512 (2x4 virtual ranks)")
    ap.add_argument("--ensemble", type=int, default=0, metavar="M",
                    help="batched ensemble of M seeded hydro members (sizes 48..96 x 96..192, "
                         "Re in [1,4], theta in [0.1,0.3]) against the same members run one "
                         "after another through the plain path")
    ap.add_argument("--adjoint", action="store_true",
                    help="time the discrete adjoint (J = final kinetic energy, checkpoints every "
                         "20 steps) against its forward run; default size 128x128x256 (2x4 "
                         "virtual ranks); the specified workload adds --mhd --steps 200")
    ap.add_argument("--checkpoint", action="store_true",
                    help="cost of an asynchronous checkpoint snapshot taken while stepping, "
                         "against plain steps and a blocking download")
    ap.add_argument("--variant", type=int, default=0, choices=(0, 1),
                    help="0: bit-exact stage kernel (default); 1: FMA-contracted stage kernel")
    ap.add_argument("--physics", default="plain", choices=sorted(PHYSICS),
                    help="plain (default); heat: the ITCON=2 embedded heat loss; density: "
                         "density diffusion ID=1 (uniform grid); relax: flux relaxation")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-secondary", action="store_true",
                    help="skip the MHD, adjoint and ensemble figures that the default run "
                         "adds to its line after the headline workload")
    ap.add_argument("--copy-calib-gib", type=float, default=0.0,
                    help="before the run, time one device-to-device copy of this many GiB "
                         "(a plain copy of known size, used to scale profiler byte counters)")
    a = ap.parse_args()
    if a.adjoint:
        sizes = (128, 128, 256, 2, 4)
    elif a.mhd:
        sizes = (252, 252, 512, 2, 4)
    else:
        sizes = (504, 504, 2048, 4, 2)
    for name, v in zip(("npx", "npy", "npz", "npey", "npez"), sizes):
        if getattr(a, name) is None:
            setattr(a, name, v)
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {a.gpus}); "
              "the multi-GPU sharding is a design, not a launched path", file=sys.stderr)
        sys.exit(2)

    if a.ensemble:
        return run_ensemble(a)
    if a.adjoint:
        return run_adjoint(a)
    if a.checkpoint:
        return run_checkpoint(a)
    extra = (MHD_FLAGS if a.mhd else []) + PHYSICS[a.physics]
    bpc = BYTES_PER_CELL_STAGE_MHD if a.mhd else BYTES_PER_CELL_STAGE
    plain = a.physics == "plain"
    # the reference CPU build refuses the two extra physics branches
    cpu = None if (a.no_cpu_baseline or a.mhd or not plain) else cpu_baseline(
        a.npx, a.npy, a.npey, a.npez, CPU_CORES)

    if a.copy_calib_gib > 0:
        import torch
        n = int(a.copy_calib_gib * 2**30 / 8)
        src = torch.ones(n, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"[bench] copy calibration: {n * 8} bytes copied once")
        del src, dst
        torch.cuda.empty_cache()

    pkg = load_package()
    args = ["--npx", str(a.npx), "--npy", str(a.npy), "--npz", str(a.npz),
            "--npey", str(a.npey), "--npez", str(a.npez)] + extra
    wall, avg_stage_s, mem = measure_steps(pkg, args, a.variant, a.steps, a.warmup)
    cells = a.npx * a.npy * a.npz
    value = cells * 3 * a.steps / wall
    achieved = bpc * cells / avg_stage_s / 1e9
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "stage_kernel_traffic.json")
    if os.path.exists(tfile):
        # HBM bytes per stage launch from the committed rocprofv3 FETCH_SIZE /
        # WRITE_SIZE passes of this workload (null for any other workload)
        try:
            for t in json.load(open(tfile)):
                if (plain and t.get("npx") == a.npx and t.get("npy") == a.npy
                        and t.get("npz") == a.npz and bool(t.get("mhd", False)) == a.mhd):
                    traffic = t.get("hbm_bytes_per_launch")
        except (OSError, ValueError, AttributeError):
            traffic = None
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": wall / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: the reference's static polytrope with its seeded ran2 "
                "temperature perturbation (no dataset involved)",
        "config": {"workload": f"{'rotating MHD' if a.mhd else 'hydro'} "
                               f"{a.npx}x{a.npy}x{a.npz}, virtual "
                               f"{a.npey}x{a.npez} subdomains on one GPU, "
                               f"{'FMA-contracted' if a.variant else 'bit-exact'} variant"
                               + ("" if plain else f", {PHYSICS_NAMES[a.physics]}"),
                   "global_cells": cells, "stages_per_step": 3,
                   "device_bytes": mem, "hbm_roofline_frac_of_value":
                       value * bpc / (HBM_PEAK_GBS * 1e9)},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic},
        "cpu_baseline": cpu,
    }
    copy = measure_copy_gbs()
    if copy:
        out["roofline"]["measured_copy_gbs"] = copy
        out["roofline"]["frac_of_copy"] = achieved / copy
    if plain and not a.mhd and a.variant == 0 and not a.no_secondary and (
            a.npx, a.npy, a.npz, a.npey, a.npez) == (504, 504, 2048, 4, 2):
        out["secondary"] = secondary(pkg)
    print(json.dumps(out), flush=True)


def measure_steps(pkg, args, variant, steps, warmup):
    """Wall seconds of `steps` timed steps after `warmup`, the stage kernel's
    average launch seconds (HIP events on the library's stream) and the
    device bytes of the run."""
    log(f"[bench] host setup {' '.join(args)}")
    t0 = time.perf_counter()
    run = pkg.Run(args)
    log(f"[bench] host setup done in {time.perf_counter() - t0:.1f}s; uploading")
    sim = run.make_sim(variant)
    run.close()
    mem = sim.memory_bytes()
    log(f"[bench] device state {mem / 2**30:.1f} GiB; warm-up {warmup} steps")
    sim.step(warmup)
    sim.sync()
    sim.set_timing(True)
    log(f"[bench] timing {steps} steps")
    t0 = time.perf_counter()
    sim.step(steps)
    sim.sync()
    wall = time.perf_counter() - t0
    n_stage, ms_stage = sim.kernel_stats(0)
    n_bound, ms_bound = sim.kernel_stats(1)
    sim.set_timing(False)
    sc = sim.scalars()
    log(f"[bench] nit={sc['nit']} dt={sc['dt']:.6g} umach={sc['umach']:.6g}; "
        f"stage kernel {n_stage} launches avg {ms_stage / max(n_stage, 1):.3f} ms, "
        f"boundary kernel avg {ms_bound / max(n_bound, 1):.3f} ms")
    sim.close()
    return wall, ms_stage / max(n_stage, 1) / 1e3, mem


def secondary(pkg):
    """The other workloads of the metric, measured after the headline one in
    the same run: rotating MHD 252x252x512 (2x4), the MHD adjoint
    128x128x256 (2x4, 20 steps) and the 64-member ensemble."""
    out = {}
    cells = 252 * 252 * 512
    args = ["--npx", "252", "--npy", "252", "--npz", "512", "--npey", "2", "--npez", "4"]
    wall, avg_stage_s, _ = measure_steps(pkg, args + MHD_FLAGS, 0, 20, 5)
    ach = BYTES_PER_CELL_STAGE_MHD * cells / avg_stage_s / 1e9
    out["mhd"] = {"workload": "rotating MHD 252x252x512, virtual 2x4, bit-exact", "steps": 20,
                  "value": cells * 3 * 20 / wall, "unit": "cell-updates/s",
                  "ms_per_step": wall / 20 * 1e3, "stage_kernel_achieved_gbs": ach,
                  "stage_kernel_frac": ach / HBM_PEAK_GBS}
    fwd_ms, bwd_ms = measure_adjoint(pkg, (128, 128, 256, 2, 4), True, 20)
    out["adjoint"] = {"workload": "MHD adjoint 128x128x256, virtual 2x4, 20 steps, checkpoints "
                                  "every 20, J = final kinetic energy",
                      "value": bwd_ms / fwd_ms, "unit": "adjoint:forward time ratio",
                      "forward_ms": fwd_ms, "backward_ms": bwd_ms}
    e = measure_ensemble(pkg, 64, 20, 5)
    ev = e["cells"] * 3 * 20 / e["wall"]
    out["ensemble"] = {"workload": "ensemble of 64 hydro members (48..96)^2 x 96..192, Re in "
                                   "[1,4], theta in [0.1,0.3], graph replay", "steps": 20,
                       "value": ev, "unit": "cell-updates/s",
                       "sequential_value": e["cells"] * 3 * 20 / e["seq"],
                       "speedup_vs_sequential": e["seq"] / e["wall"],
                       "persistent_value": e["cells"] * 3 * 20 / e["wall_persistent"],
                       "persistent_speedup_vs_sequential": e["seq"] / e["wall_persistent"],
                       "frac": BYTES_PER_CELL_STAGE * ev / (HBM_PEAK_GBS * 1e9)}
    return out


if __name__ == "__main__":
    main()
