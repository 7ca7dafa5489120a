"""Benchmark of the CGFDM time step on one MI355X.

Workload (N=1): CGFDM curvilinear grid 512 x 512 x 256 at h = 100 m, eight seeded
Gaussian hills (0.5-2 km tall, 3-15 km wide), four-layer medium with a 5%
Gaussian random perturbation (1 km correlation), 20-layer CFS-PML on five faces,
traction-image free surface, point moment source, 400 surface stations and the
PGV map, FP32 wavefield. One "step" is one full time step: source injection,
four RK4 stages (derivatives, free surface, PML, update) and the observers.
Inputs are resident in HBM before the timed region starts.

Prints one JSON line (see the keys below). Per-kernel times come from HIP events
recorded on the solver's stream inside the timed region. The default run adds,
as extra keys of the same line:
  accuracy      the 128 x 128 x 64 half-space of the plain-CGFDM workload, 1000
                steps, FP32 against FP64 (station velocity, PGV);
  cpu_baseline  the CPU oracle on the identical 512 x 512 x 256 grid, OpenMP on
                the host's cores and serial, with the core count and CPU model;
  aweno         the ME-AWENO step on the 512 x 512 x 256 basin workload;
  fwd_adjoint   forward + adjoint against the plain forward on the adjoint
                workload (256 x 256 x 128, 100 steps, checkpoint every 100;
                the stage store is allocated before the timed region);
  fwd_adjoint_2000  the same at the adjoint workload's full length: 2000 steps,
                checkpoint every 100 (20 segments; the adjoint recomputes 19);
  config4       the 1024 x 1024 x 320 grid at h = 50 m with FP32, fp16 and bf16
                coefficient storage (throughput; the per-station error against
                FP64 is in `--config4-error` and the GPU tests).

`--scheme aweno` measures the ME-AWENO step alone and prints its own line.
`--adjoint` measures forward + adjoint with its own step count and checkpoint
(`--adjoint-fd NDIR` checks its gradient against central differences).
`--aweno-accuracy STEPS` compares FP32 with FP64 on the ME-AWENO basin workload.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "cgfdm3d-eqr-sc_amd"))

METRIC = "Gcell-updates/s & %HBM peak, FP32 fwd and fwd+adjoint vs plain, at stated L2 err"
HBM_PEAK_GBS = 8000.0
W_BYTES = 36          # 9 fields x 4 B
C_BYTES = {0: 48, 1: 24, 2: 24}  # 12 coefficient planes per cell (fp32 / fp16 / bf16)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def stage_bytes(coef_storage):
    """Algorithmic HBM bytes per cell per RK stage of the fused interior kernel:
    stage 0 and 3 read W and write two arrays (3w + c), stages 1 and 2 read W, the
    RK base and the accumulator and write two arrays (5w + c). Mean over stages."""
    c = C_BYTES[coef_storage]
    return (2 * (3 * W_BYTES + c) + 2 * (5 * W_BYTES + c)) / 4.0


def traffic_from_profiles(name):
    """HBM bytes per launch of the interior kernel from the committed rocprof
    FETCH_SIZE / WRITE_SIZE summary, if present."""
    p = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(p):
        return None
    try:
        with open(p) as f:
            d = json.load(f)
        return d.get("bytes_per_launch")
    except Exception:
        return None


# ME-AWENO algorithmic bytes per cell and TVD stage (FP32): face kernel reads W
# (36 B) and 12 coefficients (48 B) and writes 27 face fluxes (108 B); the update
# reads the 27 fluxes, J, the stage input and the step-start state and writes
# the stage output (108 + 4 + 3 * 36 B).
AW_FACE_BYTES = 36 + 48 + 108
AW_UPDATE_BYTES = 108 + 4 + 3 * 36
# algorithmic bytes per cell and TVD-RK3 step of a fused ME-AWENO step: three
# stages read the stage state (and the step-start state in stages 2 and 3) and
# the 13 coefficients (9 cofactors, J, mu, lambda, 1/rho) and write the next
# state: 8 w + 3 c' = 8 x 36 + 3 x 52 B
AW_ALG_BYTES = 8 * 36 + 3 * 52


def cpu_info():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, os.cpu_count() or 1


def cpu_baseline(model, steps, threads, scheme="cgfdm", serial_steps=1):
    """The CPU oracle (C++ restatement, FP32) on `model`: `steps` steps on
    `threads` OpenMP threads, then `serial_steps` steps on one thread (the
    reference's CPU mode is plain serial loops, axis.h:159-196). Returns a dict
    of the rates and times."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    log(f"cpu baseline: oracle FP32 ({scheme}) on {model.nx}x{model.ny}x{model.nz}, "
        f"{steps} steps on {threads} threads, {serial_steps} serial")
    Orc = oracle.AwenoOracle if scheme == "aweno" else oracle.Oracle
    t0 = time.perf_counter()
    o = Orc(model, prec=4, threads=threads)
    setup_s = time.perf_counter() - t0
    o.step(1)  # first touch
    t0 = time.perf_counter()
    o.step(steps)
    dt = time.perf_counter() - t0
    out = {"omp_gcells": model.cells * steps / dt / 1e9, "omp_s": dt, "setup_s": setup_s}
    if serial_steps > 0:
        oracle.lib().orc_set_threads(1)
        t0 = time.perf_counter()
        o.step(serial_steps)
        ds = time.perf_counter() - t0
        out.update(serial_gcells=model.cells * serial_steps / ds / 1e9, serial_s=ds)
    del o
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--nx", type=int, default=512)
    ap.add_argument("--ny", type=int, default=512)
    ap.add_argument("--nz", type=int, default=256)
    ap.add_argument("--coef", type=int, default=0, help="0 fp32, 1 fp16, 2 bf16 coefficient storage")
    ap.add_argument("--cpu-steps", type=int, default=None,
                    help="OpenMP steps of the CPU baseline on the identical grid (default 3 CGFDM, 1 ME-AWENO)")
    ap.add_argument("--cpu-serial-steps", type=int, default=1, help="serial steps of the CPU baseline (0: none)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--no-aweno", action="store_true", help="skip the ME-AWENO leg of the default run")
    ap.add_argument("--no-config4", action="store_true", help="skip the 1024 x 1024 x 320 leg of the default run")
    ap.add_argument("--scheme", choices=["cgfdm", "aweno"], default="cgfdm")
    ap.add_argument("--adjoint", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip the forward+adjoint measurements of the default run")
    ap.add_argument("--checkpoint", type=int, default=20)
    ap.add_argument("--adjoint-fd", type=int, default=0, metavar="NDIR",
                    help="with --adjoint: check the gradient against central differences in NDIR random directions")
    ap.add_argument("--fd-steps", type=int, default=100, help="time steps of the finite-difference check")
    ap.add_argument("--aweno-accuracy", type=int, default=0, metavar="STEPS",
                    help="only: ME-AWENO basin workload, FP32 against FP64 over STEPS steps")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain 1 GiB device copy (reference for the FETCH_SIZE/WRITE_SIZE counters)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py runs on one GPU only: the single-rank domain holds one GPU's full share; "
              "multi-GPU sharding is a design note, not part of this build.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from cgfdm_amd import model as M
    from cgfdm_amd import solver as S

    if args.cpu_steps is None:
        args.cpu_steps = 1 if args.scheme == "aweno" else 3
    if args.copy_calibration:
        return copy_calibration()
    if args.aweno_accuracy > 0:
        return aweno_accuracy(args)
    if args.scheme == "aweno":
        return bench_aweno(args)
    if args.adjoint and args.adjoint_fd > 0:
        return adjoint_fd_check(args)
    if args.adjoint:
        return bench_adjoint(args)
    t0 = time.perf_counter()
    hmin_gpu = lambda c, md: S.grid_ranges(args.nx, args.ny, args.nz, c, md)["hmin"]
    m = M.config2(nt=args.warmup + args.steps + 1, nx=args.nx, ny=args.ny, nz=args.nz, hmin_fn=hmin_gpu)
    log(f"model {m.nx}x{m.ny}x{m.nz} built in {time.perf_counter() - t0:.1f}s, dt={m.dt:.4e}")

    dev = torch.device("cuda:0")
    stream = torch.cuda.Stream(device=dev)
    g = S.Solver(m, precision=4, coef_storage=args.coef)
    g.set_stream(stream)
    log(f"solver ready: {g.device_bytes() / 2**30:.1f} GiB on device")
    g.step(args.warmup)
    g.synchronize()

    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    g.timing(True)
    ev0.record(stream)
    t0 = time.perf_counter()
    g.step(args.steps)
    ev1.record(stream)
    g.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    tm = g.timing(False)
    ev_ms = ev0.elapsed_time(ev1)
    log(f"timed {args.steps} steps: wall {wall:.3f}s, events {ev_ms / 1e3:.3f}s; kernels {tm}")

    ms_step = wall * 1e3 / args.steps
    cells = m.cells
    gcells = cells * args.steps / wall / 1e9

    # dominant kernel: fused interior stage
    launch_ms = tm["interior_ms"] / max(tm["interior_n"], 1)
    cells_launch = m.nx * m.ny * (m.nz - 3)
    alg_bytes = stage_bytes(args.coef) * cells_launch
    achieved = alg_bytes / (launch_ms * 1e-3) / 1e9
    step_bytes = 4 * stage_bytes(args.coef) * cells
    hbm_step = step_bytes / (ms_step * 1e-3) / 1e9
    kernel_name = g.kernel_name(0)
    traffic_file = f"{kernel_name}_512x512x256_traffic.json"
    traffic = traffic_from_profiles(traffic_file)

    g.close()
    del g
    torch.cuda.empty_cache()
    traffic_src = f"profiles/{traffic_file} (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE of this kernel, copy-calibrated)"

    out = {
        "metric": METRIC,
        "value": gcells,
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {0: "f32", 1: "f32 (fp16 coefficients)", 2: "f32 (bf16 coefficients)"}[args.coef],
        "data": "synthetic",
        "config": {"workload": f"CGFDM curvilinear {m.nx}x{m.ny}x{m.nz} h=100m, 8 Gaussian hills, "
                               f"4 layers + 5% random, PML 20, free surface, point source, "
                               f"{len(m.stations)} stations + PGV, RK4",
                   "cells": cells, "dt": m.dt},
        "hbm_frac_step": hbm_step / HBM_PEAK_GBS,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_src,
                     "kernel": kernel_name, "launch_ms": launch_ms,
                     "alg_bytes_per_cell": stage_bytes(args.coef), "cells_per_launch": cells_launch},
    }
    if not args.no_accuracy:
        out["accuracy"] = accuracy_leg(args.coef)
        out["l2_err_fp32_vs_fp64"] = out["accuracy"]["station_velocity_rel_l2"]
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_leg(m, args)
    else:
        out["cpu_baseline"] = None
    del m
    if not args.no_aweno:
        out["aweno"] = aweno_leg(args)
    if not args.no_adjoint:
        for key, steps in (("fwd_adjoint", 100), ("fwd_adjoint_2000", 2000)):
            ma, t_fwd, t_rec, t_adj, ok = measure_adjoint(steps, 2, 100)
            out[key] = {
                "cost_ratio_fwd_adj_vs_fwd": (t_rec + t_adj) / t_fwd,
                "fwd_adj_gcells": ma.cells * steps / (t_rec + t_adj) / 1e9,
                "forward_gcells": ma.cells * steps / t_fwd / 1e9,
                "forward_s": t_fwd, "forward_record_s": t_rec, "adjoint_s": t_adj,
                "gradients_finite": ok,
                "workload": f"CGFDM forward+adjoint {ma.nx}x{ma.ny}x{ma.nz} h=100m, 4 hills, 4 layers + 5% random, "
                            f"PML 20, 16x8 finite fault, {len(ma.stations)} stations, {steps} steps, "
                            f"checkpoint every 100",
            }
    if not args.no_config4:
        out["config4"] = config4_leg(6, 2)
    print(json.dumps(out), flush=True)


def accuracy_leg(coef):
    """FP32 (with the given coefficient storage) against FP64 on the plain-CGFDM
    half-space (128 x 128 x 64, PML 12, double couple at 3 km, 25 receivers),
    1000 steps: relative L2 of the station velocities (all stations, and the
    worst single station) and of the PGV map."""
    import numpy as np
    from cgfdm_amd import model as M
    from cgfdm_amd import solver as S
    m1 = M.config1(nt=1000)
    a = S.Solver(m1, precision=4, coef_storage=coef)
    a.step(1000)
    b = S.Solver(m1, precision=8)
    b.step(1000)
    sa, sb = a.stations(), b.stations()
    pa, pb = a.pgv(), b.pgv()
    a.close()
    b.close()
    rel = lambda x, y: float(np.linalg.norm(x - y) / np.linalg.norm(y))
    l2 = rel(sa[:, :, :3], sb[:, :, :3])
    worst = max(rel(sa[q, :, :3], sb[q, :, :3]) for q in range(len(m1.stations)))
    log(f"accuracy: config1 1000 steps FP32 vs FP64 station velocity rel L2 = {l2:.3e} (worst station {worst:.3e})")
    return {"workload": "CGFDM half-space 128x128x64 h=100m, PML 12, double couple at 3 km, 25 receivers, 1000 steps",
            "reference": "FP64 run of the same kernels (FP64 kernels = FP64 oracle to 1e-10 in the parity tests)",
            "station_velocity_rel_l2": l2, "worst_station_rel_l2": worst, "pgv_rel_l2": rel(pa, pb)}


def aweno_accuracy(args):
    """FP32 against FP64 on the ME-AWENO basin workload (512 x 512 x 256, 400
    stations, PGV) over `--aweno-accuracy` steps: relative L2 of the station
    velocities (all stations, and the worst station carrying signal) and of the
    PGV map, and the FP32 step time."""
    import numpy as np
    from cgfdm_amd import model as M
    from cgfdm_amd import solver as S
    nst = args.aweno_accuracy
    t0 = time.perf_counter()
    m = M.config3(nt=nst + 1, hmin_fn=lambda c, md: S.grid_ranges(512, 512, 256, c, md)["hmin"])
    out = {}
    for prec in (4, 8):
        g = S.Solver(m, precision=prec, scheme=S.SCHEME_AWENO)
        t1 = time.perf_counter()
        done = 0
        while done < nst:
            n = min(250, nst - done)
            g.step(n)
            g.synchronize()
            done += n
            log(f"aweno accuracy: precision {prec}, {done}/{nst} steps")
        dt = time.perf_counter() - t1
        out[prec] = (g.stations(), g.pgv(), dt)
        g.close()
        log(f"aweno accuracy: precision {prec}, {nst} steps in {dt:.1f} s")
    (s4, p4, t4), (s8, p8, t8) = out[4], out[8]
    rel = lambda x, y: float(np.linalg.norm(x - y) / np.linalg.norm(y))
    peak = np.abs(s8[:, :, :3]).max()
    live = [q for q in range(len(m.stations)) if np.abs(s8[q, :, :3]).max() > 1e-3 * peak]
    res = {"check": "ME-AWENO FP32 vs FP64", "workload": f"ME-AWENO basin {m.nx}x{m.ny}x{m.nz}, {len(m.stations)} "
           f"stations, {nst} steps (dt {m.dt:.4e} s)", "station_velocity_rel_l2": rel(s4[:, :, :3], s8[:, :, :3]),
           "stations_with_signal": len(live),
           "worst_station_rel_l2": max(rel(s4[q, :, :3], s8[q, :, :3]) for q in live),
           "pgv_rel_l2": rel(p4, p8), "fp32_ms_per_step": t4 * 1e3 / nst, "fp64_ms_per_step": t8 * 1e3 / nst,
           "seconds": time.perf_counter() - t0}
    print(json.dumps(res), flush=True)


def cpu_leg(m, args):
    """The CPU oracle on the identical 512 x 512 x 256 grid (a bounded sample:
    `--cpu-steps` steps on the host's cores, one serial step)."""
    model_name, ncpu = cpu_info()
    threads = min(16, ncpu)
    r = cpu_baseline(m, args.cpu_steps, threads, serial_steps=args.cpu_serial_steps)
    log(f"cpu baseline {r['omp_gcells']:.4f} Gcell/s on {threads} threads, serial {r.get('serial_gcells', 0):.5f}")
    return {"value": r["omp_gcells"], "unit": "Gcell-updates/s", "cores": threads, "kind": "port",
            "serial_value": r.get("serial_gcells"), "serial_cores": 1, "nproc": ncpu, "cpu_model": model_name,
            "sample": f"CPU oracle (C++ restatement of the reference CGFDM path, FP32) on the identical "
                      f"{m.nx}x{m.ny}x{m.nz} workload: {args.cpu_steps} full steps on {threads} OpenMP threads "
                      f"({r['omp_s']:.1f} s) and {args.cpu_serial_steps} on one thread ({r.get('serial_s', 0):.1f} s)"}


def aweno_leg(args, steps=30, warmup=3):
    """ME-AWENO step on the 512 x 512 x 256 basin workload; HBM fraction of the
    algorithmic 444 B per cell-step (8 w + 3 c' with w = 36 B, c' = 52 B)."""
    import torch
    from cgfdm_amd import model as M
    from cgfdm_amd import solver as S
    nx, ny, nz = 512, 512, 256
    m = M.config3(nt=warmup + steps + 1, hmin_fn=lambda c, md: S.grid_ranges(nx, ny, nz, c, md)["hmin"])
    g = S.Solver(m, precision=4, scheme=S.SCHEME_AWENO)
    stream = torch.cuda.Stream(device=torch.device("cuda:0"))
    g.set_stream(stream)
    g.step(warmup)
    g.synchronize()
    g.timing(True)
    t0 = time.perf_counter()
    g.step(steps)
    g.synchronize()
    wall = time.perf_counter() - t0
    tm = g.timing(False)
    g.close()
    torch.cuda.empty_cache()
    ms = wall * 1e3 / steps
    gc = m.cells * steps / wall / 1e9
    hbm = AW_ALG_BYTES * m.cells / (ms * 1e-3) / 1e9
    log(f"aweno: {ms:.2f} ms/step, {gc:.3f} Gcell/s, {hbm:.0f} GB/s algorithmic")
    return {"workload": f"ME-AWENO {nx}x{ny}x{nz} h=100m, 8 hills, sharp basin, sponge 20, free surface, point "
                        f"source, {len(m.stations)} stations + PGV, TVD-RK3",
            "gcells": gc, "ms_per_step": ms, "alg_bytes_per_cell_step": AW_ALG_BYTES,
            "hbm_achieved_GBps": hbm, "hbm_frac": hbm / HBM_PEAK_GBS,
            "face_kernel_ms": tm["interior_ms"] / max(tm["interior_n"], 1),
            "update_kernel_ms": tm["surface_ms"] / max(tm["surface_n"], 1), "steps": steps}


def config4_leg(steps, warmup):
    """Throughput on the 1024 x 1024 x 320 grid (h = 50 m, 40 x 20 finite fault)
    with FP32, fp16 and bf16 coefficient storage."""
    import torch
    from cgfdm_amd import model as M
    from cgfdm_amd import solver as S
    nx, ny, nz = 1024, 1024, 320
    t0 = time.perf_counter()
    m = M.config4(nt=warmup + steps + 1, hmin_fn=lambda c, md: S.grid_ranges(nx, ny, nz, c, md)["hmin"])
    log(f"config4 model built in {time.perf_counter() - t0:.1f}s")
    res = {"workload": f"CGFDM {nx}x{ny}x{nz} h=50m, 8 hills, 4 layers + 5% random, PML 20, 40x20 von Karman "
                       f"finite fault, {len(m.stations)} stations", "steps": steps}
    for name, coef in (("fp32", 0), ("fp16", 1), ("bf16", 2)):
        g = S.Solver(m, precision=4, coef_storage=coef)
        g.step(warmup)
        g.synchronize()
        g.timing(True)
        t0 = time.perf_counter()
        g.step(steps)
        g.synchronize()
        wall = time.perf_counter() - t0
        tm = g.timing(False)
        gb = g.device_bytes() / 2 ** 30
        g.close()
        torch.cuda.empty_cache()
        lm = tm["interior_ms"] / max(tm["interior_n"], 1)
        res[name] = {"gcells": m.cells * steps / wall / 1e9, "ms_per_step": wall * 1e3 / steps,
                     "interior_launch_ms": lm, "device_GiB": gb,
                     "interior_alg_GBps": stage_bytes(coef) * nx * ny * (nz - 3) / (lm * 1e-3) / 1e9}
        log(f"config4 {name}: {res[name]}")
    return res


def measure_adjoint(steps, warmup, checkpoint):
    """Plain forward steps against forward_record + adjoint of the same steps on the
    adjoint-gradient workload (256 x 256 x 128, terrain, layered + random medium,
    16 x 8 finite fault, 64 stations), FP32. Returns the model and the three times."""
    import numpy as np
    import torch
    from cgfdm_amd import model as M
    from cgfdm_amd import solver as S

    nx, ny, nz = 256, 256, 128
    hmin_gpu = lambda c, md: S.grid_ranges(nx, ny, nz, c, md)["hmin"]
    m = M.config5(nt=steps + warmup + 3, hmin_fn=hmin_gpu)
    log(f"adjoint model {m.nx}x{m.ny}x{m.nz}, {len(m.src_cells)} fault points, {len(m.stations)} stations")
    g = S.Solver(m, precision=4)
    g.step(warmup)
    g.synchronize()
    t0 = time.perf_counter()
    g.step(steps)
    g.synchronize()
    t_fwd = time.perf_counter() - t0
    g.close()
    torch.cuda.empty_cache()
    a = S.Solver(m, precision=4)
    a.step(warmup)
    # an untimed two-step forward_record + adjoint with the same checkpoint
    # interval allocates the stage store (36 x 4 B per cell and step of a
    # segment, 145 GB here: a one-time cost of up to a few seconds) that the
    # timed pair reuses, as the iterations of an inversion do
    a.adjoint(a.forward_record(2, checkpoint))
    a.synchronize()
    t0 = time.perf_counter()
    rec = a.forward_record(steps, checkpoint)
    t_rec = time.perf_counter() - t0
    t0 = time.perf_counter()
    grads = a.adjoint(rec)
    t_adj = time.perf_counter() - t0
    ok = all(np.isfinite(grads[k]).all() for k in ("gmu", "glam", "gb"))
    a.close()
    torch.cuda.empty_cache()
    log(f"forward {t_fwd:.3f}s, forward_record {t_rec:.3f}s, adjoint {t_adj:.3f}s, "
        f"ratio {(t_rec + t_adj) / t_fwd:.2f}")
    return m, t_fwd, t_rec, t_adj, ok


def bench_adjoint(args):
    """forward_record + adjoint vs plain forward steps, FP32, same workload."""
    m, t_fwd, t_rec, t_adj, ok = measure_adjoint(args.steps, args.warmup, args.checkpoint)
    ratio = (t_rec + t_adj) / t_fwd
    out = {
        "metric": METRIC, "value": m.cells * args.steps / (t_rec + t_adj) / 1e9, "unit": "Gcell-updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": (t_rec + t_adj) * 1e3 / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"CGFDM forward+adjoint {m.nx}x{m.ny}x{m.nz} h=100m, 4 hills, 4 layers + 5% random, "
                               f"PML 20, 16x8 finite fault, {len(m.stations)} stations, checkpoint every "
                               f"{args.checkpoint} steps", "cells": m.cells},
        "forward_gcells": m.cells * args.steps / t_fwd / 1e9,
        "cost_ratio_fwd_adj_vs_fwd": ratio, "gradients_finite": ok,
    }
    print(json.dumps(out), flush=True)


def adjoint_fd_check(args):
    """Gradient check of the adjoint workload (256 x 256 x 128, terrain, layered +
    random medium, PML 20, 16 x 8 finite fault, 64 stations), FP64: L2 waveform
    misfit against a target made with a seeded 3 % Vs perturbation; gradient
    with respect to Vs, Vp, rho per cell and the moment-rate table through the
    torch wrapper over the HIP adjoint; central differences of the HIP forward
    run in NDIR random directions (each perturbing all four parameter sets)."""
    import numpy as np
    import torch
    from cgfdm_amd import autograd as A
    from cgfdm_amd import model as M
    from cgfdm_amd import solver as S

    nst = args.fd_steps
    m = M.config5(nt=nst + 1, hmin_fn=lambda c, md: S.grid_ranges(256, 256, 128, c, md)["hmin"])
    t0 = time.perf_counter()
    mt = M.config5(nt=nst + 1, hmin_fn=lambda c, md: S.grid_ranges(256, 256, 128, c, md)["hmin"])
    pert = M.gaussian_random_field(mt.medium.shape[1:], mt.dh, 1000.0, 0.03, 2024)
    mt.medium = mt.medium.copy()
    mt.medium[0] = (mt.medium[0] * (1.0 + pert)).astype(np.float32)
    target = S.Solver(mt, precision=8).forward_record(nst, nst)

    def misfit(med, rate):
        mm = M.config5(nt=nst + 1, hmin_fn=lambda c, md: S.grid_ranges(256, 256, 128, c, md)["hmin"])
        mm.medium = np.ascontiguousarray(med, np.float32)
        mm.src_rate = np.ascontiguousarray(rate, np.float32)
        rec = S.Solver(mm, precision=8).forward_record(nst, nst)
        return 0.5 * float(((rec - target) ** 2).sum())

    ts = [torch.tensor(m.medium[c], dtype=torch.float64, requires_grad=True) for c in range(3)]
    rate = torch.tensor(m.src_rate, dtype=torch.float64, requires_grad=True)
    rec = A.seismograms(*ts, rate, m, nst, precision=8, checkpoint_every=args.checkpoint)
    loss = 0.5 * ((rec - torch.as_tensor(target)) ** 2).sum()
    loss.backward()
    grads = [t.grad.numpy() for t in ts] + [rate.grad.numpy()]
    base = [m.medium[0].astype(np.float64), m.medium[1].astype(np.float64), m.medium[2].astype(np.float64),
            m.src_rate.astype(np.float64)]
    rng = np.random.default_rng(11)
    errs = []
    for n in range(args.adjoint_fd):
        d = [b * rng.uniform(-0.5, 0.5, b.shape) for b in base]
        eps = 1e-3
        vals, used = [], []
        for sg in (1, -1):
            pm = [(b + sg * eps * dd).astype(np.float32).astype(np.float64) for b, dd in zip(base, d)]
            used.append(pm)
            vals.append(misfit(np.stack(pm[:3]), pm[3]))
        fd = (vals[0] - vals[1]) / (2 * eps)
        an = sum(float((g * (u0 - u1) / (2 * eps)).sum()) for g, u0, u1 in zip(grads, used[0], used[1]))
        errs.append(abs(fd - an) / abs(fd))
        log(f"direction {n}: finite difference {fd:.6e}, adjoint {an:.6e}, relative error {errs[-1]:.2e}")
    out = {"check": "adjoint gradient vs central differences", "workload": f"config 5 {m.nx}x{m.ny}x{m.nz}, "
           f"{nst} steps, checkpoint every {args.checkpoint}, FP64, misfit against a 3% Vs-perturbed target",
           "directions": args.adjoint_fd,
           "max_rel_error": max(errs), "median_rel_error": float(np.median(errs)),
           "seconds": time.perf_counter() - t0}
    print(json.dumps(out), flush=True)


def copy_calibration():
    """Five copies of 1 GiB between two device buffers: the counter reading per
    copy dispatch divided by 2^30 B read (and written) is the scale factor of
    FETCH_SIZE (WRITE_SIZE) for this GPU."""
    import torch
    n = 1 << 28  # float32 elements = 1 GiB
    a = torch.ones(n, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(5):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / 5
    print(json.dumps({"copy_bytes": 4 * n, "copies": 5, "ms_per_copy": ms,
                      "read_plus_write_GBps": 2 * 4 * n / (ms * 1e-3) / 1e9}), flush=True)


def bench_aweno(args):
    import numpy as np
    import torch
    from cgfdm_amd import model as M
    from cgfdm_amd import solver as S

    t0 = time.perf_counter()
    hmin_gpu = lambda c, md: S.grid_ranges(args.nx, args.ny, args.nz, c, md)["hmin"]
    m = M.config3(nt=args.warmup + args.steps + 1, nx=args.nx, ny=args.ny, nz=args.nz, hmin_fn=hmin_gpu)
    log(f"model {m.nx}x{m.ny}x{m.nz} (basin) built in {time.perf_counter() - t0:.1f}s, dt={m.dt:.4e}")
    stream = torch.cuda.Stream(device=torch.device("cuda:0"))
    g = S.Solver(m, precision=4, scheme=S.SCHEME_AWENO)
    g.set_stream(stream)
    log(f"solver ready: {g.device_bytes() / 2**30:.1f} GiB on device")
    g.step(args.warmup)
    g.synchronize()
    torch.cuda.synchronize()
    g.timing(True)
    t0 = time.perf_counter()
    g.step(args.steps)
    g.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    tm = g.timing(False)
    log(f"timed {args.steps} steps: wall {wall:.3f}s; kernels {tm}")
    ms_step = wall * 1e3 / args.steps
    gcells = m.cells * args.steps / wall / 1e9
    launch_ms = tm["interior_ms"] / max(tm["interior_n"], 1)
    upd_ms = tm["surface_ms"] / max(tm["surface_n"], 1)
    # algorithmic bytes of the fused minimum, 444 B per cell-step (8 w + 3 c'),
    # over the measured step time; the face-flux buffer this build writes and
    # re-reads is overhead, reported separately
    achieved = AW_ALG_BYTES * m.cells / (ms_step * 1e-3) / 1e9
    NX, NY, NZ = m.shape_padded
    face_cells = (NX - 5) * (NY - 5) * (NZ - 5)
    flux_overhead = 2 * 27 * 4 * face_cells * 3  # 27 flux planes written and read per stage

    l2 = None
    if not args.no_accuracy:
        m1 = M.small_basin_config(nx=96, ny=96, nz=48, nt=200)
        a = S.Solver(m1, precision=4, scheme=S.SCHEME_AWENO)
        a.step(200)
        b = S.Solver(m1, precision=8, scheme=S.SCHEME_AWENO)
        b.step(200)
        sa, sb = a.stations(), b.stations()
        l2 = float(np.linalg.norm(sa[:, :, :3] - sb[:, :, :3]) / np.linalg.norm(sb[:, :, :3]))
        a.close()
        b.close()
        log(f"accuracy: basin 96x96x48, 200 steps, FP32 vs FP64 station velocity rel L2 = {l2:.3e}")
    g.close()
    torch.cuda.empty_cache()
    cpu = None
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        # bounded sample: the same workload built at half the extent per axis
        # (1/8 of the cells, about 10 s per step on 16 host threads)
        mc = M.config3(nt=args.cpu_steps + 2, nx=args.nx // 2, ny=args.ny // 2, nz=args.nz // 2)
        v, secs = cpu_baseline(mc, args.cpu_steps, threads, scheme="aweno")
        cpu = {"value": v, "unit": "Gcell-updates/s", "cores": threads, "kind": "port",
               "sample": f"CPU oracle (C++ restatement, FP32, OpenMP) of the ME-AWENO step on the same "
                         f"workload at half extent per axis ({mc.nx}x{mc.ny}x{mc.nz}: terrain, basin, "
                         f"sponge, source, stations), {args.cpu_steps} full steps ({secs:.1f}s)"}
    out = {
        "metric": METRIC, "value": gcells, "unit": "Gcell-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"ME-AWENO {m.nx}x{m.ny}x{m.nz} h=100m, 8 Gaussian hills, sharp basin "
                               f"(Vs 400 in Vs 3464), sponge 20, free surface, point source, "
                               f"{len(m.stations)} stations + PGV, TVD-RK3", "cells": m.cells, "dt": m.dt},
        "l2_err_fp32_vs_fp64": l2,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic_from_profiles("aw_faces_512x512x256_traffic.json"),
                     "kernel": "aw_faces + aw_update (whole step)", "face_launch_ms": launch_ms,
                     "update_launch_ms": upd_ms, "alg_bytes_per_cell_step": AW_ALG_BYTES, "cells": m.cells,
                     "flux_buffer_bytes_per_step": flux_overhead,
                     "note": "achieved = 444 B/cell-step (8 w + 3 c', fused minimum) x cells / step time; the "
                             "27-plane face-flux buffer traffic is overhead outside the algorithmic bytes"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
