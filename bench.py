"""Benchmark of the CGFDM time step on one MI355X.

Workload (N=1): CGFDM curvilinear grid 512 x 512 x 256 at h = 100 m, eight seeded
Gaussian hills (0.5-2 km tall, 3-15 km wide), four-layer medium with a 5%
Gaussian random perturbation (1 km correlation), 20-layer CFS-PML on five faces,
traction-image free surface, point moment source, 400 surface stations and the
PGV map, FP32 wavefield. One "step" is one full time step: source injection,
four RK4 stages (derivatives, free surface, PML, update) and the observers.
Inputs are resident in HBM before the timed region starts.

Prints one JSON line (see the keys below). Per-kernel times come from HIP events
recorded on the solver's stream inside the timed region.

`--scheme aweno` measures the ME-AWENO step instead (the 512 x 512 x 256 terrain
grid with a sharp low-velocity basin, TVD-RK3, sponge) and prints its own line.
`--adjoint` measures forward + adjoint against the plain forward on the adjoint
workload (256 x 256 x 128, terrain, layered + random medium, 16 x 8 finite
fault, 64 stations, checkpoints) and prints its own line.
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


def cpu_baseline(model, steps, threads, scheme="cgfdm"):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    log(f"cpu baseline: oracle FP32 ({scheme}) on {threads} threads, {steps} steps of the same grid")
    o = (oracle.AwenoOracle if scheme == "aweno" else oracle.Oracle)(model, prec=4, threads=threads)
    o.step(1)  # first touch
    t0 = time.perf_counter()
    o.step(steps)
    dt = time.perf_counter() - t0
    del o
    return model.cells * steps / dt / 1e9, dt


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
                    help="steps of the CPU baseline sample (default 30 CGFDM, 10 ME-AWENO: about 10 s on 16 threads)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--scheme", choices=["cgfdm", "aweno"], default="cgfdm")
    ap.add_argument("--adjoint", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip the short forward+adjoint measurement of the default run")
    ap.add_argument("--checkpoint", type=int, default=20)
    ap.add_argument("--adjoint-fd", type=int, default=0, metavar="NDIR",
                    help="with --adjoint: check the gradient against central differences in NDIR random directions")
    ap.add_argument("--fd-steps", type=int, default=100, help="time steps of the finite-difference check")
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
        args.cpu_steps = 10 if args.scheme == "aweno" else 30
    if args.copy_calibration:
        return copy_calibration()
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
    traffic = traffic_from_profiles("cg_cell_512x512x256_traffic.json")

    # accuracy leg: FP32 vs FP64 GPU on the 128 x 128 x 64 half-space (config 1), stations
    l2 = None
    if not args.no_accuracy:
        m1 = M.config1(nt=400)
        a = S.Solver(m1, precision=4, coef_storage=args.coef)
        a.step(400)
        b = S.Solver(m1, precision=8)
        b.step(400)
        sa, sb = a.stations(), b.stations()
        l2 = float(np.linalg.norm(sa[:, :, :3] - sb[:, :, :3]) / np.linalg.norm(sb[:, :, :3]))
        a.close()
        b.close()
        log(f"accuracy: config1 400 steps FP32 vs FP64 station velocity rel L2 = {l2:.3e}")

    g.close()
    torch.cuda.empty_cache()

    cpu = None
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        # bounded sample: the same workload built at half the extent per axis
        # (1/8 of the cells, a few seconds per step on 16 host threads)
        mc = M.config2(nt=args.cpu_steps + 2, nx=args.nx // 2, ny=args.ny // 2, nz=args.nz // 2)
        v, secs = cpu_baseline(mc, args.cpu_steps, threads)
        cpu = {"value": v, "unit": "Gcell-updates/s", "cores": threads, "kind": "port",
               "sample": f"CPU oracle (C++ restatement, FP32, OpenMP) of the CGFDM step on the same workload "
                         f"at half extent per axis ({mc.nx}x{mc.ny}x{mc.nz}: terrain, layered + random medium, "
                         f"PML 20, free surface, source, stations), {args.cpu_steps} full steps ({secs:.1f}s)"}
        log(f"cpu baseline {v:.4f} Gcell/s")

    dtype = {0: "f32", 1: "f32 (fp16 coefficients)", 2: "f32 (bf16 coefficients)"}[args.coef]
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
        "dtype": dtype,
        "data": "synthetic",
        "config": {"workload": f"CGFDM curvilinear {m.nx}x{m.ny}x{m.nz} h=100m, 8 Gaussian hills, "
                               f"4 layers + 5% random, PML 20, free surface, point source, "
                               f"{len(m.stations)} stations + PGV, RK4",
                   "cells": cells, "dt": m.dt},
        "hbm_frac_step": hbm_step / HBM_PEAK_GBS,
        "l2_err_fp32_vs_fp64": l2,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "cg_cell", "launch_ms": launch_ms,
                     "alg_bytes_per_cell": stage_bytes(args.coef), "cells_per_launch": cells_launch},
        "cpu_baseline": cpu,
    }
    if not args.no_adjoint:
        # the metric's second part: forward + adjoint against the plain forward, on
        # the adjoint-gradient workload (a short run; `bench.py --adjoint` for more steps)
        ma, t_fwd, t_rec, t_adj, ok = measure_adjoint(20, 2, 10)
        out["fwd_adjoint"] = {
            "cost_ratio_fwd_adj_vs_fwd": (t_rec + t_adj) / t_fwd,
            "fwd_adj_gcells": ma.cells * 20 / (t_rec + t_adj) / 1e9,
            "forward_gcells": ma.cells * 20 / t_fwd / 1e9,
            "gradients_finite": ok,
            "workload": f"CGFDM forward+adjoint {ma.nx}x{ma.ny}x{ma.nz} h=100m, 4 hills, 4 layers + 5% random, "
                        f"PML 20, 16x8 finite fault, {len(ma.stations)} stations, 20 steps, checkpoint every 10",
        }
    print(json.dumps(out), flush=True)


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
    m = M.config5(nt=steps + warmup + 1, hmin_fn=hmin_gpu)
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
    rec = A.seismograms(*ts, rate, m, nst, precision=8, checkpoint_every=max(1, nst // 10))
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
           f"{nst} steps, FP64, misfit against a 3% Vs-perturbed target", "directions": args.adjoint_fd,
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
    NX, NY, NZ = m.shape_padded
    face_cells = (NX - 5) * (NY - 5) * (NZ - 5)
    launch_ms = tm["interior_ms"] / max(tm["interior_n"], 1)
    upd_ms = tm["surface_ms"] / max(tm["surface_n"], 1)
    achieved = AW_FACE_BYTES * face_cells / (launch_ms * 1e-3) / 1e9

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
                     "traffic": traffic_from_profiles("aw_faces_512x512x256_traffic.json"), "kernel": "aw_faces",
                     "launch_ms": launch_ms, "update_launch_ms": upd_ms,
                     "alg_bytes_per_cell": AW_FACE_BYTES, "cells_per_launch": face_cells,
                     "note": "aw_faces is VALU-bound (18 WENO5 interpolations per face); the HBM fraction is "
                             "reported for the algorithmic bytes"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
