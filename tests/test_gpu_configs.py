"""The benchmark workloads at their full sizes on the MI355X.

* 128 x 128 x 64 half-space (plain CGFDM, 12-layer PML, double couple at 3 km,
  2 Hz Ricker-type rate, 25 receivers): 1000 FP32 steps, station seismograms and
  PGV against the FP64 oracle (relative L2 <= 1e-4).
* 512 x 512 x 256 curvilinear grid (8 hills, 4 layers + 5 % random medium,
  20-layer PML, 400 stations) and the same grid with the sharp basin for the
  ME-AWENO scheme: every field of the FP32 and FP64 GPU runs against the FP32
  and FP64 oracle after 2 steps from a smooth random state (every cell of the
  grid carries signal, PML and surface layers included), and FP32 against
  FP64 GPU runs of the point-source problem over 500 steps (stations, PGV).
* 1024 x 1024 x 320 at h = 50 m with fp16 and bf16 coefficient storage and a
  40 x 20 von Karman finite fault: per-station error against the FP64 run of
  the same kernels; the field offsets pass 2^31 elements here.
* 256 x 256 x 128 adjoint workload: the adjoint gradient of an L2 waveform
  misfit against central differences of the FP64 forward in 3 random
  directions.

Each test prints its case before the GPU work starts. Tolerances are stated
per test: FP64 vs FP64 1e-10 (rounding order only); FP32 vs the FP32 oracle
3e-5 (CGFDM) and 1e-5 (ME-AWENO); FP32 vs FP64 1e-4.
"""
import math

import numpy as np
import pytest

from cgfdm_amd import model as M

pytestmark = pytest.mark.gpu
H = M.HALO


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    n = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (n if n > 0 else 1.0))


def say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def S():
    from cgfdm_amd import solver
    return solver


def gpu_hmin(S, nx, ny, nz):
    return lambda c, md: S.grid_ranges(nx, ny, nz, c, md)["hmin"]


def smooth_state(shape, seed, scheme):
    """Nine smooth random components on the padded grid (zero ghost cells):
    velocities ~ 1 m/s and stresses ~ 1 MPa for CGFDM, momentum ~ 2700 kg/m^2/s
    and strains ~ 1e-4 for ME-AWENO."""
    base = [M.gaussian_random_field(shape, 1.0, 4.0, 1.0, seed + q) for q in range(3)]
    out = []
    for v in range(9):
        f = base[v % 3] * (1.0 + 0.3 * (v // 3)) + 0.2 * base[(v + 1) % 3]
        if scheme == "cgfdm":
            f = f * (1.0 if v < 3 else 1e6)
        else:
            f = f * (2700.0 if v < 3 else 1e-4)
        g = np.zeros(shape)
        g[H:-H, H:-H, H:-H] = f[H:-H, H:-H, H:-H]
        out.append(g)
    return out


def test_config1_full_fp32_vs_fp64_oracle(S, oracle_mod):
    m = M.config1(nt=1000)
    say(f"config1 {m.nx}x{m.ny}x{m.nz}, PML {m.npml}, 1000 steps: GPU FP32 vs FP64 oracle")
    g = S.Solver(m, precision=4)
    g.step(1000)
    sg, pg = g.stations(), g.pgv()
    g.close()
    o = oracle_mod.Oracle(m, prec=8, threads=16)
    o.step(1000)
    so, po = o.stations(), o.pgv()
    errs = [rel_l2(sg[s, :, :3], so[s, :, :3]) for s in range(len(m.stations))]
    say(f"config1: station velocity rel L2 max {max(errs):.2e} median {np.median(errs):.2e}; "
        f"all fields {rel_l2(sg, so):.2e}; PGV {rel_l2(pg, po):.2e}")
    assert max(errs) < 1e-4
    # velocities and the in-plane stresses per component; the traction components
    # Tzz, Txz, Tyz vanish on the free surface, so their error is measured against
    # the size of the whole stress record
    for c in (0, 1, 2, 3, 4, 6):
        assert rel_l2(sg[:, :, c], so[:, :, c]) < 1e-4, c
    stress = np.linalg.norm(so[:, :, 3:])
    for c in (5, 7, 8):
        assert np.linalg.norm(sg[:, :, c] - so[:, :, c]) / stress < 1e-4, c
    assert rel_l2(pg, po) < 1e-4


@pytest.mark.parametrize("scheme", ["cgfdm", "aweno"])
def test_config2_config3_full_grid_two_steps_vs_oracle(scheme, S, oracle_mod):
    nx, ny, nz = 512, 512, 256
    build = M.config2 if scheme == "cgfdm" else M.config3
    m = build(nt=4, hmin_fn=gpu_hmin(S, nx, ny, nz))
    sch = S.SCHEME_CGFDM if scheme == "cgfdm" else S.SCHEME_AWENO
    Orc = oracle_mod.Oracle if scheme == "cgfdm" else oracle_mod.AwenoOracle
    init = smooth_state(m.shape_padded, 31, scheme)
    tol32 = 3e-5 if scheme == "cgfdm" else 1e-5
    for prec, tol in ((8, 1e-10), (4, tol32)):
        say(f"{m.name} {scheme} {nx}x{ny}x{nz}: smooth random state, 2 steps, GPU vs oracle, precision {prec}")
        g = S.Solver(m, precision=prec, scheme=sch)
        for v in range(9):
            g.set_field(v, init[v])
        g.step(2)
        g.synchronize()
        wg = [g.field(v) for v in range(9)]
        g.close()
        o = Orc(m, prec=prec, threads=16)
        for v in range(9):
            o.set_field(v, init[v])
        o.step(2)
        for v in range(9):
            err = rel_l2(wg[v], o.field(v))
            say(f"  field {v}: rel L2 {err:.2e}")
            assert err < tol, (prec, v, err)
        del o, wg


@pytest.mark.parametrize("scheme", ["cgfdm", "aweno"])
def test_config2_config3_fp32_vs_fp64_500_steps(scheme, S):
    nx, ny, nz = 512, 512, 256
    build = M.config2 if scheme == "cgfdm" else M.config3
    m = build(nt=500, hmin_fn=gpu_hmin(S, nx, ny, nz))
    sch = S.SCHEME_CGFDM if scheme == "cgfdm" else S.SCHEME_AWENO
    say(f"{m.name} {scheme} {nx}x{ny}x{nz}: point source, 500 steps, GPU FP32 vs GPU FP64")
    out = {}
    for prec in (4, 8):
        g = S.Solver(m, precision=prec, scheme=sch)
        g.step(500)
        out[prec] = (g.stations(), g.pgv())
        g.close()
    (s4, p4), (s8, p8) = out[4], out[8]
    live = [s for s in range(len(m.stations)) if np.abs(s8[s, :, :3]).max() > 1e-3 * np.abs(s8[:, :, :3]).max()]
    errs = [rel_l2(s4[s, :, :3], s8[s, :, :3]) for s in live]
    say(f"{m.name}: {len(live)} stations with signal, velocity rel L2 max {max(errs):.2e}; "
        f"all stations {rel_l2(s4[:, :, :3], s8[:, :, :3]):.2e}; PGV {rel_l2(p4, p8):.2e}")
    assert len(live) >= 10
    assert rel_l2(s4[:, :, :3], s8[:, :, :3]) < 1e-4
    assert max(errs) < 1e-3
    assert rel_l2(p4, p8) < 1e-4


def test_config4_16bit_coefficients_vs_fp64(S):
    nx, ny, nz = 1024, 1024, 320
    steps = 300
    m = M.config4(nt=steps, hmin_fn=gpu_hmin(S, nx, ny, nz))
    N = int(np.prod(m.shape_padded))
    say(f"config4 {nx}x{ny}x{nz} h=50m, {len(m.src_cells)} fault points, {len(m.stations)} stations, "
        f"{steps} steps: fp16 / bf16 coefficients vs FP64 (9N = {9 * N} > 2^31)")
    assert 9 * N > 2 ** 31
    rec = {}
    for key, prec, coef in (("fp64", 8, 0), ("fp16", 4, 1), ("bf16", 4, 2), ("fp32", 4, 0)):
        g = S.Solver(m, precision=prec, coef_storage=coef)
        g.step(steps)
        rec[key] = g.stations()
        g.close()
        say(f"  {key} done")
    ref = rec["fp64"]
    live = [s for s in range(len(m.stations)) if np.abs(ref[s, :, :3]).max() > 1e-3 * np.abs(ref[:, :, :3]).max()]
    assert len(live) >= 8, len(live)
    worst = {}
    for key in ("fp32", "fp16", "bf16"):
        errs = [rel_l2(rec[key][s, :, :3], ref[s, :, :3]) for s in live]
        worst[key] = max(errs)
        say(f"  {key}: {len(live)} stations with signal, velocity rel L2 max {max(errs):.2e} "
            f"median {np.median(errs):.2e}")
    assert worst["fp32"] < 1e-4
    assert worst["fp16"] < 1e-3
    assert worst["bf16"] < 1e-2


def test_config5_adjoint_vs_central_differences(S):
    from cgfdm_amd import autograd as A
    import torch
    # 300 steps (about 1 s): the P waves of the fault reach the stations, so the
    # misfit and its gradient are not the numerical tail ahead of the wave front
    steps, ndir = 300, 3
    hm = gpu_hmin(S, 256, 256, 128)
    m = M.config5(nt=steps + 1, hmin_fn=hm)
    say(f"config5 {m.nx}x{m.ny}x{m.nz}, {len(m.src_cells)} fault points, {len(m.stations)} stations, "
        f"{steps} steps, FP64: adjoint vs central differences in {ndir} directions")
    mt = M.config5(nt=steps + 1, hmin_fn=hm)
    pert = M.gaussian_random_field(mt.medium.shape[1:], mt.dh, 1000.0, 0.03, 2024)
    mt.medium = mt.medium.copy()
    mt.medium[0] = (mt.medium[0] * (1.0 + pert)).astype(np.float32)
    target = S.Solver(mt, precision=8).forward_record(steps, steps)
    assert np.abs(target).max() > 1e-12, "no signal at the stations"

    def misfit(med, rate):
        mm = M.config5(nt=steps + 1, hmin_fn=hm)
        mm.medium = np.ascontiguousarray(med, np.float32)
        mm.src_rate = np.ascontiguousarray(rate, np.float32)
        r = S.Solver(mm, precision=8).forward_record(steps, steps)
        return 0.5 * float(((r - target) ** 2).sum())

    ts = [torch.tensor(m.medium[c], dtype=torch.float64, requires_grad=True) for c in range(3)]
    rate = torch.tensor(m.src_rate, dtype=torch.float64, requires_grad=True)
    rec = A.seismograms(*ts, rate, m, steps, precision=8, checkpoint_every=100)
    loss = 0.5 * ((rec - torch.as_tensor(target)) ** 2).sum()
    loss.backward()
    grads = [t.grad.numpy() for t in ts] + [rate.grad.numpy()]
    base = [m.medium[c].astype(np.float64) for c in range(3)] + [m.src_rate.astype(np.float64)]
    rng = np.random.default_rng(17)
    for n in range(ndir):
        d = [b * rng.uniform(-0.5, 0.5, b.shape) for b in base]
        eps = 1e-3
        vals, used = [], []
        for sg in (1, -1):
            pm = [(b + sg * eps * dd).astype(np.float32).astype(np.float64) for b, dd in zip(base, d)]
            used.append(pm)
            vals.append(misfit(np.stack(pm[:3]), pm[3]))
        fd = (vals[0] - vals[1]) / (2 * eps)
        an = sum(float((g * (u0 - u1) / (2 * eps)).sum()) for g, u0, u1 in zip(grads, used[0], used[1]))
        err = abs(fd - an) / abs(fd)
        say(f"  direction {n}: central difference {fd:.6e}, adjoint {an:.6e}, relative error {err:.2e}")
        assert err < 1e-3
