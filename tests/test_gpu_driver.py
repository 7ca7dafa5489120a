"""The params.json driver (cgfdm_run) on the MI355X against the CPU oracle.

cgfdm_run reads params.json/station.json, builds the grid, medium and source,
steps the HIP library and writes the reference's files split over a 2 x 1 x 2
rank layout. The same inputs (grid and medium from the driver's own builders,
the source table from its single-source builder with J and mu of the oracle's
setup) are stepped by the FP64 oracle; the files, reassembled as the
reference's plotting scripts do, must match it: station seismograms, the PGV
map and the velocity slices within 1e-4 relative L2 (FP32 GPU vs FP64 oracle),
for the CGFDM and the ME-AWENO builds.
"""
import ctypes
import json
import os
import subprocess

import numpy as np
import pytest

from cgfdm_amd import model as M
import test_host_driver as T

pytestmark = pytest.mark.gpu
H = M.HALO
NX_, NY_, NZ_ = 28, 24, 20
PXYZ = (2, 1, 2)


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def setup_case(tmp_path, scheme):
    lay = tmp_path / "layers.bin"
    rows = [(0.0, 2.0, 3.6, 2.2), (0.4, 3.0, 5.2, 2.6), (1.0, 3.4, 6.0, 2.8)]
    with open(lay, "wb") as f:
        f.write(np.int32(len(rows)).tobytes())
        f.write(np.array(rows, np.float32).tobytes())
    p = dict(TMAX=0.3, DT=0.008, DH=100.0, NX=NX_, NY=NY_, NZ=NZ_, PX=PXYZ[0], PY=PXYZ[1], PZ=PXYZ[2], centerX=14,
             centerY=12, IT_SKIP=10, sliceX=9, sliceY=12, sliceZ=14, sliceFreeSurf=1, nPML=5, gauss_hill=1,
             useTerrain=0, useMedium=1, LayeredModel=1, LayeredFileName=str(lay), useMultiSource=0, sourceX=14,
             sourceY=12, sourceZ=12, rickerfc=4.0, out="./out")
    p["useSingleSource(ricker)"] = 1
    (tmp_path / "params.json").write_text(json.dumps(p))
    st = {"0": [7, 6, NZ_ - 1], "1": [20, 6, NZ_ - 1], "2": [7, 18, NZ_ - 1], "3": [21, 17, NZ_ - 1],
          "4": [14, 12, 8], "5": [3, 20, NZ_ - 1]}
    (tmp_path / "station.json").write_text(json.dumps({"point": 1, "station(point)": st}))
    return p, np.array(list(st.values()), np.int32)


def run_driver(tmp_path, scheme, extra=()):
    cmd = [T.RUNNER, "--quiet", "--scheme", scheme] + list(extra)
    print("driver:", " ".join(cmd), "in", tmp_path, flush=True)
    r = subprocess.run(cmd, cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    return r


def oracle_model(lib, tmp_path, scheme, oracle_mod, nt):
    nx, ny, nz = NX_, NY_, NZ_
    rc, coord, terrain, medium = T.model(lib, str(tmp_path / "params.json"), nx, ny, nz)
    assert rc == 0
    m = M.Model(nx=nx, ny=ny, nz=nz, dh=100.0, dt=float(np.float32(0.008)), nt=nt, npml=5, coord=coord,
                medium=medium, free_surface=True)
    cls = oracle_mod.Oracle if scheme == "cgfdm" else oracle_mod.AwenoOracle
    probe = cls(m, prec=8, threads=8)
    jac, mu = probe.setup(9), probe.setup(10)
    del probe
    rc, cells, rate, dt = T.single(lib, str(tmp_path / "params.json"), nt, np.ascontiguousarray(jac),
                                   np.ascontiguousarray(mu))
    assert rc == 0
    m.src_cells, m.src_rate, m.src_dt = cells, rate, dt
    return m, cls


def gather(path_fmt, spans_a, spans_b, ranks, shape, fast_axis, slow_axis):
    """Reassemble rank files into the whole plane (plotXY.py:80-112)."""
    out = np.zeros(shape, np.float32)
    found = 0
    for (X, Y, Z) in ranks:
        f = path_fmt.format(X=X, Y=Y, Z=Z)
        if not os.path.exists(f):
            continue
        found += 1
        idx = {"X": X, "Y": Y, "Z": Z}
        a0, an = spans_a[idx[fast_axis]]
        b0, bn = spans_b[idx[slow_axis]]
        out[b0:b0 + bn, a0:a0 + an] = np.fromfile(f, np.float32).reshape(bn, an)
    return out, found


@pytest.mark.parametrize("scheme", ["cgfdm", "aweno"])
def test_driver_outputs_match_oracle(scheme, tmp_path, oracle_mod):
    lib = T.host_lib()
    p, st = setup_case(tmp_path, scheme)
    nt = int(0.3 / float(np.float32(0.008)))
    run_driver(tmp_path, scheme)
    out = tmp_path / "out"
    m, cls = oracle_model(lib, tmp_path, scheme, oracle_mod, nt)
    m.stations = (st + H).astype(np.int32)
    o = cls(m, prec=8, threads=8)
    sx, sy, sz = T.spans(NX_, PXYZ[0]), T.spans(NY_, PXYZ[1]), T.spans(NZ_, PXYZ[2])
    ranks = [(X, Y, Z) for X in range(PXYZ[0]) for Y in range(PXYZ[1]) for Z in range(PXYZ[2])]
    # velocity slices after step 20 (file index 20) and the free-surface slice
    o.step(21)
    vx = o.field(0)
    xs, n1 = gather(str(out / "Vx_20_X_mpi_{X}_{Y}_{Z}.bin"), sy, sz, ranks, (NZ_, NY_), "Y", "Z")
    zs, n2 = gather(str(out / "Vx_20_Z_mpi_{X}_{Y}_{Z}.bin"), sx, sy, ranks, (NY_, NX_), "X", "Y")
    fs, n3 = gather(str(out / "FreeSurfVx_20_Z_mpi_{X}_{Y}_{Z}.bin"), sx, sy, ranks, (NY_, NX_), "X", "Y")
    assert n1 == 2 and n2 == 2 and n3 == 2
    if scheme == "aweno":
        # the AWENO oracle holds rho v: velocity = (rho v) / rho
        vx = vx * o.setup(12)
    assert rel_l2(xs, vx[H:-H, H:-H, H + 9]) < 1e-4
    assert rel_l2(zs, vx[H + 14, H:-H, H:-H]) < 1e-4
    assert rel_l2(fs, vx[H + NZ_ - 1, H:-H, H:-H]) < 1e-4
    o.step(nt - 21)
    # station files: indices and seismograms of the stations each rank holds
    so = o.stations()
    got = {}
    for (X, Y, Z) in ranks:
        f = out / f"station_mpi_{X}_{Y}_{Z}.bin"
        if not f.exists():
            continue
        raw = f.read_bytes()
        n = (len(raw)) // (12 + 4 * nt * 9)
        xyz = np.frombuffer(raw[:12 * n], np.int32).reshape(n, 3)
        wave = np.frombuffer(raw[12 * n:], np.float32).reshape(n, nt, 9)
        for q in range(n):
            got[tuple(xyz[q])] = wave[q]
    assert len(got) == len(st)
    for s, (i, j, k) in enumerate(st):
        w = got[(i, j, k)]
        for fam in (slice(0, 3), slice(3, 9)):
            assert rel_l2(w[:, fam], so[s, :, fam]) < 1e-4, (scheme, s, fam)
    # PGV map of the surface ranks
    pg = np.zeros((NY_, NX_, 2), np.float32)
    for (X, Y, Z) in ranks:
        f = out / f"PGV_Z_mpi_{X}_{Y}_{Z}.bin"
        assert f.exists() == (Z == PXYZ[2] - 1)
        if f.exists():
            pg[sy[Y][0]:sum(sy[Y]), sx[X][0]:sum(sx[X])] = np.fromfile(f, np.float32).reshape(sy[Y][1], sx[X][1], 2)
    assert rel_l2(pg, o.pgv()) < 1e-4
    # model slices written at setup
    vs, nv = gather(str(out / "Vs_Z_mpi_{X}_{Y}_{Z}.bin"), sx, sy, ranks, (NY_, NX_), "X", "Y")
    assert nv == 2 and np.array_equal(vs, m.medium[0, H + 14, H:-H, H:-H])


def test_driver_fault_source_matches_oracle(tmp_path, oracle_mod):
    """useMultiSource: a synthetic fault file (own projection, srcCoord2Index, moment
    rates) through the driver against the oracle stepping the driver's own table."""
    lib = T.host_lib()
    p, st = setup_case(tmp_path, "cgfdm")
    lon0, lat0 = 101.246376, 37.7942405
    off = [(0.0, 0.0, -800.0), (0.004, -0.003, -1200.0), (-0.006, 0.002, -1500.0), (0.0, 0.0, -800.0)]
    nt_src = 60
    t = np.arange(nt_src) * 0.01
    rate = np.stack([np.exp(-((t - 0.15 - 0.02 * q) / 0.05) ** 2) for q in range(len(off))])
    rake = np.full((len(off), nt_src), 30.0)
    os.makedirs(tmp_path / "src", exist_ok=True)
    T.write_fault(tmp_path / "src" / "f.bin", [lon0 + o[0] for o in off], [lat0 + o[1] for o in off],
                  [o[2] for o in off], np.full(len(off), 1e6), np.full(len(off), 40.0), np.full(len(off), 70.0),
                  rake, rate, 0.01)
    p.update(useMultiSource=1, sourceDir=str(tmp_path / "src"), sourceFile="f.bin", degree2radian=1,
             centerLongitude=lon0, centerLatitude=lat0)
    p["useSingleSource(ricker)"] = 0
    (tmp_path / "params.json").write_text(json.dumps(p))
    run_driver(tmp_path, "cgfdm")
    out = tmp_path / "out"
    nt = int(0.3 / float(np.float32(0.008)))
    rc, coord, terrain, medium = T.model(lib, str(tmp_path / "params.json"), NX_, NY_, NZ_)
    assert rc == 0
    cells = np.zeros(16, np.int64)
    rt = np.zeros(16 * nt_src * 6, np.float32)
    xyz = np.zeros(3 * len(off), np.float32)
    n, ntv, dtv = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
    rc = lib.h_fault_source(str(tmp_path / "params.json").encode(), T.P(cells, ctypes.c_longlong),
                            T.P(rt, ctypes.c_float), ctypes.byref(n), ctypes.byref(ntv), ctypes.byref(dtv),
                            T.P(xyz, ctypes.c_float), 16)
    assert rc == 0 and n.value == len(off)
    m = M.Model(nx=NX_, ny=NY_, nz=NZ_, dh=100.0, dt=float(np.float32(0.008)), nt=nt, npml=5, coord=coord,
                medium=medium, free_surface=True)
    m.src_cells = cells[:n.value]
    m.src_rate = rt[:n.value * nt_src * 6].reshape(n.value, nt_src, 6)
    m.src_dt = dtv.value
    m.stations = (st + H).astype(np.int32)
    o = oracle_mod.Oracle(m, prec=8, threads=8)
    o.step(nt)
    so = o.stations()
    got = {}
    for (X, Y, Z) in [(X, Y, Z) for X in range(PXYZ[0]) for Y in range(PXYZ[1]) for Z in range(PXYZ[2])]:
        f = out / f"station_mpi_{X}_{Y}_{Z}.bin"
        if not f.exists():
            continue
        raw = f.read_bytes()
        k = len(raw) // (12 + 4 * nt * 9)
        xyzs = np.frombuffer(raw[:12 * k], np.int32).reshape(k, 3)
        wave = np.frombuffer(raw[12 * k:], np.float32).reshape(k, nt, 9)
        for q in range(k):
            got[tuple(xyzs[q])] = wave[q]
    assert len(got) == len(st)
    for s_, (i, j, k) in enumerate(st):
        for fam in (slice(0, 3), slice(3, 9)):
            assert rel_l2(got[(i, j, k)][:, fam], so[s_, :, fam]) < 1e-4, (s_, fam)
    # the source files of the reference layout: one merged cell fewer than the points
    nfiles = [f for f in os.listdir(out) if f.startswith("source_mpi_")]
    total = 0
    for f in nfiles:
        total += int(np.frombuffer((out / f).read_bytes()[:4], np.int32)[0])
    assert total == len(set(cells[:n.value].tolist())) == len(off) - 1


@pytest.mark.parametrize("extra,tol", [(("--precision", "8"), 1e-9), (("--coef", "1"), 5e-3)])
def test_driver_precision_options(extra, tol, tmp_path, oracle_mod):
    """--precision 8 (FP64 wavefield) and --coef 1 (FP16 coefficient storage) through the
    driver: station files against the FP64 oracle on the same inputs."""
    lib = T.host_lib()
    p, st = setup_case(tmp_path, "cgfdm")
    nt = int(0.3 / float(np.float32(0.008)))
    run_driver(tmp_path, "cgfdm", extra)
    m, cls = oracle_model(lib, tmp_path, "cgfdm", oracle_mod, nt)
    m.stations = (st + H).astype(np.int32)
    o = cls(m, prec=8, threads=8)
    o.step(nt)
    so = o.stations()
    out = tmp_path / "out"
    got = {}
    for X in range(PXYZ[0]):
        for Y in range(PXYZ[1]):
            for Z in range(PXYZ[2]):
                f = out / f"station_mpi_{X}_{Y}_{Z}.bin"
                if not f.exists():
                    continue
                raw = f.read_bytes()
                k = len(raw) // (12 + 4 * nt * 9)
                xyz = np.frombuffer(raw[:12 * k], np.int32).reshape(k, 3)
                wave = np.frombuffer(raw[12 * k:], np.float32).reshape(k, nt, 9)
                for q in range(k):
                    got[tuple(xyz[q])] = wave[q]
    for s_, (i, j, k) in enumerate(st):
        for fam in (slice(0, 3), slice(3, 9)):
            # the files hold float32 samples: the FP64 run agrees to float rounding
            assert rel_l2(got[(i, j, k)][:, fam], so[s_, :, fam]) < max(tol, 1e-6), (extra, s_, fam)


def read_station_files(out, nt):
    got = {}
    for X in range(PXYZ[0]):
        for Y in range(PXYZ[1]):
            for Z in range(PXYZ[2]):
                f = out / f"station_mpi_{X}_{Y}_{Z}.bin"
                if not f.exists():
                    continue
                raw = f.read_bytes()
                k = len(raw) // (12 + 4 * nt * 9)
                xyz = np.frombuffer(raw[:12 * k], np.int32).reshape(k, 3)
                wave = np.frombuffer(raw[12 * k:], np.float32).reshape(k, nt, 9)
                for q in range(k):
                    got[tuple(xyz[q])] = wave[q]
    return got


def srtm_tile(tdir):
    """A 5-degree SRTM tile (100-105 E, 35-40 N, 6000 x 6000 floats): a gentle ramp
    with latitude plus a smooth 300 m hill of 1 km width under the grid centre."""
    P = 6000
    J = np.arange(P, dtype=np.float32)
    tile = np.repeat((100.0 + 2000.0 * J / (P - 1))[:, None], P, axis=1).astype(np.float32)
    r0, c0 = int(round(2.5 / 5.0 * P)), int(round(2.5 / 5.0 * P))
    rr = np.arange(-60, 61)
    bump = 300.0 * np.exp(-0.5 * (rr[:, None] ** 2 + rr[None, :] ** 2) / 12.0 ** 2)
    tile[r0 - 60:r0 + 61, c0 - 60:c0 + 61] += bump.astype(np.float32)
    os.makedirs(tdir, exist_ok=True)
    tile.tofile(tdir / "srtm_35N100E.bin")


def shen_files(mdir):
    nlayer = int((2.0e3 + 3e4) / 100.0)
    os.makedirs(mdir, exist_ok=True)
    for a in range(3):
        for b in range(3):
            lon, lat = 102.0 + 0.5 * a, 37.0 + 0.5 * b
            K = np.arange(nlayer + 5)
            vs = 2.0 + 0.01 * K + 0.3 * (lon - 102.0) + 0.2 * (lat - 37.0)
            np.savetxt(mdir / f"{lon:.1f}_{lat:.1f}.mod", np.stack([K * 0.1, vs, np.full_like(vs, 0.01)], axis=1),
                       fmt="%.6f")


def crust1_files(cdir):
    nlon, nlat, nl = 360, 180, 9
    tops = np.array([0.0, -0.3, -1.3, -1.3, -2.4, -6.0, -7.5, -9.0, -12.0])  # 0.3 km top layer is skipped
    I = np.arange(nlon)[None, :].repeat(nlat, 0).ravel()
    base = 1.6 + 0.25 * np.arange(nl)
    vs = base[None, :] + 0.001 * I[:, None]
    os.makedirs(cdir, exist_ok=True)
    for name, arr in (("vs", vs), ("vp", 1.8 * vs), ("rho", 2.0 + 0.1 * np.arange(nl)[None, :] + 0 * I[:, None]),
                      ("bnds", tops[None, :] + 0 * I[:, None])):
        np.savetxt(cdir / f"crust1.{name}", arr, fmt="%.5f")


REAL_DATA = {
    # SRTM terrain-following grid with the layered file (layers measured from the local surface)
    "srtm_terrain": lambda d: (srtm_tile(d / "tiles"), dict(
        useTerrain=1, SRTM90=1, TerrainDir=str(d / "tiles"), lonStart=100, latStart=35, blockX=1, blockY=1,
        centerLongitude=102.5, centerLatitude=37.5, gauss_hill=0, **{"Depth(km)": 2.0}))[1],
    # Shen model medium from .mod profiles on a 0.5-degree grid
    "shen": lambda d: (shen_files(d / "shen"), dict(
        LayeredModel=0, ShenModel=1, MLonStart=102.0, MLatStart=37.0, MLonEnd=103.0, MLatEnd=38.0, MLonStep=0.5,
        MLatStep=0.5, MVeticalStep=100.0, MediumDir=str(d / "shen"), centerLongitude=102.5,
        centerLatitude=37.5))[1],
    # CRUST1.0 medium (9-layer 1-degree files)
    "crust1": lambda d: (crust1_files(d / "crust"), dict(
        LayeredModel=0, Crust_1Model=1, CrustDir=str(d / "crust"), CrustLonStep=1.0, CrustLatStep=1.0,
        centerLongitude=102.5, centerLatitude=37.5))[1],
    # Gaussian-STF double couple (singleSource.cpp:217-359) instead of the Ricker explosion
    "double_couple": lambda d: {"useSingleSource(ricker)": 0, "useSingleSource(double_couple)": 1,
                                "strike": 30.0, "dip": 60.0, "rake": 90.0, "Mw": 4.0, "duration": 0.2},
}


@pytest.mark.parametrize("case", list(REAL_DATA))
def test_driver_real_data_builders_match_oracle(case, tmp_path, oracle_mod):
    """Synthetic SRTM tile, Shen and CRUST1.0 files and the double-couple single
    source through the driver on the MI355X; the FP64 oracle steps the grid,
    medium and source table that the driver's own builders make from the same
    files: station files within 1e-4 relative L2."""
    lib = T.host_lib()
    p, st = setup_case(tmp_path, "cgfdm")
    p.update(REAL_DATA[case](tmp_path))
    (tmp_path / "params.json").write_text(json.dumps(p))
    nt = int(0.3 / float(np.float32(0.008)))
    run_driver(tmp_path, "cgfdm")
    m, cls = oracle_model(lib, tmp_path, "cgfdm", oracle_mod, nt)
    if case == "srtm_terrain":
        relief = float(np.ptp(m.coord[2, H + NZ_ - 1]))
        assert relief > 50.0, relief  # the hill is in the grid
    m.stations = (st + H).astype(np.int32)
    o = cls(m, prec=8, threads=8)
    o.step(nt)
    so = o.stations()
    got = read_station_files(tmp_path / "out", nt)
    assert len(got) == len(st)
    for s_, (i, j, k) in enumerate(st):
        for fam in (slice(0, 3), slice(3, 9)):
            assert rel_l2(got[(i, j, k)][:, fam], so[s_, :, fam]) < 1e-4, (case, s_, fam)
