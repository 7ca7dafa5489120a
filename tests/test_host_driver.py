"""Host side of the params.json driver (cgfdm3d-eqr-sc_amd/host/), on the CPU.

The reference's own example inputs (examples/Example_finite_fault/params.json
and station.json, kept as fixtures under tests/golden/reference_example/) are
read through the parser; the builders are checked against closed forms and the
writers by reassembling their rank-split files the way the reference's plotting
scripts do (plotPGV.py:85-117, plotWaveform.py:60-110).
"""
import ctypes
import json
import math
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "cgfdm3d-eqr-sc_amd", "host")
BUILD = os.path.join(ROOT, "tests", "host", "_build")
LIB = os.path.join(BUILD, "libhost_capi.so")
GOLD = os.path.join(ROOT, "tests", "golden", "reference_example")
RUNNER = os.path.join(ROOT, "cgfdm3d-eqr-sc_amd", "cgfdm_amd", "cgfdm_run")
H = 3
SRCS = ["json_lite.cpp", "params.cpp", "model_build.cpp", "source_build.cpp", "terrain.cpp", "io_files.cpp"]


_LIB = []


def host_lib():
    """Build (when stale) and load the host-side test library."""
    if _LIB:
        return _LIB[0]
    os.makedirs(BUILD, exist_ok=True)
    srcs = [os.path.join(HOST, s) for s in SRCS] + [os.path.join(ROOT, "tests", "host", "host_capi.cpp")]
    newest = max(os.path.getmtime(s) for s in srcs + [os.path.join(HOST, f) for f in os.listdir(HOST)])
    if not os.path.exists(LIB) or os.path.getmtime(LIB) < newest:
        subprocess.check_call(["g++", "-std=c++17", "-O2", "-shared", "-fPIC", "-I", HOST, "-o", LIB] + srcs)
    L = ctypes.CDLL(LIB)
    L.h_last_msg.restype = ctypes.c_char_p
    L.h_aeqd.argtypes = [ctypes.c_double] * 6 + [ctypes.POINTER(ctypes.c_double)]
    L.h_double_couple.argtypes = [ctypes.c_float] * 3 + [ctypes.POINTER(ctypes.c_float)]
    _LIB.append(L)
    return L


@pytest.fixture(scope="module")
def lib():
    return host_lib()


def P(a, t):
    return a.ctypes.data_as(ctypes.POINTER(t))


def params(lib, path):
    buf = ctypes.create_string_buffer(16384)
    n = lib.h_params(path.encode(), buf, len(buf))
    if n < 0:
        return -n, None
    return 0, json.loads(buf.value.decode())


def write_params(tmp_path, **kw):
    base = dict(TMAX=0.3, DT=0.01, DH=100.0, NX=20, NY=18, NZ=16, PX=1, PY=1, PZ=1, centerX=10, centerY=9,
                IT_SKIP=5, sliceX=10, sliceY=9, sliceZ=8, sliceFreeSurf=1, nPML=4, gauss_hill=0, useTerrain=0,
                useMedium=0, useMultiSource=0, sourceX=10, sourceY=9, sourceZ=8, rickerfc=2.0, out=str(tmp_path / "out"))
    base["useSingleSource(ricker)"] = 1
    base.update(kw)
    p = tmp_path / "params.json"
    p.write_text(json.dumps(base))
    return str(p)


# ---------------------------------------------------------------------------
# params.json
# ---------------------------------------------------------------------------
def test_reference_example_params_parse(lib):
    rc, p = params(lib, os.path.join(GOLD, "params.json"))
    assert rc == 0
    assert (p["NX"], p["NY"], p["NZ"], p["PX"], p["PY"], p["PZ"]) == (801, 601, 401, 2, 2, 2)
    assert p["DH"] == 200.0 and p["DT"] == pytest.approx(1.3e-2) and p["TMAX"] == pytest.approx(60.005)
    # lower-case "out" and "srtm90" fill OUT and SRTM90 (cJSON key lookup ignores case)
    assert p["OUT"] == "./output" and p["SRTM90"] == 1
    assert p["useMultiSource"] == 1 and p["sourceFile"] == "Chenkejie_source106.bin"
    assert p["ShenModel"] if "ShenModel" in p else True
    assert p["Depth"] == 80.0 and p["useTerrain"] == 1 and p["degree2radian"] == 1
    assert p["strike"] == pytest.approx(103) and p["Mw"] == pytest.approx(6.6, rel=1e-7)
    assert p["CrustDir"] == "./crust1.0" and p["MLonStep"] == 0.5
    # NT = TMAX / DT with DT held in a float
    assert p["NT"] == int(60.005 / float(np.float32(1.3e-2))) == 4615


def test_params_defaults_and_errors(lib, tmp_path):
    path = write_params(tmp_path, useTerrain=0)
    rc, p = params(lib, path)
    assert rc == 0
    # without terrain the depth is DH * NZ (km)
    assert p["Depth"] == pytest.approx(100.0 * 16 * 1e-3)
    assert p["ricker"] == 1 and p["double_couple"] == 0
    assert params(lib, str(tmp_path / "missing.json"))[0] == 100
    bad = tmp_path / "bad.json"
    bad.write_text('{"NX": 10, "NY": }')
    assert params(lib, str(bad))[0] == 1001
    wrong = tmp_path / "wrong.json"
    wrong.write_text('{"out": 3}')
    assert params(lib, str(wrong))[0] == 1001


def test_reference_example_stations(lib, tmp_path):
    path = write_params(tmp_path, NX=801, NY=601, NZ=401)
    g = np.zeros(3 * 16, np.int32)
    n = lib.h_stations(path.encode(), os.path.join(GOLD, "station.json").encode(), P(g, ctypes.c_int), 16)
    assert n == 4
    assert g[:12].tolist() == [388, 66, 400, 453, 136, 400, 474, 183, 400, 470, 217, 400]
    # stations outside the grid are dropped, as readStationIndex does
    path = write_params(tmp_path, NX=460, NY=601, NZ=401)
    assert lib.h_stations(path.encode(), os.path.join(GOLD, "station.json").encode(), P(g, ctypes.c_int), 16) == 2
    assert lib.h_stations(path.encode(), str(tmp_path / "none.json").encode(), P(g, ctypes.c_int), 16) == 0


# ---------------------------------------------------------------------------
# grid and medium
# ---------------------------------------------------------------------------
def model(lib, path, nx, ny, nz):
    NX, NY, NZ = nx + 2 * H, ny + 2 * H, nz + 2 * H
    c = np.zeros((3, NZ, NY, NX), np.float32)
    t = np.zeros((NY, NX), np.float32)
    m = np.zeros((3, NZ, NY, NX), np.float32)
    rc = lib.h_model(path.encode(), P(c, ctypes.c_float), P(t, ctypes.c_float), P(m, ctypes.c_float))
    return rc, c, t, m


def test_flat_grid_and_homogeneous_medium(lib, tmp_path):
    path = write_params(tmp_path)
    rc, c, t, m = model(lib, path, 20, 18, 16)
    assert rc == 0
    # x = 0 at the centre column, z = 0 on the top interior layer, spacing DH
    assert c[0, 5, 5, H + 10] == 0 and c[1, 5, H + 9, 5] == 0
    assert np.all(c[2, H + 16 - 1] == 0)
    assert np.allclose(np.diff(c[0], axis=2), 100) and np.allclose(np.diff(c[1], axis=1), 100)
    assert np.allclose(np.diff(c[2], axis=0), 100)
    assert np.all(t == 0)
    assert np.all(m[0] == 3300) and np.all(m[1] == 5716) and np.all(m[2] == 2700)
    # SRTM terrain without its tiles: the reference's abort code 1001
    assert model(lib, write_params(tmp_path, useTerrain=1, SRTM90=1, TerrainDir=str(tmp_path / "none"),
                                   lonStart=100, latStart=35, blockX=1, blockY=1), 20, 18, 16)[0] == 1001


def test_gauss_hill_grid(lib, tmp_path):
    path = write_params(tmp_path, gauss_hill=1)
    rc, c, t, m = model(lib, path, 20, 18, 16)
    assert rc == 0
    depth = 100.0 * 16
    # hill of height 0.2 depth at the centre; bottom interior layer at -depth;
    # top interior layer on the surface; uniform spacing in each column
    assert t[H + 9, H + 10] == pytest.approx(0.2 * depth, rel=1e-6)
    x = c[0, 0]
    y = c[1, 0]
    a = 0.1 * depth
    assert np.allclose(t, 0.2 * depth * np.exp(-0.5 * (x * x + y * y) / (a * a)), rtol=1e-5, atol=1e-3)
    assert np.allclose(c[2, H], -depth, atol=1e-3)
    assert np.allclose(c[2, H + 15], t, atol=1e-2)
    dz = np.diff(c[2], axis=0)
    assert np.allclose(dz, dz[:1], rtol=1e-4)


def test_layered_medium_follows_the_surface(lib, tmp_path):
    lay = tmp_path / "layers.bin"
    # layers: tops at 0, 0.35, 0.75 km; Vs/Vp/rho in km/s, g/cm^3
    rows = [(0.0, 1.0, 2.0, 2.0), (0.35, 2.0, 3.5, 2.4), (0.75, 3.0, 5.2, 2.7)]
    with open(lay, "wb") as f:
        f.write(np.int32(len(rows)).tobytes())
        f.write(np.array(rows, np.float32).tobytes())
    for hill in (0, 1):
        path = write_params(tmp_path, gauss_hill=hill, useMedium=1, LayeredModel=1, LayeredFileName=str(lay))
        rc, c, t, m = model(lib, path, 20, 18, 16)
        assert rc == 0
        depth_below = (t[None] - c[2])  # depth below the local surface
        vs = m[0]
        # above the first interface (with the 1 m tolerance) layer 0, between, layer 1, below, layer 2
        assert np.all(vs[depth_below < 350 - 1.5] == 1000)
        assert np.all(vs[(depth_below > 350 + 1.5) & (depth_below < 750 - 1.5)] == 2000)
        assert np.all(vs[depth_below > 750 + 1.5] == 3000)
        assert np.all(m[1][vs == 3000] == 5200) and np.all(m[2][vs == 2000] == 2400)


# ---------------------------------------------------------------------------
# projection
# ---------------------------------------------------------------------------
def aeqd(lib, lon0, lat0, lon, lat, a=6378137.0, f=1 / 298.257223563):
    xy = (ctypes.c_double * 2)()
    assert lib.h_aeqd(lon0, lat0, lon, lat, a, f, xy) == 0
    return xy[0], xy[1]


def test_aeqd_sphere_matches_closed_form(lib):
    # f = 0: the spherical azimuthal equidistant projection (Snyder 1987, eq. 25-2..25-4)
    R = 6371000.0
    lon0, lat0 = 101.25, 37.8
    for lon, lat in [(101.9, 38.3), (100.2, 37.1), (101.25, 39.0), (103.0, 37.8), (99.0, 35.5)]:
        p0, l0, p, l = map(math.radians, (lat0, lon0, lat, lon))
        cosc = math.sin(p0) * math.sin(p) + math.cos(p0) * math.cos(p) * math.cos(l - l0)
        cc = math.acos(cosc)
        k = cc / math.sin(cc)
        xe = R * k * math.cos(p) * math.sin(l - l0)
        ye = R * k * (math.cos(p0) * math.sin(p) - math.sin(p0) * math.cos(p) * math.cos(l - l0))
        x, y = aeqd(lib, lon0, lat0, lon, lat, a=R, f=0.0)
        assert x == pytest.approx(xe, abs=1e-6) and y == pytest.approx(ye, abs=1e-6)


def test_aeqd_ellipsoid_meridian_equator_and_symmetry(lib):
    a, f = 6378137.0, 1 / 298.257223563
    e2 = f * (2 - f)
    # central meridian: x = 0, y = meridian arc length (Simpson on M(phi))
    lat0, lat = 37.8, 38.9
    n = 2000
    ph = np.linspace(math.radians(lat0), math.radians(lat), n + 1)
    M = a * (1 - e2) / (1 - e2 * np.sin(ph) ** 2) ** 1.5
    h = (ph[-1] - ph[0]) / n
    arc = h / 3 * (M[0] + M[-1] + 4 * M[1:-1:2].sum() + 2 * M[2:-1:2].sum())
    x, y = aeqd(lib, 101.0, lat0, 101.0, lat)
    assert abs(x) < 1e-6 and y == pytest.approx(arc, abs=1e-4)
    # equator from an equatorial centre: the geodesic is the equator, x = a dlon
    x, y = aeqd(lib, 10.0, 0.0, 11.5, 0.0)
    assert x == pytest.approx(a * math.radians(1.5), abs=1e-4) and abs(y) < 1e-6
    # mirror symmetry about the central meridian
    x1, y1 = aeqd(lib, 101.0, 37.0, 101.7, 36.2)
    x2, y2 = aeqd(lib, 101.0, 37.0, 100.3, 36.2)
    assert x1 == pytest.approx(-x2, abs=1e-6) and y1 == pytest.approx(y2, abs=1e-6)


def test_aeqd_published_geodesic(lib):
    # Flinders Peak -> Buninyong (Geoscience Australia geodesy test line, GRS80):
    # s = 54972.271 m, forward azimuth 306 52' 05.37"
    def dms(d, m, s):
        return math.copysign(abs(d) + m / 60 + s / 3600, d)
    lat1, lon1 = dms(-37, 57, 3.72030), dms(144, 25, 29.52440)
    lat2, lon2 = dms(-37, 39, 10.15610), dms(143, 55, 35.38390)
    x, y = aeqd(lib, lon1, lat1, lon2, lat2, a=6378137.0, f=1 / 298.257222101)
    assert math.hypot(x, y) == pytest.approx(54972.271, abs=2e-3)
    az = math.degrees(math.atan2(x, y)) % 360
    assert az == pytest.approx(dms(306, 52, 5.37), abs=1e-5)


# ---------------------------------------------------------------------------
# sources
# ---------------------------------------------------------------------------
def aki_richards(strike, dip, rake):
    """Moment tensor n d^T + d n^T in (north, east, down) from the fault normal and
    slip vectors (Aki & Richards 2002, box 4.4), in the reference's axis order
    (x east, y north, z up): (M_EE, M_NN, M_UU, M_EN, M_EU, M_NU)."""
    s, d, r = strike, dip, rake
    n = np.array([-math.sin(d) * math.sin(s), math.sin(d) * math.cos(s), -math.cos(d)])
    u = np.array([math.cos(r) * math.cos(s) + math.cos(d) * math.sin(r) * math.sin(s),
                  math.cos(r) * math.sin(s) - math.cos(d) * math.sin(r) * math.cos(s),
                  -math.sin(r) * math.sin(d)])
    M = np.outer(n, u) + np.outer(u, n)
    # NED -> ENU: E = 1, N = 0, U = -D
    return np.array([M[1, 1], M[0, 0], M[2, 2], M[0, 1], -M[1, 2], -M[0, 2]])


def test_double_couple_matches_fault_vectors(lib):
    rng = np.random.default_rng(3)
    for _ in range(20):
        s, d, r = rng.uniform(0, 2 * math.pi), rng.uniform(0.05, math.pi / 2), rng.uniform(-math.pi, math.pi)
        m = (ctypes.c_float * 6)()
        lib.h_double_couple(s, d, r, m)
        np.testing.assert_allclose(np.array(m[:]), aki_richards(np.float32(s), np.float32(d), np.float32(r)),
                                   atol=2e-6)


def single(lib, path, nsteps, jac, mu):
    cells = np.zeros(343, np.int64)
    rate = np.zeros(343 * (nsteps + 1) * 6, np.float32)
    n, nt, dt = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
    rc = lib.h_single_source(path.encode(), nsteps, P(jac, ctypes.c_double), P(mu, ctypes.c_double),
                             P(cells, ctypes.c_longlong), P(rate, ctypes.c_float), ctypes.byref(n), ctypes.byref(nt),
                             ctypes.byref(dt))
    return rc, cells[:n.value], rate[:n.value * nt.value * 6].reshape(n.value, nt.value, 6), dt.value


def test_single_ricker_source_table(lib, tmp_path):
    path = write_params(tmp_path)
    NX, NY, NZ = 26, 24, 22
    jac = np.ones((NZ, NY, NX))
    mu = np.ones((NZ, NY, NX))
    rc, cells, rate, dt = single(lib, path, 30, jac, mu)
    assert rc == 0 and len(cells) == 343 and rate.shape == (343, 31, 6)
    assert dt == pytest.approx(np.float32(0.01))
    # Gaussian weights: normalised so that they sum to one over the 7^3 cells
    w = rate[:, :, 0].sum(axis=0)
    centre = int(np.argmax(np.abs(rate[:, 10, 0])))
    c0 = cells[centre]
    assert c0 == (H + 10) + NX * ((H + 9) + NY * (H + 8))
    s_t = w  # sum of weights times the time function
    peak = rate[centre, :, 0]
    ratio = s_t[np.abs(peak) > 1e-3 * np.abs(peak).max()] / peak[np.abs(peak) > 1e-3 * np.abs(peak).max()]
    g1 = np.exp(-(np.arange(-3, 4) / 1.5) ** 2) / (1.5 * 1.772453850905516)
    # sum over the cells / centre weight = (sum of the 1-D weights / centre weight)^3
    assert np.allclose(ratio, (g1.sum() / g1[3]) ** 3, rtol=1e-5)
    assert (g1.sum() ** 3) / 0.998125703461425 == pytest.approx(1.0, abs=1e-6)
    # isotropic: Txx = Tyy = Tzz, no shear
    assert np.all(rate[:, :, 0] == rate[:, :, 1]) and np.all(rate[:, :, 0] == rate[:, :, 2])
    assert np.all(rate[:, :, 3:] == 0)
    # time function: odd about t = 1/fc = 0.5 s (steps 40 and 60 of dt = 0.01 around step 50)
    rc, cells, rate, dt = single(lib, path, 100, jac, mu)
    tf = rate[centre, :, 0].astype(np.float64)
    assert tf[50] == pytest.approx(0.0, abs=1e-3 * np.abs(tf).max())
    for q in (5, 10, 20):
        assert tf[50 + q] == pytest.approx(-tf[50 - q], rel=1e-3)
    # J and mu enter as J / (J_centre mu): the library multiplies back by mu / J
    jac2 = jac.copy()
    mu2 = mu.copy()
    jac2[:] = 2.0
    mu2[H + 8, H + 9, H + 11] = 4.0
    rc, cells2, rate2, _ = single(lib, path, 100, jac2, mu2)
    k = int(np.where(cells2 == (H + 11) + NX * ((H + 9) + NY * (H + 8)))[0][0])
    np.testing.assert_allclose(rate2[k], rate[k] / 4.0, rtol=1e-6)


def test_single_double_couple_source_table(lib, tmp_path):
    path = write_params(tmp_path, **{"useSingleSource(ricker)": 0, "useSingleSource(double_couple)": 1,
                                     "strike": 30.0, "dip": 60.0, "rake": 90.0, "Mw": 4.0, "duration": 0.4})
    NX, NY, NZ = 26, 24, 22
    rc, cells, rate, dt = single(lib, path, 60, np.ones((NZ, NY, NX)), np.ones((NZ, NY, NX)))
    assert rc == 0
    m = aki_richards(*[np.float32(v * np.float32(3.141592657) / 180.0) for v in (30.0, 60.0, 90.0)])
    centre = int(np.argmax(np.abs(rate[:, :, 0]).max(axis=1)))
    t = int(np.argmax(np.abs(rate[centre, :, 0])))
    # radiation pattern, sign: W -= m * rate with rate = -g amp m0 / J, i.e. table = -m g amp m0
    np.testing.assert_allclose(rate[centre, t] / rate[centre, t, 0] * m[0], m, rtol=2e-4, atol=2e-6)
    m0 = 1.5 * 10 ** (4.0 + 6)
    g1 = np.exp(-(np.arange(-3, 4) / 1.5) ** 2) / (1.5 * 1.772453850905516)
    amp = g1[3] ** 3 / 0.998125703461425
    tt = np.float32(t) * np.float32(0.01)
    gt = math.exp(-(tt - 1.2) ** 2 / 0.16) / (math.sqrt(math.pi) * 0.4)
    assert rate[centre, t, 0] == pytest.approx(-m[0] * gt * amp * m0, rel=1e-4)
    # both single sources at once: refused (modelChecking code 180)
    path = write_params(tmp_path, **{"useSingleSource(double_couple)": 1})
    assert single(lib, path, 10, np.ones((NZ, NY, NX)), np.ones((NZ, NY, NX)))[0] == 180


def write_fault(path, lon, lat, z, area, strike, dip, rake, rate, dt):
    npts, nt = rate.shape
    with open(path, "wb") as f:
        f.write(np.array([npts, nt], np.int32).tobytes())
        f.write(np.float32(dt).tobytes())
        for p in range(npts):
            f.write(np.array([lon[p], lat[p], z[p], area[p], strike[p], dip[p]], np.float32).tobytes())
            f.write(np.asarray(rake[p], np.float32).tobytes())
            f.write(np.asarray(rate[p], np.float32).tobytes())


def test_fault_file_source_table(lib, tmp_path):
    nx, ny, nz, dh = 40, 36, 30, 200.0
    lon0, lat0 = 101.246376, 37.7942405
    # points: 2 at the same spot (merged by the library), one off the grid
    off = [(0.0, 0.0, -1000.0), (0.0, 0.0, -1000.0), (0.011, -0.006, -2500.0), (0.02, 0.013, -4000.0),
           (0.5, 0.0, -3000.0)]
    lon = np.array([lon0 + o[0] for o in off])
    lat = np.array([lat0 + o[1] for o in off])
    z = np.array([o[2] for o in off])
    nt = 7
    rng = np.random.default_rng(11)
    rate = rng.uniform(0.1, 1.0, (len(off), nt))
    rake = rng.uniform(-30, 30, (len(off), nt))
    strike = np.array([103.0, 103.0, 80.0, 120.0, 90.0])
    dip = np.array([88.0, 88.0, 70.0, 60.0, 45.0])
    area = np.array([1e6, 2e6, 1.5e6, 1e6, 1e6])
    os.makedirs(tmp_path / "src", exist_ok=True)
    write_fault(tmp_path / "src" / "f.bin", lon, lat, z, area, strike, dip, rake, rate, 0.05)
    path = write_params(tmp_path, NX=nx, NY=ny, NZ=nz, DH=dh, centerX=20, centerY=18, centerLongitude=lon0,
                        centerLatitude=lat0, useMultiSource=1, sourceDir=str(tmp_path / "src"), sourceFile="f.bin",
                        degree2radian=1, **{"useSingleSource(ricker)": 0})
    cells = np.zeros(16, np.int64)
    rt = np.zeros(16 * nt * 6, np.float32)
    xyz = np.zeros(3 * len(off), np.float32)
    n, ntv, dtv = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
    rc = lib.h_fault_source(path.encode(), P(cells, ctypes.c_longlong), P(rt, ctypes.c_float), ctypes.byref(n),
                            ctypes.byref(ntv), ctypes.byref(dtv), P(xyz, ctypes.c_float), 16)
    assert rc == 0, lib.h_last_msg()
    assert n.value == 4 and ntv.value == nt and dtv.value == pytest.approx(0.05)
    NX, NY = nx + 2 * H, ny + 2 * H
    sc = float(np.float32(3.141592657)) / 180.0 / (math.pi / 180.0)
    for q in range(4):
        # centre: the float parameters printed with %lf into the PROJ string
        c0 = [float("%f" % float(np.float32(v))) for v in (lon0, lat0)]
        x, y = aeqd(lib, c0[0], c0[1], float(np.float32(lon[q])) * sc, float(np.float32(lat[q])) * sc)
        i = int(np.float32(x) / np.float32(dh) + 0.5) + 20 + H
        j = int(np.float32(y) / np.float32(dh) + 0.5) + 18 + H
        # flat grid: the shallowest point (z = -1000) sits at 0 after the shift, then
        # DZ (NZ-1) - Depth = -DH lifts every point by -DH; cell centres at
        # z_k = (k - H + 1 - NZ) DH with half-spacing intervals
        zz = (z[q] - z.max()) - dh
        k = int(math.floor(zz / dh + 0.5)) + nz - 1 + H
        assert cells[q] == i + NX * (j + NY * k), q
        assert xyz[q] == pytest.approx(x, abs=0.01) and xyz[len(off) + q] == pytest.approx(y, abs=0.01)
        s, d = np.float32(strike[q] * sc * math.pi / 180), np.float32(dip[q] * sc * math.pi / 180)
        for t in (0, 3, 6):
            r = np.float32(np.float32(rake[q, t]) * sc * math.pi / 180)
            m = aki_richards(s, d, r)
            np.testing.assert_allclose(rt.reshape(16, nt, 6)[q, t],
                                       m * np.float32(area[q]) * np.float32(rate[q, t]), rtol=3e-5, atol=1e-2)
    assert cells[0] == cells[1]


# ---------------------------------------------------------------------------
# writers, rank split as the reference's ranks write
# ---------------------------------------------------------------------------
def spans(N, P):
    n, res = divmod(N, P)
    out = []
    for r in range(P):
        out.append((r * (n + 1), n + 1) if r < res else (res * (n + 1) + (r - res) * n, n))
    return out


def test_writers_split_over_ranks(lib, tmp_path):
    nx, ny, nz = 21, 17, 13
    out = tmp_path / "out"
    os.makedirs(out)
    path = write_params(tmp_path, NX=nx, NY=ny, NZ=nz, PX=2, PY=3, PZ=2, sliceX=5, sliceY=12, sliceZ=7)
    rng = np.random.default_rng(5)
    # stations
    gijk = np.array([[1, 1, 12], [15, 2, 12], [3, 16, 0], [20, 16, 12], [10, 8, 6]], np.int32)
    NT = 11
    rec = rng.standard_normal((len(gijk), NT, 9)).astype(np.float32)
    assert lib.h_write_stations(path.encode(), P(gijk, ctypes.c_int), len(gijk), NT, P(rec, ctypes.c_float)) == 0
    sx, sy, sz = spans(nx, 2), spans(ny, 3), spans(nz, 2)
    seen = []
    for X in range(2):
        for Y in range(3):
            for Z in range(2):
                inside = [s for s, (i, j, k) in enumerate(gijk)
                          if sx[X][0] <= i < sum(sx[X]) and sy[Y][0] <= j < sum(sy[Y]) and sz[Z][0] <= k < sum(sz[Z])]
                f = out / f"station_mpi_{X}_{Y}_{Z}.bin"
                assert f.exists() == bool(inside)
                if inside:
                    raw = np.fromfile(f, np.int32, count=3 * len(inside)).reshape(-1, 3)
                    np.testing.assert_array_equal(raw, gijk[inside])
                    wave = np.fromfile(f, np.float32, offset=12 * len(inside)).reshape(len(inside), NT, 9)
                    np.testing.assert_array_equal(wave, rec[inside])
                    seen += inside
    assert sorted(seen) == list(range(len(gijk)))
    # PGV: surface ranks only, reassembled as plotPGV.py does
    pgv = rng.standard_normal((ny, nx, 2)).astype(np.float32)
    assert lib.h_write_pgv(path.encode(), P(pgv, ctypes.c_float)) == 0
    full = np.zeros((ny, nx, 2), np.float32)
    for X in range(2):
        for Y in range(3):
            assert not (out / f"PGV_Z_mpi_{X}_{Y}_0.bin").exists()
            d = np.fromfile(out / f"PGV_Z_mpi_{X}_{Y}_1.bin", np.float32).reshape(sy[Y][1], sx[X][1], 2)
            full[sy[Y][0]:sum(sy[Y]), sx[X][0]:sum(sx[X])] = d
    np.testing.assert_array_equal(full, pgv)
    # slices: X slice [nz][ny], Y slice [nz][nx], Z slice [ny][nx] per rank holding the plane
    vol = rng.standard_normal((nz, ny, nx)).astype(np.float32)
    name = str(out / "Vx_0")
    assert lib.h_write_slices(path.encode(), name.encode(), P(vol, ctypes.c_float)) == 0
    xs = np.zeros((nz, ny), np.float32)
    ys = np.zeros((nz, nx), np.float32)
    zs = np.zeros((ny, nx), np.float32)
    nfiles = 0
    for X in range(2):
        for Y in range(3):
            for Z in range(2):
                fx = f"{name}_X_mpi_{X}_{Y}_{Z}.bin"
                if os.path.exists(fx):
                    nfiles += 1
                    assert sx[X][0] <= 5 < sum(sx[X])
                    xs[sz[Z][0]:sum(sz[Z]), sy[Y][0]:sum(sy[Y])] = np.fromfile(fx, np.float32).reshape(sz[Z][1], sy[Y][1])
                fy = f"{name}_Y_mpi_{X}_{Y}_{Z}.bin"
                if os.path.exists(fy):
                    nfiles += 1
                    ys[sz[Z][0]:sum(sz[Z]), sx[X][0]:sum(sx[X])] = np.fromfile(fy, np.float32).reshape(sz[Z][1], sx[X][1])
                fz = f"{name}_Z_mpi_{X}_{Y}_{Z}.bin"
                if os.path.exists(fz):
                    nfiles += 1
                    zs[sy[Y][0]:sum(sy[Y]), sx[X][0]:sum(sx[X])] = np.fromfile(fz, np.float32).reshape(sy[Y][1], sx[X][1])
    assert nfiles == 6 + 4 + 6
    np.testing.assert_array_equal(xs, vol[:, :, 5])
    np.testing.assert_array_equal(ys, vol[:, 12, :])
    np.testing.assert_array_equal(zs, vol[7])


def test_source_file_writer_merges_and_uses_rank_indices(lib, tmp_path):
    nx, ny, nz = 12, 10, 8
    out = tmp_path / "out"
    os.makedirs(out)
    path = write_params(tmp_path, NX=nx, NY=ny, NZ=nz, PX=2, PY=1, PZ=1)
    NX, NY = nx + 2 * H, ny + 2 * H
    cells = np.array([3 + 2 + NX * (H + 4 + NY * (H + 5)), 3 + 2 + NX * (H + 4 + NY * (H + 5)),
                      3 + 9 + NX * (H + 1 + NY * (H + 2))], np.int64)
    nt = 4
    rate = np.arange(3 * nt * 6, dtype=np.float32).reshape(3, nt, 6)
    assert lib.h_write_source_files(path.encode(), P(cells, ctypes.c_longlong), P(rate, ctypes.c_float), 3, nt,
                                    ctypes.c_float(0.25)) == 0
    for X, (i, j, k, rows) in enumerate([(2, 4, 5, [0, 1]), (9 - 6, 1, 2, [2])]):
        raw = (out / f"source_mpi_{X}_0_0.bin").read_bytes()
        n, ntf = np.frombuffer(raw[:8], np.int32)
        assert (n, ntf) == (1, nt) and np.frombuffer(raw[8:12], np.float32)[0] == 0.25
        idx = np.frombuffer(raw[12:20], np.int64)[0]
        lnx, lny = 6 + 2 * H, ny + 2 * H
        assert idx == (i + H) + lnx * ((j + H) + lny * (k + H))
        data = np.frombuffer(raw[20:], np.float32).reshape(6, nt)
        np.testing.assert_array_equal(data, rate[rows].sum(axis=0).T)


# ---------------------------------------------------------------------------
# the driver's error exits (before any GPU work)
# ---------------------------------------------------------------------------
@pytest.mark.skipif(not os.path.exists(RUNNER), reason="cgfdm_run not built")
def test_runner_error_codes(tmp_path):
    def run(**kw):
        d = tmp_path / f"r{len(os.listdir(tmp_path))}"
        os.makedirs(d)
        if kw.pop("_no_params", False) is False:
            write_params(d, **kw)
        return subprocess.run([RUNNER, "--quiet"], cwd=d, capture_output=True, text=True, timeout=120)
    r = run(_no_params=True)
    assert r.returncode == 100 and "params.json" in r.stderr
    r = run(**{"useSingleSource(double_couple)": 1})
    assert r.returncode == 180
    r = run(**{"useSingleSource(ricker)": 0})
    assert r.returncode == 180 and "source" in r.stderr
    r = run(useTerrain=1, SRTM90=1, TerrainDir="none", lonStart=100, latStart=35, blockX=1, blockY=1)
    assert r.returncode == 1001 % 256 and "no such file" in r.stderr


# ---------------------------------------------------------------------------
# SRTM terrain
# ---------------------------------------------------------------------------
def aeqd_inv(lib, lon0, lat0, x, y, a=6378137.0, f=1 / 298.257223563):
    ll = (ctypes.c_double * 2)()
    assert lib.h_aeqd_inv(lon0, lat0, x, y, a, f, ll) == 0
    return ll[0], ll[1]


def test_aeqd_inverse_round_trip_and_published_direct(lib):
    lib.h_aeqd_inv.argtypes = [ctypes.c_double] * 6 + [ctypes.POINTER(ctypes.c_double)]
    for x, y in [(0.0, 1000.0), (25000.0, -13000.0), (-80000.0, 40000.0), (3.0, 0.5)]:
        lon, lat = aeqd_inv(lib, 101.25, 37.79, x, y)
        x2, y2 = aeqd(lib, 101.25, 37.79, lon, lat)
        assert x2 == pytest.approx(x, abs=1e-6) and y2 == pytest.approx(y, abs=1e-6)
    # Flinders Peak, azimuth 306 52' 05.37", 54972.271 m -> Buninyong (GRS80)
    def dms(d, m, s):
        return math.copysign(abs(d) + m / 60 + s / 3600, d)
    az = math.radians(dms(306, 52, 5.37))
    s = 54972.271
    lon, lat = aeqd_inv(lib, dms(144, 25, 29.52440), dms(-37, 57, 3.72030), s * math.sin(az), s * math.cos(az),
                        f=1 / 298.257222101)
    assert lat == pytest.approx(dms(-37, 39, 10.15610), abs=2e-8)
    assert lon == pytest.approx(dms(143, 55, 35.38390), abs=2e-8)


def test_srtm_terrain_from_a_ramp_tile(lib, tmp_path):
    lib.h_aeqd_inv.argtypes = [ctypes.c_double] * 6 + [ctypes.POINTER(ctypes.c_double)]
    P = 6000
    J = np.arange(P, dtype=np.float64)
    ramp = (100.0 + 2000.0 * J / (P - 1)).astype(np.float32)     # rises with the tile row (latitude)
    tile = np.repeat(ramp[:, None], P, axis=1)                  # [row = lat][col = lon]
    tile[P // 2 + 40, P // 2 + 30] += 5000.0                     # one spike under the grid
    tdir = tmp_path / "tiles"
    os.makedirs(tdir)
    tile.tofile(tdir / "srtm_35N100E.bin")
    nx, ny, nz = 30, 26, 12
    path = write_params(tmp_path, NX=nx, NY=ny, NZ=nz, DH=90.0, centerX=15, centerY=13, useTerrain=1, SRTM90=1,
                        TerrainDir=str(tdir), lonStart=100, latStart=35, blockX=1, blockY=1,
                        centerLongitude=102.5, centerLatitude=37.5, **{"Depth(km)": 2.0})
    rc, c, t, m = model(lib, path, nx, ny, nz)
    assert rc == 0, lib.h_last_msg()
    # closed form: the reference maps latitude to the fractional tile row with a spacing of
    # 5/6000 degree, the tile holds a linear ramp in the row, bilinear interpolation and
    # the symmetric Gaussian smoothings (weights summing to F each) keep a linear field
    g1 = np.exp(-((np.arange(11) - 5) / 2.5) ** 2) / (2.5 * 1.772453850905516)
    F = g1.sum() ** 2 / 0.996768
    refdeg = math.pi / float(np.float32(3.141592657))
    worst = 0.0
    for j in range(0, ny + 6, 5):
        for i in range(0, nx + 6, 7):
            x = float(np.float32((i - 3) * 90.0) - np.float32(15 * 90.0))
            y = float(np.float32((j - 3) * 90.0) - np.float32(13 * 90.0))
            lon, lat = aeqd_inv(lib, 102.5, 37.5, x, y)
            row = (lat * refdeg - 35.0) / (5.0 / P)
            expect = F * F * (100.0 + 2000.0 * row / (P - 1))
            worst = max(worst, abs(t[j, i] - expect))
    assert worst < 0.05, worst
    # the spike is gone (despiked where the central gradient exceeds 100 m, then smoothed)
    assert t.max() < 100.0 + 2000.0 * (37.6 - 35.0) / 5.0 * 1.01
    # terrain-following grid from -Depth to the surface
    assert np.allclose(c[2, 3], -2000.0, atol=1e-3)
    assert np.allclose(c[2, 3 + nz - 1], t, atol=1e-2)
    # a grid reaching beyond the tile is refused with 1001
    path = write_params(tmp_path, NX=nx, NY=ny, NZ=nz, DH=90.0, centerX=15, centerY=13, useTerrain=1, SRTM90=1,
                        TerrainDir=str(tdir), lonStart=100, latStart=35, blockX=1, blockY=1,
                        centerLongitude=100.001, centerLatitude=37.5, **{"Depth(km)": 2.0})
    assert model(lib, path, nx, ny, nz)[0] == 1001


def test_shen_model_medium(lib, tmp_path):
    lib.h_aeqd_inv.argtypes = [ctypes.c_double] * 6 + [ctypes.POINTER(ctypes.c_double)]
    mdir = tmp_path / "shen"
    os.makedirs(mdir)
    nlayer = int((1.6e3 + 3e4) / 100.0)
    for a in range(3):
        for b in range(3):
            lon, lat = 101.0 + 0.5 * a, 37.0 + 0.5 * b
            K = np.arange(nlayer + 5)
            vs = 2.0 + 0.01 * K + 0.5 * (lon - 101.0)          # km/s, deeper is faster, eastward faster
            rows = np.stack([K * 0.1, vs, np.full_like(vs, 0.01)], axis=1)
            np.savetxt(mdir / f"{lon:.1f}_{lat:.1f}.mod", rows, fmt="%.6f")
    nx, ny, nz = 20, 18, 16
    path = write_params(tmp_path, NX=nx, NY=ny, NZ=nz, DH=100.0, centerX=10, centerY=9, useMedium=1, ShenModel=1,
                        MLonStart=101.0, MLatStart=37.0, MLonEnd=102.0, MLatEnd=38.0, MLonStep=0.5, MLatStep=0.5,
                        MVeticalStep=100.0, MediumDir=str(mdir), centerLongitude=101.5, centerLatitude=37.5)
    rc, c, t, m = model(lib, path, nx, ny, nz)
    assert rc == 0, lib.h_last_msg()
    vs, vp, rho = m[0].astype(np.float64), m[1].astype(np.float64), m[2].astype(np.float64)
    # one cell deeper is one 100 m layer deeper: +10 m/s (below the clamped top layers)
    d = vs[3:12] - vs[4:13]
    assert np.allclose(d, 10.0, atol=0.05)
    # laterally: 500 m/s per degree of longitude, through the surface longitude of the column
    refdeg = math.pi / float(np.float32(3.141592657))
    for (i, j) in [(3, 3), (12, 7), (20, 15)]:
        x = float(np.float32((i - 3) * 100.0) - np.float32(10 * 100.0))
        y = float(np.float32((j - 3) * 100.0) - np.float32(9 * 100.0))
        lon, lat = aeqd_inv(lib, 101.5, 37.5, x, y)
        col = vs[5, j, i] - vs[5, 3 + 9, 3 + 10]
        lon0, _ = aeqd_inv(lib, 101.5, 37.5, 0.0, 0.0)
        assert col == pytest.approx(500.0 * (lon - lon0) * refdeg, abs=0.05)
    # Vp and rho from Vs by Brocher's (2005) regressions (Vp(Vs) and Nafe-Drake rho(Vp), km/s, g/cm^3)
    VS = vs / 1e3
    VP = 0.9409 + 2.0947 * VS - 0.8206 * VS ** 2 + 0.2683 * VS ** 3 - 0.0251 * VS ** 4
    RHO = 1.6612 * VP - 0.4721 * VP ** 2 + 0.0671 * VP ** 3 - 0.0043 * VP ** 4 + 0.000106 * VP ** 5
    np.testing.assert_allclose(vp, VP * 1e3, rtol=2e-6)
    np.testing.assert_allclose(rho, RHO * 1e3, rtol=2e-6)
    # a model that does not cover the grid is refused with 1001
    path = write_params(tmp_path, NX=nx, NY=ny, NZ=nz, DH=100.0, centerX=10, centerY=9, useMedium=1, ShenModel=1,
                        MLonStart=101.0, MLatStart=37.0, MLonEnd=102.0, MLatEnd=38.0, MLonStep=0.5, MLatStep=0.5,
                        MVeticalStep=100.0, MediumDir=str(mdir), centerLongitude=101.001, centerLatitude=37.5)
    assert model(lib, path, nx, ny, nz)[0] == 1001


def test_crust1_medium(lib, tmp_path):
    lib.h_aeqd_inv.argtypes = [ctypes.c_double] * 6 + [ctypes.POINTER(ctypes.c_double)]
    cdir = tmp_path / "crust"
    os.makedirs(cdir)
    nlon, nlat, nl = 360, 180, 9
    tops = np.array([0.0, -0.5, -3.0, -3.0, -4.5, -6.0, -7.5, -9.0, -12.0])   # km; layer 0 is 0.5 km thin, 2 is empty
    I = np.arange(nlon)[None, :].repeat(nlat, 0).ravel()                    # longitude cell of each record
    base = 1.0 + 0.5 * np.arange(nl)
    vs = base[None, :] + 0.01 * I[:, None]                                   # eastward gradient, per layer
    vp = 1.8 * vs
    rho = 2.0 + 0.1 * np.arange(nl)[None, :] + 0 * I[:, None]
    bnds = tops[None, :] + 0 * I[:, None]
    for name, arr in (("vs", vs), ("vp", vp), ("rho", rho), ("bnds", bnds)):
        np.savetxt(cdir / f"crust1.{name}", arr, fmt="%.5f")
    nx, ny, nz = 20, 18, 40
    path = write_params(tmp_path, NX=nx, NY=ny, NZ=nz, DH=300.0, centerX=10, centerY=9, useMedium=1, Crust_1Model=1,
                        CrustDir=str(cdir), CrustLonStep=1.0, CrustLatStep=1.0, centerLongitude=101.5,
                        centerLatitude=37.5)
    rc, c, t, m = model(lib, path, nx, ny, nz)
    assert rc == 0, lib.h_last_msg()
    refdeg = math.pi / float(np.float32(3.141592657))
    H3 = 3
    for (i, j) in [(3, 3), (15, 11), (22, 20)]:
        x = float(np.float32((i - 3) * 300.0) - np.float32(10 * 300.0))
        y = float(np.float32((j - 3) * 300.0) - np.float32(9 * 300.0))
        lon, lat = aeqd_inv(lib, 101.5, 37.5, x, y)
        dI = (lon * refdeg - (-179.5)) * 0.01                                # interpolated eastward term
        for k in range(H3, H3 + nz):
            z = c[2, k, j, i] / 1e3
            # the layer whose top is the lowest one above the cell; layer 0 is thinner than 1 km,
            # so its cells (and any above the surface) take layer 1's values; the empty layer 2
            # (same top as layer 3) is never chosen
            if z > -0.5:
                L = 1
            elif z <= -12.0:
                L = 8
            else:
                L = max(q for q in range(nl - 1) if tops[q] >= z > tops[q + 1])
            assert m[0, k, j, i] == pytest.approx((base[L] + dI) * 1e3, abs=0.5), (i, j, k, z, L)
            assert m[2, k, j, i] == pytest.approx((2.0 + 0.1 * L) * 1e3, abs=0.05)
