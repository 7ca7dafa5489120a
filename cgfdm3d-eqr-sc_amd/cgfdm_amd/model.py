"""Model description and synthetic inputs for the CGFDM / ME-AWENO solvers.

A `Model` carries exactly what the reference's setup stage hands to its time
loop (reference run.cpp:28-138): curvilinear grid coordinates on the padded
grid, the medium (Vs, Vp, rho) on the padded grid, the finite-fault moment-rate
table and the station list.  The builders below follow the reference's own
constructions:

* flat grid            coord.cpp:100-135   (construct_flat_coord)
* terrain-following z  coord.cpp:205-257   (construct_terrain_coord)
* layered medium       medium.cpp:89-213   (generate_1D_medium)
* moment tensor        multiSource.cpp:450-535 (solveMomentRate)
* Ricker-type rate     singleSource.cpp:42-68  (sourceFunction / 1e9)
* CFL time step        calc_CFL.cpp:255-349 (Hmin by point-to-plane distance)

Arrays are float32, SoA, x fastest: coord[3, NZ, NY, NX] with NX = nx + 6
(HALO = 3 on every side, reference macro.h:39).
"""
from __future__ import annotations

import dataclasses
import math
from typing import Optional

import numpy as np

HALO = 3


@dataclasses.dataclass
class Model:
    nx: int
    ny: int
    nz: int
    dh: float
    dt: float
    nt: int                      # time steps the station buffers hold
    npml: int
    coord: np.ndarray            # float32 [3, NZ, NY, NX]
    medium: np.ndarray           # float32 [3, NZ, NY, NX]  (Vs, Vp, rho)
    free_surface: bool = True
    src_cells: Optional[np.ndarray] = None   # int64 [np] padded linear cell index
    src_rate: Optional[np.ndarray] = None    # float32 [np, src_nt, 6] = M_ij * area * slip-rate
    src_dt: float = 0.0
    stations: Optional[np.ndarray] = None    # int32 [ns, 3] padded (i, j, k)
    name: str = ""
    terrain: Optional[np.ndarray] = None     # float32 [NY, NX] surface elevation
    src_mt: Optional[np.ndarray] = None      # [np, 6] moment tensor * area per finite-fault point

    @property
    def shape_padded(self):
        return (self.nz + 2 * HALO, self.ny + 2 * HALO, self.nx + 2 * HALO)

    @property
    def cells(self) -> int:
        return self.nx * self.ny * self.nz

    def cell_index(self, i, j, k):
        NZ, NY, NX = self.shape_padded
        return np.int64(i) + np.int64(NX) * (np.int64(j) + np.int64(NY) * np.int64(k))


# ---------------------------------------------------------------------------
# grids
# ---------------------------------------------------------------------------
def flat_coord(nx, ny, nz, dh, center_x=0, center_y=0):
    """coord.cpp:100-135: x = (I-H)*DH - cX*DH, z = (K-H+1)*DH - NZ*DH (single rank)."""
    NX, NY, NZ = nx + 2 * HALO, ny + 2 * HALO, nz + 2 * HALO
    I = np.arange(NX, dtype=np.float64)
    J = np.arange(NY, dtype=np.float64)
    K = np.arange(NZ, dtype=np.float64)
    c = np.empty((3, NZ, NY, NX), np.float32)
    c[0] = ((I - HALO) * dh - center_x * dh)[None, None, :]
    c[1] = ((J - HALO) * dh - center_y * dh)[None, :, None]
    c[2] = ((K - HALO + 1) * dh - nz * dh)[:, None, None]
    terrain = np.zeros((NY, NX), np.float32)
    return c, terrain


def terrain_coord(nx, ny, nz, dh, terrain, depth, center_x=0, center_y=0):
    """coord.cpp:205-257: z = -|D| + (h + |D|)/(NZ-1) * (k - H), x/y as the flat grid."""
    c, _ = flat_coord(nx, ny, nz, dh, center_x, center_y)
    NZ = nz + 2 * HALO
    K = np.arange(NZ, dtype=np.float64) - HALO
    dz = (terrain.astype(np.float64) + abs(depth)) / (nz - 1)
    c[2] = (-abs(depth) + dz[None, :, :] * K[:, None, None]).astype(np.float32)
    return c


def gaussian_hills(nx, ny, dh, npml, nhills, seed, hmin=500.0, hmax=2000.0,
                   smin=3000.0, smax=15000.0, center_x=0, center_y=0):
    """Sum of seeded Gaussian hills h*exp(-r^2/(2 s^2)), centres inside the PML frame."""
    rng = np.random.default_rng(seed)
    NX, NY = nx + 2 * HALO, ny + 2 * HALO
    x = (np.arange(NX) - HALO) * dh - center_x * dh
    y = (np.arange(NY) - HALO) * dh - center_y * dh
    X, Y = np.meshgrid(x, y)
    lo_x, hi_x = x[HALO + npml], x[HALO + nx - 1 - npml]
    lo_y, hi_y = y[HALO + npml], y[HALO + ny - 1 - npml]
    t = np.zeros((NY, NX))
    for _ in range(nhills):
        h = rng.uniform(hmin, hmax)
        s = rng.uniform(smin, smax)
        cx = rng.uniform(lo_x, hi_x)
        cy = rng.uniform(lo_y, hi_y)
        t += h * np.exp(-0.5 * ((X - cx) ** 2 + (Y - cy) ** 2) / s ** 2)
    return t.astype(np.float32)


# ---------------------------------------------------------------------------
# media
# ---------------------------------------------------------------------------
def homogeneous_medium(shape, vs, vp, rho):
    m = np.empty((3,) + tuple(shape), np.float32)
    m[0], m[1], m[2] = vs, vp, rho
    return m


def layered_medium(coord, terrain, tops_m, vs, vp, rho, tolerance=1.0):
    """medium.cpp:89-213 restated: layer F occupies z < terrain - top_F - tol."""
    z = coord[2].astype(np.float64)
    surf = terrain.astype(np.float64)[None, :, :]
    idx = np.zeros(z.shape, np.int32)
    for F in range(1, len(tops_m)):
        idx += (z < surf - tops_m[F] - tolerance).astype(np.int32)
    m = np.empty((3,) + z.shape, np.float32)
    m[0] = np.asarray(vs, np.float64)[idx]
    m[1] = np.asarray(vp, np.float64)[idx]
    m[2] = np.asarray(rho, np.float64)[idx]
    return m


def basin_medium(coord, terrain, center_xy, semi_axes, inside=(400.0, 1200.0, 1900.0),
                 outside=(3464.0, 6000.0, 2700.0)):
    """Homogeneous medium with a sharp ellipsoidal low-velocity basin: a cell
    is inside when ((x-xc)/a)^2 + ((y-yc)/b)^2 + (depth/c)^2 <= 1, depth being
    measured down from the terrain surface above it."""
    x = coord[0].astype(np.float64)
    y = coord[1].astype(np.float64)
    depth = terrain.astype(np.float64)[None, :, :] - coord[2].astype(np.float64)
    a, b, c = semi_axes
    r = ((x - center_xy[0]) / a) ** 2 + ((y - center_xy[1]) / b) ** 2 + (np.maximum(depth, 0.0) / c) ** 2
    m = np.empty((3,) + x.shape, np.float32)
    ins = r <= 1.0
    for q in range(3):
        m[q] = np.where(ins, inside[q], outside[q])
    return m


def gaussian_random_field(shape, dh, corr_len, sigma, seed, clip=3.0):
    """Zero-mean Gaussian field with Gaussian autocorrelation (FFT-filtered noise)."""
    import scipy.fft as sfft
    rng = np.random.default_rng(seed)
    noise = rng.standard_normal(shape, dtype=np.float32)
    spec = sfft.rfftn(noise, workers=-1)
    ks = [np.fft.fftfreq(n, d=dh) for n in shape[:-1]] + [np.fft.rfftfreq(shape[-1], d=dh)]
    filt = np.ones(spec.shape, np.float32)
    for ax, k in enumerate(ks):
        sh = [1] * len(shape)
        sh[ax] = len(k)
        # autocorrelation exp(-r^2/a^2) <=> spectrum exp(-(pi a k)^2); amplitude = sqrt
        filt = filt * np.exp(-0.5 * (math.pi * corr_len * k) ** 2).astype(np.float32).reshape(sh)
    spec *= filt
    f = sfft.irfftn(spec, s=shape, workers=-1).astype(np.float32)
    f *= sigma / max(float(f.std()), 1e-30)
    return np.clip(f, -clip * sigma, clip * sigma).astype(np.float32)


# ---------------------------------------------------------------------------
# sources
# ---------------------------------------------------------------------------
def moment_tensor(strike_deg, dip_deg, rake_deg):
    """multiSource.cpp:507-519: (Mxx, Myy, Mzz, Mxy, Mxz, Myz), x = east, y = north, z = up."""
    d2r = math.pi / 180.0
    s, d, r = strike_deg * d2r, dip_deg * d2r, rake_deg * d2r
    M11 = -(math.sin(d) * math.cos(r) * math.sin(2 * s) + math.sin(2 * d) * math.sin(r) * math.sin(s) ** 2)
    M22 = math.sin(d) * math.cos(r) * math.sin(2 * s) - math.sin(2 * d) * math.sin(r) * math.cos(s) ** 2
    M33 = -(M11 + M22)
    M12 = math.sin(d) * math.cos(r) * math.cos(2 * s) + 0.5 * math.sin(2 * d) * math.sin(r) * math.sin(2 * s)
    M13 = -(math.cos(d) * math.cos(r) * math.cos(s) + math.cos(2 * d) * math.sin(r) * math.sin(s))
    M23 = -(math.cos(d) * math.cos(r) * math.sin(s) - math.cos(2 * d) * math.sin(r) * math.cos(s))
    return np.array([M22, M11, M33, M12, -M23, -M13], np.float64)


def ricker_rate(fc, t):
    """singleSource.cpp:42-68 time function divided by its 1e9 amplitude."""
    r = math.pi * fc * (np.asarray(t, np.float64) - 1.0 / fc)
    rr = r * r
    return r * (3.0 - 2.0 * rr) * np.exp(-rr) * math.sqrt(math.pi) * 0.5 * fc * math.pi


def triangle_rate(t, t0, rise, slip):
    """Triangular slip rate of duration `rise` starting at t0, integral = slip."""
    t = np.asarray(t, np.float64)
    half = 0.5 * rise
    peak = 2.0 * slip / rise
    u = t - t0
    out = np.where((u >= 0) & (u < half), peak * u / half, 0.0)
    out = np.where((u >= half) & (u < rise), peak * (rise - u) / half, out)
    return out


def point_source(model: Model, i, j, k, strike, dip, rake, m0, fc, nt_src=None):
    """One double couple delivered as a one-point moment-rate table (area = 1,
    slip-rate = M0 g(t) / mu) so it goes through the finite-fault path."""
    NZ, NY, NX = model.shape_padded
    mu = float(model.medium[2, k, j, i]) * float(model.medium[0, k, j, i]) ** 2
    nt_src = nt_src or model.nt + 2
    t = np.arange(nt_src) * model.dt
    g = ricker_rate(fc, t)
    M = moment_tensor(strike, dip, rake)
    rate = (M[None, :] * (m0 * g / mu)[:, None]).astype(np.float32)
    model.src_cells = np.array([model.cell_index(i, j, k)], np.int64)
    model.src_rate = rate[None]
    model.src_dt = model.dt
    return model


def nearest_cell(model: Model, x, y, z):
    """Padded (i, j, k) of the grid point next to (x, y, z) in model coordinates
    (srcCoord2Index, multiSource.cpp:173-314, for a local Cartesian point): i, j
    from x, y on the regular horizontal grid, k the interior cell of that column
    whose z is closest."""
    NZ, NY, NX = model.shape_padded
    x0 = float(model.coord[0, 0, 0, HALO])
    y0 = float(model.coord[1, 0, HALO, 0])
    i = int(round((x - x0) / model.dh)) + HALO
    j = int(round((y - y0) / model.dh)) + HALO
    i = min(max(i, HALO), NX - HALO - 1)
    j = min(max(j, HALO), NY - HALO - 1)
    col = model.coord[2, HALO:NZ - HALO, j, i].astype(np.float64)
    k = int(np.argmin(np.abs(col - z))) + HALO
    return i, j, k


def von_karman_slip(nl, nw, dl, dw, a_s, a_d, hurst, mean, seed):
    """Seeded random slip on an nw x nl subfault grid with a von Karman spectrum
    P(k) ~ (1 + k_s^2 a_s^2 + k_d^2 a_d^2)^-(H+1), shifted and scaled to the
    given mean and clipped at zero."""
    rng = np.random.default_rng(seed)
    ks = np.fft.fftfreq(nl, d=dl) * 2 * math.pi
    kd = np.fft.fftfreq(nw, d=dw) * 2 * math.pi
    KD, KS = np.meshgrid(kd, ks, indexing="ij")
    amp = (1.0 + (KS * a_s) ** 2 + (KD * a_d) ** 2) ** (-(hurst + 1.0) / 2.0)
    noise = np.fft.fft2(rng.standard_normal((nw, nl)))
    f = np.real(np.fft.ifft2(noise * amp))
    f = (f - f.mean()) / (f.std() + 1e-30)
    slip = mean * (1.0 + 0.5 * f)
    slip = np.clip(slip, 0.0, None)
    return slip * (mean / slip.mean())


def finite_fault(model: Model, x_top, y_top, depth_top, strike, dip, rake, length, width, nl, nw, slip, vr,
                 rise, hypo=(0.5, 0.5), nt_src=None):
    """Planar finite fault of nl x nw subfaults (one point each), the input of
    solveMomentRate (multiSource.cpp:450-535): rate[p, t] = M(strike, dip, rake)
    * area * slip_rate_p(t), with a triangular slip rate of duration `rise`
    starting at the rupture time distance(hypocentre) / vr and integrating to
    the subfault slip. x_top, y_top: centre of the top edge; depth below z = 0.
    Sets model.src_cells / src_rate / src_dt; returns M * area per point [np, 6]
    (for slip-rate gradients)."""
    d2r = math.pi / 180.0
    s, d = strike * d2r, dip * d2r
    u_s = np.array([math.sin(s), math.cos(s), 0.0])                       # along strike
    u_d = np.array([math.cos(s) * math.cos(d), -math.sin(s) * math.cos(d), -math.sin(d)])  # down dip
    dl, dw = length / nl, width / nw
    top = np.array([x_top, y_top, -depth_top])
    hyp = top + (hypo[0] - 0.5) * length * u_s + hypo[1] * width * u_d
    nt_src = nt_src or model.nt + 2
    t = np.arange(nt_src) * model.dt
    M = moment_tensor(strike, dip, rake)
    cells, rates, mts = [], [], []
    slip = np.broadcast_to(np.asarray(slip, np.float64), (nw, nl))
    for b in range(nw):
        for a in range(nl):
            P = top + ((a + 0.5) * dl - 0.5 * length) * u_s + (b + 0.5) * dw * u_d
            i, j, k = nearest_cell(model, P[0], P[1], P[2])
            tr = float(np.linalg.norm(P - hyp)) / vr
            sr = triangle_rate(t, tr, rise, float(slip[b, a]))
            area = dl * dw
            cells.append(model.cell_index(i, j, k))
            rates.append(M[None, :] * area * sr[:, None])
            mts.append(M * area)
    model.src_cells = np.array(cells, np.int64)
    model.src_rate = np.ascontiguousarray(np.array(rates), dtype=np.float32)
    model.src_dt = model.dt
    return np.array(mts)


# ---------------------------------------------------------------------------
# time step from the CFL estimate
# ---------------------------------------------------------------------------
def hmin_point_plane(coord):
    """calc_CFL.cpp:94-163 over the interior: min distance from a point to the
    plane through its -/+ neighbours in x, y, z (8 combinations)."""
    H = HALO
    NZ = coord.shape[1]
    best = np.inf
    chunk = 8
    for k0 in range(H, NZ - H, chunk):
        k1 = min(k0 + chunk, NZ - H)
        X = coord[:, k0 - 1:k1 + 1].astype(np.float64)   # one plane of margin each side
        P = X[:, 1:-1, H:-H, H:-H]
        for a in (-1, 1):
            for b in (-1, 1):
                for c in (-1, 1):
                    A = np.roll(X, a, axis=3)[:, 1:-1, H:-H, H:-H]
                    B = np.roll(X, b, axis=2)[:, 1:-1, H:-H, H:-H]
                    C = np.roll(X, c, axis=1)[:, 1:-1, H:-H, H:-H]
                    n = np.cross(B - A, C - A, axis=0)
                    nn = np.sqrt((n * n).sum(0))
                    d = np.abs(((P - A) * n).sum(0) / nn)
                    best = min(best, float(d.min()))
    return best


# ---------------------------------------------------------------------------
# the five benchmark configurations (scaled copies for tests)
# ---------------------------------------------------------------------------
def surface_station_grid(nx, ny, nz, npml, n_per_side):
    I = np.round(np.linspace(npml, nx - 1 - npml, n_per_side)).astype(int) + HALO
    J = np.round(np.linspace(npml, ny - 1 - npml, n_per_side)).astype(int) + HALO
    K = nz - 1 + HALO
    st = np.array([[i, j, K] for j in J for i in I], np.int32)
    return st


def config1(nt=1000, nx=128, ny=128, nz=64, dh=100.0, npml=12):
    """Plain CGFDM: flat half-space Vp 6000 / Vs 3464 / rho 2700, double couple
    (30, 60, 90, M0 1e16) at 3 km depth, 2 Hz Ricker-type rate, CFL 0.5."""
    coord, terrain = flat_coord(nx, ny, nz, dh)
    med = homogeneous_medium(coord.shape[1:], 3464.0, 6000.0, 2700.0)
    hmin = dh / math.sqrt(3.0)
    dt = 0.5 * hmin / 6000.0
    m = Model(nx, ny, nz, dh, dt, nt, npml, coord, med, True, name="config1", terrain=terrain)
    depth_cells = int(round(3000.0 / dh))
    i, j, k = nx // 2 + HALO, ny // 2 + HALO, nz - 1 - depth_cells + HALO
    point_source(m, i, j, k, 30.0, 60.0, 90.0, 1e16, 2.0)
    picks = [int(round(f * nx)) for f in (20 / 128, 42 / 128, 64 / 128, 86 / 128, 108 / 128)]
    m.stations = np.array([[a + HALO, b + HALO, nz - 1 + HALO] for b in picks for a in picks], np.int32)
    return m


def config2(nt=3000, nx=512, ny=512, nz=256, dh=100.0, npml=20, seed=1002, nhills=8,
            perturb=True, hmin_fn=None):
    """CGFDM curvilinear: 8 seeded Gaussian hills, 4-layer medium with a 5%
    Gaussian random perturbation (1 km correlation), 20-layer PML, point
    moment source at 5 km depth, 400 surface stations, PGV map."""
    depth = nz * dh
    terrain = gaussian_hills(nx, ny, dh, npml, nhills, seed)
    coord = terrain_coord(nx, ny, nz, dh, terrain, depth)
    vp = np.array([3000.0, 4500.0, 6000.0, 7000.0])
    med = layered_medium(coord, terrain, [0.0, 2000.0, 8000.0, 18000.0], vp / 1.73, vp,
                         [2300.0, 2500.0, 2700.0, 3000.0])
    if perturb:
        delta = gaussian_random_field(med.shape[1:], dh, 1000.0, 0.05, seed + 7)
        med[0] *= 1.0 + delta
        med[1] *= 1.0 + delta
        del delta
    hmin = (hmin_fn or (lambda c, md: hmin_point_plane(c)))(coord, med)
    dt = 0.5 * hmin / float(med[1].max())
    m = Model(nx, ny, nz, dh, dt, nt, npml, coord, med, True, name="config2", terrain=terrain)
    depth_cells = int(round(5000.0 / dh))
    i, j, k = nx // 2 + HALO, ny // 2 + HALO, nz - 1 - depth_cells + HALO
    point_source(m, i, j, k, 30.0, 60.0, 90.0, 1e16, 2.0)
    nside = 20 if min(nx, ny) - 2 * npml >= 20 else max(2, min(nx, ny) - 2 * npml)
    m.stations = surface_station_grid(nx, ny, nz, npml, nside)
    return m


def config3(nt=4000, nx=512, ny=512, nz=256, dh=100.0, seed=1002, nhills=8, hmin_fn=None):
    """ME-AWENO: the 512x512x256 terrain grid of config2 with a sharp
    low-velocity basin (Vs 400, Vp 1200, rho 1900 inside the ellipsoid of
    semi-axes 10 x 5 x 1 km centred on the surface at the domain centre, i.e. a
    20 x 10 x 2 km ellipsoid cut by the surface; Vs 3464, Vp 6000, rho 2700
    outside), point moment source at 5 km depth, 400 surface stations.
    Absorption is the scheme's 20-layer exponential sponge."""
    depth = nz * dh
    terrain = gaussian_hills(nx, ny, dh, 20, nhills, seed)
    coord = terrain_coord(nx, ny, nz, dh, terrain, depth)
    med = basin_medium(coord, terrain, ((nx // 2) * dh, (ny // 2) * dh), (10000.0, 5000.0, 1000.0))
    hmin = (hmin_fn or (lambda c, md: hmin_point_plane(c)))(coord, med)
    dt = 0.5 * hmin / float(med[1].max())
    m = Model(nx, ny, nz, dh, dt, nt, 0, coord, med, True, name="config3", terrain=terrain)
    depth_cells = int(round(5000.0 / dh))
    i, j, k = nx // 2 + HALO, ny // 2 + HALO, nz - 1 - depth_cells + HALO
    point_source(m, i, j, k, 30.0, 60.0, 90.0, 1e16, 2.0)
    nside = 20 if min(nx, ny) - 40 >= 20 else max(2, min(nx, ny) - 40)
    m.stations = surface_station_grid(nx, ny, nz, 20, nside)
    return m


def config4(nt=2000, nx=1024, ny=1024, nz=320, dh=50.0, npml=20, seed=1004, hmin_fn=None):
    """Mixed-precision large domain: 1024 x 1024 x 320 at h = 50 m (16 km deep,
    335.5 M cells), 8 seeded Gaussian hills, the 4-layer medium of config2 with a
    5% Gaussian random perturbation (1 km correlation), 20-layer PML; a finite
    fault of 40 x 20 subfaults on a 40 x 20 km plane (strike 30, dip 40, rake 90,
    top edge at 1 km depth, hypocentre at its centre) with seeded von Karman slip
    (a_s 10 km, a_d 5 km, H 0.75, mean 2 m), rupture at 0.8 Vs and a 1 s
    triangular slip rate; 400 surface stations on a 20 x 20 grid plus 16 buried
    stations on a 4 x 4 grid of 1 km spacing, 1 km above the hypocentre (they
    see the first arrivals within a few hundred steps). Meant to be stepped with
    16-bit coefficient storage; the FP64 run of the same kernels is its
    reference."""
    depth = nz * dh
    terrain = gaussian_hills(nx, ny, dh, npml, 8, seed)
    coord = terrain_coord(nx, ny, nz, dh, terrain, depth)
    vp = np.array([3000.0, 4500.0, 6000.0, 7000.0])
    med = layered_medium(coord, terrain, [0.0, 2000.0, 8000.0, 12000.0], vp / 1.73, vp,
                         [2300.0, 2500.0, 2700.0, 3000.0])
    delta = gaussian_random_field(med.shape[1:], dh, 1000.0, 0.05, seed + 7)
    med[0] *= 1.0 + delta
    med[1] *= 1.0 + delta
    del delta
    hmin = (hmin_fn or (lambda c, md: hmin_point_plane(c)))(coord, med)
    dt = 0.5 * hmin / float(med[1].max())
    m = Model(nx, ny, nz, dh, dt, nt, npml, coord, med, True, name="config4", terrain=terrain)
    nl, nw, length, width, dip = 40, 20, 40000.0, 20000.0, 40.0
    slip = von_karman_slip(nl, nw, length / nl, width / nw, 10000.0, 5000.0, 0.75, 2.0, seed)
    xc, yc = (nx // 2) * dh, (ny // 2) * dh
    # rupture speed from the shear speed at the hypocentre's depth range
    hyp_depth = 1000.0 + 0.5 * width * math.sin(math.radians(dip))
    i0, j0, k0 = nearest_cell(m, xc, yc, -hyp_depth)
    vs_src = float(np.median(med[0, max(k0 - 20, 0):k0 + 21, max(j0 - 20, 0):j0 + 21, max(i0 - 20, 0):i0 + 21]))
    m.src_mt = finite_fault(m, xc, yc, 1000.0, 30.0, dip, 90.0, length, width, nl, nw, slip, 0.8 * vs_src, 1.0)
    # hypocentre: centre of the plane (finite_fault's default hypo = (0.5, 0.5))
    s_, d_ = math.radians(30.0), math.radians(dip)
    u_d = np.array([math.cos(s_) * math.cos(d_), -math.sin(s_) * math.cos(d_), -math.sin(d_)])
    hyp = np.array([xc, yc, -1000.0]) + 0.5 * width * u_d
    buried = []
    for b in range(4):
        for a in range(4):
            x = hyp[0] + (a - 1.5) * 1000.0
            y = hyp[1] + (b - 1.5) * 1000.0
            buried.append(nearest_cell(m, x, y, hyp[2] + 1000.0))
    m.stations = np.concatenate([surface_station_grid(nx, ny, nz, npml, 20), np.array(buried, np.int32)])
    return m


def config5(nt=2000, nx=256, ny=256, nz=128, dh=100.0, npml=20, seed=1005, hmin_fn=None):
    """Adjoint workload: 256 x 256 x 128 at h = 100 m, 4 seeded Gaussian hills,
    4-layer medium with a 5% Gaussian random perturbation (1 km correlation),
    20-layer PML, a 16 x 8 subfault finite fault (16 x 8 km, strike 30, dip 60,
    rake 90, top at 2 km depth, uniform 1 m slip, rupture 0.8 Vs, 1 s rise),
    64 surface stations on an 8 x 8 grid."""
    depth = nz * dh
    terrain = gaussian_hills(nx, ny, dh, npml, 4, seed)
    coord = terrain_coord(nx, ny, nz, dh, terrain, depth)
    vp = np.array([3000.0, 4500.0, 6000.0, 7000.0])
    med = layered_medium(coord, terrain, [0.0, 2000.0, 6000.0, 10000.0], vp / 1.73, vp,
                         [2300.0, 2500.0, 2700.0, 3000.0])
    delta = gaussian_random_field(med.shape[1:], dh, 1000.0, 0.05, seed + 7)
    med[0] *= 1.0 + delta
    med[1] *= 1.0 + delta
    hmin = (hmin_fn or (lambda c, md: hmin_point_plane(c)))(coord, med)
    dt = 0.5 * hmin / float(med[1].max())
    m = Model(nx, ny, nz, dh, dt, nt, npml, coord, med, True, name="config5", terrain=terrain)
    xc, yc = (nx // 2) * dh, (ny // 2) * dh
    vs_src = float(np.median(med[0]))
    m.src_mt = finite_fault(m, xc, yc, 2000.0, 30.0, 60.0, 90.0, 16000.0, 8000.0, 16, 8, 1.0, 0.8 * vs_src, 1.0)
    m.stations = surface_station_grid(nx, ny, nz, npml, 8)
    return m


def small_basin_config(nx=48, ny=44, nz=20, dh=100.0, nt=12, seed=5, hills=2):
    """Small terrain grid with a sharp low-velocity basin (Vs 400 / Vp 1200 /
    rho 1900 in an ellipsoid 1.6 x 1.2 x 0.6 km under the surface centre)."""
    depth = nz * dh
    terrain = gaussian_hills(nx, ny, dh, 4, hills, seed, hmin=100.0, hmax=300.0, smin=300.0, smax=900.0)
    coord = terrain_coord(nx, ny, nz, dh, terrain, depth)
    med = basin_medium(coord, terrain, ((nx // 2) * dh, (ny // 2) * dh), (800.0, 600.0, 600.0))
    hmin = hmin_point_plane(coord)
    dt = 0.5 * hmin / float(med[1].max())
    m = Model(nx, ny, nz, dh, dt, nt, 0, coord, med, True, name="small_basin", terrain=terrain)
    i, j, k = nx // 2 + HALO + 3, ny // 2 + HALO - 2, nz // 2 + HALO - 2
    point_source(m, i, j, k, 30.0, 60.0, 90.0, 1e16, 8.0)
    m.stations = np.array([[nx // 2 + HALO, ny // 2 + HALO, nz - 1 + HALO], [HALO + 1, HALO + 2, nz - 1 + HALO],
                           [i + 2, j - 3, k], [nx + HALO - 1, ny + HALO - 1, HALO]], np.int32)
    return m


def small_config(nx=24, ny=20, nz=16, dh=100.0, npml=4, nt=12, seed=7, hills=2, layered=True,
                 free_surface=True):
    """Small curvilinear layered case used by the parity tests."""
    depth = nz * dh
    if hills:
        terrain = gaussian_hills(nx, ny, dh, npml, hills, seed, hmin=100.0, hmax=300.0,
                                 smin=300.0, smax=900.0)
    else:
        terrain = np.zeros((ny + 2 * HALO, nx + 2 * HALO), np.float32)
    coord = terrain_coord(nx, ny, nz, dh, terrain, depth)
    if layered:
        vp = np.array([3000.0, 4500.0, 6000.0])
        med = layered_medium(coord, terrain, [0.0, 400.0, 900.0], vp / 1.73, vp, [2300.0, 2500.0, 2700.0])
        delta = gaussian_random_field(med.shape[1:], dh, 300.0, 0.05, seed + 1)
        med[0] *= 1.0 + delta
        med[1] *= 1.0 + delta
    else:
        med = homogeneous_medium(coord.shape[1:], 3464.0, 6000.0, 2700.0)
    hmin = hmin_point_plane(coord)
    dt = 0.5 * hmin / float(med[1].max())
    m = Model(nx, ny, nz, dh, dt, nt, npml, coord, med, free_surface, name="small", terrain=terrain)
    i, j, k = nx // 2 + HALO, ny // 2 + HALO, nz // 2 + HALO
    point_source(m, i, j, k, 30.0, 60.0, 90.0, 1e16, 8.0)
    m.stations = np.array([[i, j, nz - 1 + HALO], [HALO + 1, HALO + 2, nz - 1 + HALO],
                           [i + 2, j - 3, k], [nx + HALO - 1, ny + HALO - 1, HALO]], np.int32)
    return m
