"""Pins of the CPU oracle against what the mathematics fixes.

The reference ships no tests or golden vectors for this path, so each oracle
function is checked against an independent closed form, an invariant or a
textbook formula, chosen so that a plausible slip (a wrong sign, a transposed
metric, a flipped F/B direction, a dropped term) fails one of them.
"""
import math

import numpy as np
import pytest

from cgfdm_amd import model as M

H = M.HALO


def affine_model(nx=14, ny=12, nz=10, dh=100.0, shear=(0.25, 0.0, -0.25, 0.125), fs=False, npml=0,
                 vs=3464.0, vp=6000.0, rho=2700.0, nt=4):
    """x = dh (i + s0 k), y = dh (j + s3 i), z = dh (k + s1 i + s2 j) on the padded grid
    (an affine map; every coordinate is exactly representable in float32)."""
    NX, NY, NZ = nx + 2 * H, ny + 2 * H, nz + 2 * H
    I = np.arange(NX)[None, None, :] - H
    J = np.arange(NY)[None, :, None] - H
    K = np.arange(NZ)[:, None, None] - H
    c = np.empty((3, NZ, NY, NX), np.float32)
    c[0] = dh * (I + shear[0] * K)
    c[1] = dh * (J + shear[3] * I)
    c[2] = dh * (K + shear[1] * I + shear[2] * J)
    med = M.homogeneous_medium((NZ, NY, NX), vs, vp, rho)
    dt = 0.5 * M.hmin_point_plane(c) / vp
    return M.Model(nx, ny, nz, dh, dt, nt, npml, c, med, fs, name="affine")


def elastic_rhs(G, lam, mu, b):
    """Velocity-stress elastodynamics for linear fields W_v = G[v] . x (textbook
    form); G may carry trailing grid axes."""
    out = [None] * 9
    out[0] = b * (G[3, 0] + G[6, 1] + G[7, 2])
    out[1] = b * (G[6, 0] + G[4, 1] + G[8, 2])
    out[2] = b * (G[7, 0] + G[8, 1] + G[5, 2])
    div = G[0, 0] + G[1, 1] + G[2, 2]
    out[3] = lam * div + 2 * mu * G[0, 0]
    out[4] = lam * div + 2 * mu * G[1, 1]
    out[5] = lam * div + 2 * mu * G[2, 2]
    out[6] = mu * (G[0, 1] + G[1, 0])
    out[7] = mu * (G[0, 2] + G[2, 0])
    out[8] = mu * (G[1, 2] + G[2, 1])
    return np.array(out)


ALL_FB = [(a, b, c) for a in (-1, 1) for b in (-1, 1) for c in (-1, 1)]


S0, S1 = -6e-5, 0.99998   # sum(a_m), sum(m a_m) of the MacCormack coefficients


def affine_A(shear):
    """d(x, y, z)/d(i, j, k) / dh of affine_model: A[a][d]."""
    return np.array([[1.0, 0.0, shear[0]], [shear[3], 1.0, 0.0], [shear[1], shear[2], 1.0]])


@pytest.mark.parametrize("shear", [(0.0, 0.0, 0.0, 0.0), (0.25, 0.125, -0.25, 0.125)])
def test_linear_fields_give_exact_elastic_rhs(oracle_mod, shear):
    """On an affine (sheared) grid, fields linear in x,y,z must give the
    elastodynamic right-hand side: exactly for the mean over the 8 F/B
    combinations, and for each combination up to the operators' consistency term
    FB_d S0 W / dh (sum of coefficients S0 = -6e-5). Pins the metric
    cofactors/Jacobian, the chain rule, the stress-velocity assembly and the
    F/B direction of every axis."""
    m = affine_model(shear=shear)
    o = oracle_mod.Oracle(m, prec=8)
    rng = np.random.default_rng(3)
    G = rng.standard_normal((9, 3))
    x = m.coord.astype(np.float64)
    W = [G[v, 0] * x[0] + G[v, 1] * x[1] + G[v, 2] * x[2] for v in range(9)]
    for v in range(9):
        o.set_field(v, W[v])
    lam = 2700.0 * (6000.0 ** 2 - 2 * 3464.0 ** 2)
    mu = 2700.0 * 3464.0 ** 2
    b = 1.0 / 2700.0
    want = elastic_rhs(G, lam, mu, b)
    Minv = np.linalg.inv(S1 * affine_A(shear))   # computed contravariant metrics, rows = xi_d
    sl = (slice(H, -H), slice(H, -H), slice(H, -H))
    mean = np.zeros((9,) + W[0][sl].shape)
    scale = abs(want).max()
    for fb in ALL_FB:
        h = o.rhs(fb, with_boundaries=False) / m.dt
        corr = np.array([sum(Minv[d, a] * fb[d] for d in range(3)) for a in range(3)])
        Gc = np.array([[corr[a] * S0 * W[v][sl] / m.dh for a in range(3)] for v in range(9)])
        extra = elastic_rhs(Gc, lam, mu, b)
        for v in range(9):
            np.testing.assert_allclose(h[v][sl], want[v] + extra[v], rtol=1e-9, atol=1e-9 * scale)
            mean[v] += h[v][sl] / 8
    for v in range(9):
        np.testing.assert_allclose(mean[v], want[v], rtol=1e-9, atol=1e-9 * scale)


def test_flat_metrics_closed_form(oracle_mod):
    """Flat grid: x_xi = sum(m a_m) = 0.99998 for the MacCormack pair, so the
    contravariant metrics are diag(1/0.99998) and J = 0.99998^3."""
    m = affine_model(shear=(0, 0, 0, 0))
    o = oracle_mod.Oracle(m, prec=8)
    s = 0.99998
    sl = (slice(H, -H), slice(H, -H), slice(H, -H))
    eye = [1, 0, 0, 0, 1, 0, 0, 0, 1]
    for c in range(9):
        np.testing.assert_allclose(o.setup(c)[sl], eye[c] / s, atol=1e-12)
    np.testing.assert_allclose(o.setup(9)[sl], s ** 3, rtol=1e-12)


def test_sheared_metrics_are_inverse_jacobian(oracle_mod):
    """Affine map: (xi_x ... zt_z) must be the inverse of d(x,y,z)/d(i,j,k) / dh."""
    shear = (0.25, 0.125, -0.25, 0.125)
    m = affine_model(shear=shear)
    o = oracle_mod.Oracle(m, prec=8)
    A = affine_A(shear) * S1
    Ainv = np.linalg.inv(A)  # rows: xi, eta, zeta ; columns: x, y, z
    p = (H + 4, H + 5, H + 6)
    got = np.array([o.setup(c)[p[2], p[1], p[0]] for c in range(9)]).reshape(3, 3)
    np.testing.assert_allclose(got, Ainv, rtol=1e-12)
    np.testing.assert_allclose(o.setup(9)[p[2], p[1], p[0]], np.linalg.det(A), rtol=1e-12)


@pytest.mark.parametrize("fb", [1, -1])
def test_maccormack_quadratic_moments(oracle_mod, fb):
    """Vx = x^2: the one-sided operators differ from 2x by FB*(S0 x^2/dh + S2 dh)/S1
    with S0 = sum a = -6e-5, S1 = sum m a = 0.99998, S2 = sum m^2 a = -0.03422;
    the forward/backward mean cancels the even moments."""
    m = affine_model(shear=(0, 0, 0, 0))
    o = oracle_mod.Oracle(m, prec=8)
    x = m.coord[0].astype(np.float64)
    o.set_field(0, x * x)
    lam = 2700.0 * (6000.0 ** 2 - 2 * 3464.0 ** 2)
    sl = (slice(H, -H), slice(H, -H), slice(H, -H))
    h = o.rhs((fb, 1, 1), with_boundaries=False)
    dvdx = h[4][sl] / (lam * m.dt)  # Tyy rate = lambda * dVx/dx
    xs = x[sl]
    S2 = -0.03422
    want = 2 * xs + fb * (S0 * xs * xs / m.dh + S2 * m.dh) / S1
    np.testing.assert_allclose(dvdx, want, rtol=1e-9, atol=1e-9)


def test_rk4_amplification(oracle_mod):
    """y' = z y through the four (W, t, m) stages gives the RK4 polynomial."""
    for z in (-0.3, 0.7, -1.9 + 0.0):
        w = np.array([1.0])
        t = np.zeros(1)
        mm = np.zeros(1)
        for irk in range(4):
            h = z * w.copy()
            oracle_mod.rk_stage(8, irk, h, w, t, mm)
        want = 1 + z + z * z / 2 + z ** 3 / 6 + z ** 4 / 24
        # the reference's weights are 0.16666667 / 0.33333333, not exact thirds
        assert abs(w[0] - want) < 5e-8 * (1 + abs(z) ** 4)


def smooth_fields(m, seed=1):
    rng = np.random.default_rng(seed)
    x, y, z = (m.coord[c].astype(np.float64) for c in range(3))
    out = []
    for v in range(9):
        a, b, c, ph = rng.uniform(0.5, 2.0, 4)
        f = np.sin(a * x / 800.0 + ph) * np.cos(b * y / 700.0) * np.cos(c * z / 900.0 + 0.3)
        out.append(f * (1.0 if v < 3 else 1e6))
    return out


@pytest.mark.parametrize("hills", [0, 2])
def test_free_surface_traction_rate_vanishes(oracle_mod, hills):
    """On the surface layer the traction-image scheme imposes d/dt (n . sigma) = 0
    with n = grad(zeta): pins the -DZ^-1 DX / -DZ^-1 DY matrices and the surface
    stress assembly on flat and curvilinear surfaces.

    The medium is a layered Poisson solid (lambda = mu): the reference builds the
    diagonal of DZ, DX, DY with chi = 2 lambda + mu (contravariant.cpp:376) where
    the traction condition has lambda + 2 mu; the two agree only for lambda = mu,
    and the oracle keeps the reference's form."""
    m = M.small_config(nx=20, ny=18, nz=12, npml=0, hills=hills, layered=True)
    m.medium[0] = (m.medium[1].astype(np.float64) / math.sqrt(3.0)).astype(np.float32)
    o = oracle_mod.Oracle(m, prec=8)
    for v, f in enumerate(smooth_fields(m)):
        o.set_field(v, f)
    k = H + m.nz - 1
    zt = [o.setup(6 + c)[k] for c in range(3)]
    for fb in ALL_FB:
        h = o.rhs(fb, with_boundaries=True)
        sxx, syy, szz, sxy, sxz, syz = (h[3 + c][k] for c in range(6))
        t1 = zt[0] * sxx + zt[1] * sxy + zt[2] * sxz
        t2 = zt[0] * sxy + zt[1] * syy + zt[2] * syz
        t3 = zt[0] * sxz + zt[1] * syz + zt[2] * szz
        sl = (slice(H, -H), slice(H, -H))
        scale = max(abs(sxx[sl]).max(), abs(syy[sl]).max())
        for t in (t1, t2, t3):
            assert abs(t[sl]).max() < 2e-6 * scale


def aki_richards_tensor(strike, dip, rake):
    """M/M0 = n d^T + d n^T from the fault normal n and slip d (north, east, down),
    returned in (east, north, up) order (Mxx, Myy, Mzz, Mxy, Mxz, Myz)."""
    s, d, r = (math.radians(a) for a in (strike, dip, rake))
    n = np.array([-math.sin(d) * math.sin(s), math.sin(d) * math.cos(s), -math.cos(d)])
    u = np.array([math.cos(r) * math.cos(s) + math.cos(d) * math.sin(r) * math.sin(s),
                  math.cos(r) * math.sin(s) - math.cos(d) * math.sin(r) * math.cos(s),
                  -math.sin(r) * math.sin(d)])
    Mned = np.outer(n, u) + np.outer(u, n)
    P = np.array([[0, 1, 0], [1, 0, 0], [0, 0, -1]], float)  # NED -> ENU
    Menu = P @ Mned @ P.T
    return np.array([Menu[0, 0], Menu[1, 1], Menu[2, 2], Menu[0, 1], Menu[0, 2], Menu[1, 2]])


@pytest.mark.parametrize("sdr", [(30, 60, 90), (103, 88, -6), (0, 90, 0), (217, 35, 141)])
def test_moment_tensor_matches_aki_richards(sdr):
    np.testing.assert_allclose(M.moment_tensor(*sdr), aki_richards_tensor(*sdr), atol=1e-12)


def test_pml_profiles(oracle_mod):
    """Profiles: beta and d grow quadratically and alpha linearly in the distance
    l into the layer, d0 = -(p+1) Cp ln R / (2L) with p = 2, Cp = 6000; outside
    the layer alpha = d = 0, beta = 1; no layer on top under a free surface."""
    npml = 6
    m = affine_model(nx=20, ny=18, nz=14, shear=(0, 0, 0, 0), fs=True, npml=npml)
    o = oracle_mod.Oracle(m, prec=8)
    dh = m.dh
    L = dh * (npml - 1)
    R = 10 ** (-(math.log10(npml) - 1) / math.log10(2) - 3)
    d0 = -3 * 6000.0 * math.log(R) / (2 * L)
    a = o.setup(40)
    binv = o.setup(41)
    dd = o.setup(42)
    for i in range(H, H + npml):
        l = (H + npml - 1 - i) * dh
        beta = 1 + 2 * (l / L) ** 2
        assert binv[i] == pytest.approx(1 / beta, rel=1e-12)
        assert dd[i] == pytest.approx(d0 * (l / L) ** 2 / beta, rel=1e-12)
        assert a[i] == pytest.approx(3.141592657 * 0.5 * (1 - l / L), rel=1e-12, abs=1e-12)
    mid = H + m.nx // 2
    assert (a[mid], binv[mid], dd[mid]) == (0.0, 1.0, 0.0)
    # mirror symmetry in x
    np.testing.assert_allclose(dd[H:-H], dd[H:-H][::-1], rtol=1e-12)
    # z: bottom layer only
    dz = o.setup(48)
    assert dz[H] > 0 and dz[-H - 1] == 0.0


def test_flat_hmin_is_dh_over_sqrt3(oracle_mod):
    m = affine_model(shear=(0, 0, 0, 0))
    o = oracle_mod.Oracle(m, prec=8)
    assert o.ranges()["hmin"] == pytest.approx(m.dh / math.sqrt(3), rel=1e-12)


def test_step_is_linear(oracle_mod):
    """Without a source the whole step (PML, surface, RK4) is linear in W."""
    m = M.small_config(nx=16, ny=14, nz=12, npml=3, hills=2)
    m.src_cells = None
    f1 = smooth_fields(m, 1)
    f2 = smooth_fields(m, 2)
    res = []
    for comb in ((1.0, 0.0), (0.0, 1.0), (1.0, -2.5)):
        o = oracle_mod.Oracle(m, prec=8)
        for v in range(9):
            g = comb[0] * f1[v] + comb[1] * f2[v]
            g[:H] = 0; g[-H:] = 0; g[:, :H] = 0; g[:, -H:] = 0; g[:, :, :H] = 0; g[:, :, -H:] = 0
            o.set_field(v, g)
        o.step(3)
        res.append(o.wavefield())
    lin = res[0] - 2.5 * res[1]
    np.testing.assert_allclose(res[2], lin, atol=1e-9 * abs(lin).max())


def test_pml_absorbs(oracle_mod):
    """A stress pulse in a full space: with CFS-PML on all faces the energy that
    remains after the wave has crossed the boundary is far below what the bare
    zero-ghost boundary leaves behind."""
    late = {}
    for npml in (0, 6):
        m = affine_model(nx=28, ny=28, nz=28, shear=(0, 0, 0, 0), fs=False, npml=npml, nt=1)
        o = oracle_mod.Oracle(m, prec=4, threads=8)
        x, y, z = (m.coord[c].astype(np.float64) for c in range(3))
        c0 = [float(np.mean(x)), float(np.mean(y)), float(np.mean(z))]
        r2 = (x - c0[0]) ** 2 + (y - c0[1]) ** 2 + (z - c0[2]) ** 2
        g = np.exp(-r2 / (2 * 200.0 ** 2))
        g[:H] = 0; g[-H:] = 0; g[:, :H] = 0; g[:, -H:] = 0; g[:, :, :H] = 0; g[:, :, -H:] = 0
        for c in (3, 4, 5):
            o.set_field(c, 1e6 * g)

        def energy():
            w = o.wavefield()
            sl = (slice(H, -H),) * 3
            return float(sum((w[v][sl] ** 2).sum() for v in range(3)))

        o.step(20)
        peak = energy()
        o.step(80)
        late[npml] = energy() / peak
    assert late[6] < 1e-3
    assert late[6] < 0.1 * late[0]
