"""Pins of the ME-AWENO (SCFDM) part of the CPU oracle.

The reference ships no tests or golden vectors for this path. Each function of
the oracle is checked against what the mathematics fixes: polynomial
exactness of the WENO interpolation and of the alternative-flux difference,
closed forms of the curvilinear metrics on affine grids, the stability
polynomial of TVD-RK3, the invariants of the characteristic free-surface
condition (zero traction, preserved outgoing characteristics, idempotence,
cubic extrapolation), Hooke's law for the strain-rate source, the Cerjan
sponge profile and an analytic plane P wave.
"""
import math

import numpy as np
import pytest

from cgfdm_amd import model as M

H = M.HALO
VS, VP, RHO = 3464.0, 6000.0, 2700.0
MU = RHO * VS * VS
LAM = RHO * (VP * VP - 2 * VS * VS)


def affine_model(nx=16, ny=14, nz=16, dh=100.0, A=None, fs=True, nt=4):
    """Padded-grid coordinates x = dh * A (i - H, j - H, k - H) (an affine map)."""
    A = np.eye(3) if A is None else np.asarray(A, np.float64)
    NX, NY, NZ = nx + 2 * H, ny + 2 * H, nz + 2 * H
    K, J, I = np.meshgrid(np.arange(NZ) - H, np.arange(NY) - H, np.arange(NX) - H, indexing="ij")
    c = np.empty((3, NZ, NY, NX), np.float32)
    for a in range(3):
        c[a] = dh * (A[a, 0] * I + A[a, 1] * J + A[a, 2] * K)
    med = M.homogeneous_medium((NZ, NY, NX), VS, VP, RHO)
    dt = 0.4 * M.hmin_point_plane(c) / VP
    return M.Model(nx, ny, nz, dh, dt, nt, 0, c, med, fs, name="affine")


A_SHEAR = [[1.0, 0.0, 0.25], [0.125, 1.0, 0.0], [0.25, -0.25, 1.0]]


# --------------------------------------------------------------------------- WENO


def test_weno_reproduces_quadratics(oracle_mod):
    """Every 3-point candidate is exact for quadratics, so any convex
    combination is: the value at x = 2.5 of a quadratic sampled at 0..4."""
    rng = np.random.default_rng(0)
    for _ in range(20):
        a, b, c = rng.standard_normal(3) * 10 ** rng.uniform(-2, 3)
        p = lambda x: a + b * x + c * x * x
        got = oracle_mod.weno([p(x) for x in range(5)])
        assert abs(got - p(2.5)) <= 1e-12 * max(1.0, abs(a) + abs(b) + abs(c) * 10)


def test_weno_small_amplitude_is_the_quartic_interpolant(oracle_mod):
    """With smoothness indicators far below the 1e-6 regulariser the weights
    are the linear ones (5/16, 10/16, 1/16): the result is the 5-point Lagrange
    interpolant (3, -20, 90, 60, -5)/128, exact for quartics."""
    rng = np.random.default_rng(1)
    for _ in range(10):
        co = rng.standard_normal(5) * 1e-9
        p = lambda x: sum(co[n] * x ** n for n in range(5))
        u = [p(x) for x in range(5)]
        got = oracle_mod.weno(u)
        assert abs(got - p(2.5)) <= 1e-6 * max(abs(q) for q in u)
        # an O(1) quartic is not reproduced (the weights leave the linear ones)
        co1 = co * 1e9
        u1 = [sum(co1[n] * x ** n for n in range(5)) for x in range(5)]
        lin = np.dot([3, -20, 90, 60, -5], u1) / 128
        assert abs(lin - sum(co1[n] * 2.5 ** n for n in range(5))) < 1e-9 * max(abs(q) for q in u1)


def test_weno_is_essentially_non_oscillatory(oracle_mod):
    """A unit step between samples 3 and 4 (or 2 and 3): the interpolant picks
    the smooth side and does not overshoot."""
    v = oracle_mod.weno([0.0, 0.0, 0.0, 1.0, 1.0])
    assert 0.0 <= v < 1e-6 or abs(v) < 1e-6
    v = oracle_mod.weno([0.0, 0.0, 1.0, 1.0, 1.0])
    assert abs(v - 1.0) < 1e-6
    v = oracle_mod.weno([1.0, 1.0, 1.0, 0.0, 0.0])
    assert abs(v - 1.0) < 1e-6


# ------------------------------------------------------------------------ metrics


@pytest.mark.parametrize("A", [np.eye(3), A_SHEAR])
def test_metrics_are_the_affine_closed_forms(oracle_mod, A):
    """Central differences are exact on an affine map: covariant basis = A,
    J = det A, cofactor rows = J * grad(xi_d) = rows of det(A) A^-T; written on
    the reference's index ranges ([1, N-1) for J and x_xi, x_eta; [2, N-2) for
    the cofactors) and zero elsewhere."""
    A = np.asarray(A, np.float64)
    m = affine_model(A=A)
    o = oracle_mod.AwenoOracle(m, prec=8)
    NZ, NY, NX = m.shape_padded
    det = np.linalg.det(A)
    cof = det * np.linalg.inv(A)  # cof[d, a] = J * d xi_d / d x_a
    inner1 = (slice(1, NZ - 1), slice(1, NY - 1), slice(1, NX - 1))
    inner2 = (slice(2, NZ - 2), slice(2, NY - 2), slice(2, NX - 2))
    for d in range(3):
        for a in range(3):
            f = o.setup(3 * d + a)
            np.testing.assert_allclose(f[inner2], cof[d, a], atol=1e-13)
            g = f.copy()
            g[inner2] = 0
            assert not g.any()
    J = o.setup(9)
    np.testing.assert_allclose(J[inner1], det, rtol=1e-13)
    g = J.copy()
    g[inner1] = 0
    assert not g.any()
    for q in range(6):
        f = o.setup(13 + q)
        np.testing.assert_allclose(f[inner1], A[q % 3, q // 3], atol=1e-13)
    np.testing.assert_allclose(o.setup(10), MU, rtol=1e-12)
    np.testing.assert_allclose(o.setup(11), LAM, rtol=1e-12)
    np.testing.assert_allclose(o.setup(12), 1 / RHO, rtol=1e-12)
    assert o.alpha == pytest.approx(VP)


# ------------------------------------------------------------------- flux update


def quadratic_state(x, rng):
    """Conservative fields quadratic in x, y, z, with their exact gradients and
    Hessians; returns (W[9], dW[9][3], d2W[9][3][3]) as functions of x."""
    c0 = rng.standard_normal(9)
    G = rng.standard_normal((9, 3))
    Q = rng.standard_normal((9, 3, 3))
    Q = 0.5 * (Q + Q.transpose(0, 2, 1))
    scale = np.array([RHO, RHO, RHO, 1 / VP, 1 / VP, 1 / VP, 1 / VP, 1 / VP, 1 / VP]) * 1e-3
    L = 5000.0
    W = np.array([scale[v] * (c0[v] + np.einsum("a,a...->...", G[v], x) / L
                              + np.einsum("ab,a...,b...->...", Q[v], x, x) / L ** 2) for v in range(9)])
    dW = np.array([[scale[v] * (G[v, a] + 2 * np.einsum("b,b...->...", Q[v, a], x) / L) / L for a in range(3)]
                   for v in range(9)])
    return W, dW


def elastic_rate(dW):
    """Exact d/dt of the conservative fields (rho v, e_xx, e_yy, e_zz, g_yz,
    g_xz, g_xy) for a homogeneous medium, from their spatial gradients dW[v][a]."""
    b = 1.0 / RHO
    dS = lambda v, a: dW[v][a]
    def dsig(c, a):  # d sigma_c / d x_a with c in xx,yy,zz,xy,xz,yz
        tr = dS(3, a) + dS(4, a) + dS(5, a)
        return {"xx": LAM * tr + 2 * MU * dS(3, a), "yy": LAM * tr + 2 * MU * dS(4, a),
                "zz": LAM * tr + 2 * MU * dS(5, a), "xy": MU * dS(8, a), "xz": MU * dS(7, a),
                "yz": MU * dS(6, a)}[c]
    dv = lambda c, a: b * dW[c][a]
    out = [dsig("xx", 0) + dsig("xy", 1) + dsig("xz", 2),
           dsig("xy", 0) + dsig("yy", 1) + dsig("yz", 2),
           dsig("xz", 0) + dsig("yz", 1) + dsig("zz", 2),
           dv(0, 0), dv(1, 1), dv(2, 2),
           dv(1, 2) + dv(2, 1), dv(0, 2) + dv(2, 0), dv(0, 1) + dv(1, 0)]
    return np.array(out)


def elastic_rate_dir(dW, A, d):
    """Contribution of the computational direction d to the exact rate: the
    conservative form splits the physical divergence as sum_d (grad xi_d) . d/dxi_d,
    i.e. the gradient restricted to its component along xi_d."""
    Ainv = np.linalg.inv(A)  # d xi / d x (per dh)
    # derivative along xi_d of the field: dW . (A[:, d]); its physical gradient part
    part = np.zeros_like(dW)
    for v in range(9):
        dxi = sum(dW[v][a] * A[a, d] for a in range(3))  # d W_v / d xi_d (times 1/dh)
        for a in range(3):
            part[v][a] = Ainv[d, a] * dxi
    return elastic_rate(part)


@pytest.mark.parametrize("A", [np.eye(3), A_SHEAR])
def test_flux_update_is_exact_for_quadratic_fields(oracle_mod, A):
    """WENO reproduces quadratics at the faces (so the Lax-Friedrichs jump
    vanishes), the D2/D4 corrections of quadratic face values cancel in the
    difference, and on an affine grid the cofactor metrics are constant: the
    right-hand side equals DT times the exact elastodynamic rate at every cell
    the stencils reach with written faces. Pins flux signs, component wiring,
    cofactor layout and the 1/J scaling."""
    A = np.asarray(A, np.float64)
    m = affine_model(A=A)
    o = oracle_mod.AwenoOracle(m, prec=8)
    x = m.coord.astype(np.float64)
    W, dW = quadratic_state(x, np.random.default_rng(5))
    for v in range(9):
        o.set_field(v, W[v])
    h = o.rhs()
    want = m.dt * elastic_rate(dW)
    NZ, NY, NX = m.shape_padded
    box = (slice(5, NZ - 6), slice(5, NY - 5), slice(5, NX - 5))
    for v in range(9):
        scale = np.abs(want[v][box]).max()
        np.testing.assert_allclose(h[v][box], want[v][box], rtol=0, atol=1e-9 * scale, err_msg=f"comp {v}")


@pytest.mark.parametrize("A", [np.eye(3), A_SHEAR])
def test_top_layer_flux_update_is_exact_for_quadratic_fields(oracle_mod, A):
    """The three layers below the surface use 6-point corrections of the
    physical fluxes of neighbouring cells; for quadratic fields they are exact
    too. x and y are checked on all three layers; z on the deepest of them,
    whose 7-cell column stays inside the written metrics."""
    A = np.asarray(A, np.float64)
    m = affine_model(A=A)
    o = oracle_mod.AwenoOracle(m, prec=8)
    x = m.coord.astype(np.float64)
    W, dW = quadratic_state(x, np.random.default_rng(6))
    for v in range(9):
        o.set_field(v, W[v])
    NZ, NY, NX = m.shape_padded
    for d in range(3):
        h = o.rhs_dir(d)
        want = m.dt * elastic_rate_dir(dW, A, d)
        ks = slice(NZ - 6, NZ - 3) if d < 2 else slice(NZ - 6, NZ - 5)
        box = (ks, slice(5, NY - 5), slice(5, NX - 5))
        for v in range(9):
            scale = np.abs(m.dt * elastic_rate(dW)[v][box]).max()
            np.testing.assert_allclose(h[v][box], want[v][box], rtol=0, atol=1e-9 * scale,
                                       err_msg=f"dir {d} comp {v}")


def test_direction_split_sums_to_the_full_update(oracle_mod):
    m = M.small_config(nx=16, ny=14, nz=12)
    o = oracle_mod.AwenoOracle(m, prec=8)
    rng = np.random.default_rng(2)
    for v in range(9):
        o.set_field(v, rng.standard_normal(m.shape_padded))
    parts = sum(o.rhs_dir(d) for d in range(3))
    full = o.rhs()
    np.testing.assert_allclose(full, parts, rtol=1e-12, atol=1e-12 * np.abs(full).max())


# ----------------------------------------------------------------------- TVD-RK3


@pytest.mark.parametrize("z", [-0.3, 0.1 + 0.0j, -1.2, 0.7])
def test_tvd_rk3_stability_polynomial(oracle_mod, z):
    """For y' = lambda y with h = z W, the three stages give
    1 + z + z^2/2 + z^3/6 (Shu-Osher SSP-RK3)."""
    z = float(np.real(z))
    n = 9 * 4
    w = np.linspace(0.5, 2.0, n)
    m = np.zeros(n)
    w0 = w.copy()
    for irk in range(3):
        h = z * w
        oracle_mod.tvd_stage(8, irk, np.ascontiguousarray(h), w, m)
    np.testing.assert_allclose(w, w0 * (1 + z + z * z / 2 + z ** 3 / 6), rtol=1e-14)


# ------------------------------------------------------------------ free surface


def surface_frame(coord, k):
    """Unit normal grad(zeta), unit x_xi and x_eta at layer k from central
    differences of the padded coordinates (numpy, independent of the oracle)."""
    c = coord.astype(np.float64)
    d = lambda a, ax: 0.5 * (np.roll(c[a], -1, ax) - np.roll(c[a], 1, ax))
    xi = np.stack([d(a, 2) for a in range(3)])[:, k]
    et = np.stack([d(a, 1) for a in range(3)])[:, k]
    zt = np.stack([d(a, 0) for a in range(3)])[:, k]
    n = np.cross(xi, et, axis=0)
    n /= np.linalg.norm(n, axis=0)
    s = xi / np.linalg.norm(xi, axis=0)
    t = et / np.linalg.norm(et, axis=0)
    return n, s, t


def physical(W, lam, mu, b):
    """(v[3], sigma[3][3]) from conservative fields."""
    v = np.stack([W[0] * b, W[1] * b, W[2] * b])
    tr = W[3] + W[4] + W[5]
    S = np.empty((3, 3) + W.shape[1:])
    for a in range(3):
        S[a, a] = lam * tr + 2 * mu * W[3 + a]
    S[0, 1] = S[1, 0] = mu * W[8]
    S[0, 2] = S[2, 0] = mu * W[7]
    S[1, 2] = S[2, 1] = mu * W[6]
    return v, S


def frame_components(v, S, n, s, t):
    """Components in the (possibly non-orthogonal) frame B = [n s t]:
    v = B v', S = B S' B^T."""
    B = np.stack([n, s, t], axis=1)  # B[a, col]
    Bm = np.moveaxis(B, (0, 1), (-2, -1))
    Binv = np.linalg.inv(Bm)
    vv = np.einsum("...ij,j...->i...", Binv, v)
    SS = np.einsum("...ia,ab...,...jb->ij...", Binv, S, Binv)
    return vv, SS


@pytest.fixture(scope="module")
def fs_case(oracle_mod):
    m = M.small_config(nx=18, ny=16, nz=12, hills=3, seed=11)
    o = oracle_mod.AwenoOracle(m, prec=8)
    rng = np.random.default_rng(9)
    W0 = rng.standard_normal((9,) + m.shape_padded)
    W0[:3] *= 1e3
    W0[3:] *= 1e-6
    for v in range(9):
        o.set_field(v, W0[v])
    o.char_free_surface()
    W1 = o.wavefield()
    return m, o, W0, W1


def test_char_free_surface_zero_traction(fs_case):
    """After the condition the traction sigma . n on the surface vanishes."""
    m, o, W0, W1 = fs_case
    NZ, NY, NX = m.shape_padded
    k = NZ - H - 1
    I = (slice(H, NY - H), slice(H, NX - H))
    lam, mu, b = (o.setup(q)[k][I] for q in (11, 10, 12))
    n, s, t = (f[(slice(None),) + I] for f in surface_frame(m.coord, k))
    v1, S1 = physical(W1[:, k][(slice(None),) + I], lam, mu, b)
    v0, S0 = physical(W0[:, k][(slice(None),) + I], lam, mu, b)
    trac = np.einsum("ab...,b...->a...", S1, n)
    assert np.abs(trac).max() <= 1e-9 * np.abs(S0).max()


def test_char_free_surface_keeps_outgoing_characteristics(fs_case):
    """The update only removes the incoming characteristics: v'_n - T'_nn/(rho vp),
    v'_s - T'_ns/(rho vs), v'_t - T'_nt/(rho vs), T'_ss - l/(l+2m) T'_nn,
    T'_tt - l/(l+2m) T'_nn and T'_st are unchanged; T'_nn = T'_ns = T'_nt = 0 after."""
    m, o, W0, W1 = fs_case
    NZ, NY, NX = m.shape_padded
    k = NZ - H - 1
    I = (slice(H, NY - H), slice(H, NX - H))
    lam, mu, b = (o.setup(q)[k][I] for q in (11, 10, 12))
    rho = 1 / b
    vp = np.sqrt((lam + 2 * mu) * b)
    vs = np.sqrt(mu * b)
    n, s, t = (f[(slice(None),) + I] for f in surface_frame(m.coord, k))
    vf0, Sf0 = frame_components(*physical(W0[:, k][(slice(None),) + I], lam, mu, b), n, s, t)
    vf1, Sf1 = frame_components(*physical(W1[:, k][(slice(None),) + I], lam, mu, b), n, s, t)
    r = lam / (lam + 2 * mu)
    inv = lambda vf, Sf: np.stack([vf[0] - Sf[0, 0] / (rho * vp), vf[1] - Sf[0, 1] / (rho * vs),
                                   vf[2] - Sf[0, 2] / (rho * vs), Sf[1, 1] - r * Sf[0, 0],
                                   Sf[2, 2] - r * Sf[0, 0], Sf[1, 2]])
    a, b_ = inv(vf0, Sf0), inv(vf1, Sf1)
    for q in range(6):
        np.testing.assert_allclose(b_[q], a[q], rtol=1e-9, atol=1e-9 * np.abs(a[q]).max(), err_msg=f"inv {q}")
    for a_, c_ in ((0, 0), (0, 1), (0, 2)):
        assert np.abs(Sf1[a_, c_]).max() <= 1e-9 * np.abs(Sf0).max()


def test_char_free_surface_is_idempotent_and_extrapolates_cubically(fs_case, oracle_mod):
    m, o, W0, W1 = fs_case
    NZ, NY, NX = m.shape_padded
    k = NZ - H - 1
    o2 = oracle_mod.AwenoOracle(m, prec=8)
    for v in range(9):
        o2.set_field(v, W1[v])
    o2.char_free_surface()
    W2 = o2.wavefield()
    I = (slice(None), slice(H, NY - H), slice(H, NX - H))
    np.testing.assert_allclose(W2[:, k][I], W1[:, k][I], rtol=1e-9, atol=1e-12)
    # ghosts: cubic through the surface cell and the three cells below
    col = W1[:, k - 3:k + 1][(slice(None), slice(None)) + I[1:]]
    zs = np.arange(-3, 1, dtype=np.float64)
    V = np.vander(zs, 4)
    coef = np.linalg.solve(V, col.reshape(9, 4, -1).transpose(1, 0, 2).reshape(4, -1))
    for hgt in (1, 2, 3):
        pred = (np.vander(np.array([float(hgt)]), 4) @ coef).reshape(9, NY - 2 * H, NX - 2 * H)
        got = W1[:, k + hgt][I]
        np.testing.assert_allclose(got, pred, rtol=1e-8, atol=1e-8 * np.abs(col).max())
    # nothing below the surface changes
    np.testing.assert_array_equal(W1[:, :k], W0[:, :k])


# ------------------------------------------------------------------------ source


def test_source_is_hookes_law_inverse(oracle_mod):
    """The injected strain decrement, pushed through Hooke's law, is the scaled
    moment rate: lam tr(de) I + 2 mu de = -DT * M * mu / (J dh^3), with the
    tensor components (xx, yy, zz, xy, xz, yz); two points in one cell add."""
    m = affine_model(A=A_SHEAR, nt=8)
    i, j, k = 8, 9, 10
    cell = m.cell_index(i, j, k)
    Mt = np.array([1.0, -2.0, 0.5, 0.7, -0.3, 1.1])
    rate = np.tile(Mt, (2, 8, 1)).astype(np.float32)
    rate[1] *= 0.5
    m.src_cells = np.array([cell, cell], np.int64)
    m.src_rate = rate
    m.src_dt = m.dt
    o = oracle_mod.AwenoOracle(m, prec=8)
    o.source_only()
    W = o.wavefield()
    de = W[3:, k, j, i]
    assert np.count_nonzero(W) == 6
    J = np.linalg.det(np.asarray(A_SHEAR))
    scaled = 1.5 * Mt * MU / (J * m.dh ** 3)
    tr = de[0] + de[1] + de[2]
    sig = np.array([LAM * tr + 2 * MU * de[0], LAM * tr + 2 * MU * de[1], LAM * tr + 2 * MU * de[2],
                    MU * de[5], MU * de[4], MU * de[3]])
    np.testing.assert_allclose(sig, -m.dt * scaled, rtol=1e-6)


# ------------------------------------------------------------------------ sponge


@pytest.mark.parametrize("fs", [True, False])
def test_sponge_is_the_cerjan_profile(oracle_mod, fs):
    """W = 1 becomes the product over the boundaries within 20 cells of
    exp(-(0.015 p)^2), p = 20 - distance; no damping at the free surface."""
    m = affine_model(nx=48, ny=44, nz=46, fs=fs)
    o = oracle_mod.AwenoOracle(m, prec=8)
    NZ, NY, NX = m.shape_padded
    for v in range(9):
        o.set_field(v, np.ones(m.shape_padded))
    o.sponge()
    W = o.wavefield()
    g = lambda p: np.exp(-(0.015 * p) ** 2)
    def axis_factor(n, hi_side=True):
        q = np.arange(n, dtype=np.float64)
        f = np.where(q <= 20, g(20 - q), 1.0)
        if hi_side:
            f = f * np.where(q >= n - 20, g(20 - (n - q)), 1.0)
        return f
    fx, fy, fz = axis_factor(m.nx), axis_factor(m.ny), axis_factor(m.nz, hi_side=not fs)
    want = fz[:, None, None] * fy[None, :, None] * fx[None, None, :]
    inner = (slice(H, NZ - H), slice(H, NY - H), slice(H, NX - H))
    for v in (0, 5, 8):
        np.testing.assert_allclose(W[v][inner], want, rtol=1e-15)
    halo = W[0].copy()
    halo[inner] = 1.0
    assert np.all(halo == 1.0)


# --------------------------------------------------------------------- observers


def test_observers_convert_to_physical(oracle_mod):
    """Stations record (v, sxx, syy, szz, sxy, sxz, syz) from the conservative
    fields; PGV is max |v_h|, |v| on the surface; displacement integrates v DT."""
    m = M.small_config(nx=16, ny=14, nz=12)
    o = oracle_mod.AwenoOracle(m, prec=8)
    rng = np.random.default_rng(4)
    W = rng.standard_normal((9,) + m.shape_padded)
    for v in range(9):
        o.set_field(v, W[v])
    o.observe()
    lam, mu, b = (o.setup(q) for q in (11, 10, 12))
    vv, S = physical(W, lam, mu, b)
    st = o.stations()[:, 0]
    for s_, (i, j, k) in enumerate(m.stations):
        want = [vv[0, k, j, i], vv[1, k, j, i], vv[2, k, j, i], S[0, 0, k, j, i], S[1, 1, k, j, i],
                S[2, 2, k, j, i], S[0, 1, k, j, i], S[0, 2, k, j, i], S[1, 2, k, j, i]]
        np.testing.assert_allclose(st[s_], want, rtol=2e-6)
    NZ, NY, NX = m.shape_padded
    k = NZ - H - 1
    I = (slice(H, NY - H), slice(H, NX - H))
    vs_ = vv[:, k][(slice(None),) + I]
    pgv = o.pgv()
    np.testing.assert_allclose(pgv[..., 0], np.hypot(vs_[0], vs_[1]), rtol=2e-6)
    np.testing.assert_allclose(pgv[..., 1], np.sqrt((vs_ ** 2).sum(0)), rtol=2e-6)
    np.testing.assert_allclose(o.surface_dis(), vs_ * m.dt, rtol=2e-6)


# -------------------------------------------------------------- plane P wave


def test_plane_p_wave_matches_the_analytic_solution(oracle_mod):
    """A smooth P pulse travelling along x in a homogeneous box (no free
    surface): after 12 steps the centre of the box, farther from the sponge
    layers than the pulse can carry their influence, matches the translated
    pulse v_x = f(x - vp t), e_xx = -f/vp. Pins the assembly of the whole step
    (flux update, TVD-RK3) and that the sponge leaves the interior alone."""
    nx, ny, nz = 72, 56, 56
    m = affine_model(nx=nx, ny=ny, nz=nz, fs=False, nt=12)
    m.dt = 0.4 * m.dh / VP
    o = oracle_mod.AwenoOracle(m, prec=8)
    x = m.coord[0].astype(np.float64)
    x0, w = 33 * m.dh, 6 * m.dh
    f = lambda s: 1e-3 * np.exp(-((s - x0) / w) ** 2)
    W = np.zeros((9,) + m.shape_padded)
    W[0] = RHO * f(x)
    W[3] = -f(x) / VP
    for v in range(9):
        o.set_field(v, W[v])
    o.step(m.nt)
    T = m.nt * m.dt
    got = o.wavefield()
    NZ, NY, NX = m.shape_padded
    box = (slice(NZ // 2 - 1, NZ // 2 + 2), slice(NY // 2 - 1, NY // 2 + 2), slice(H + 22, NX - H - 22))
    vx = got[0][box] / RHO
    want = f(x[box] - VP * T)
    err = np.linalg.norm(vx - want) / np.linalg.norm(want)
    assert err < 1e-3, err
    np.testing.assert_allclose(-VP * got[3][box], want, atol=1e-3 * np.abs(want).max())
    for v in (1, 2, 6, 7, 8):
        assert np.abs(got[v][box]).max() < 1e-5 * np.abs(got[0][box]).max()
