"""cgfdm_get_slice on the MI355X: every component on the three plane orientations.

CGFDM slices are the stored wavefield (stress in physical units) on the plane.
ME-AWENO slices convert the conservative state as wave_conserv2phy does
(data_io.cpp:435-483): v = (rho v) / rho, Txx = (lam + 2 mu) exx + lam (eyy + ezz)
and its permutations, Tyz = mu gyz, Txz = mu gxz, Txy = mu gxy, with the state
order (rho vx, rho vy, rho vz, exx, eyy, ezz, gyz, gxz, gxy) and mu = rho Vs^2,
lam = rho (Vp^2 - 2 Vs^2) of the medium. Each plane must match the full field
read through cgfdm_get_field (converted here in FP64 from the model's medium for
ME-AWENO) relative to the larger of the plane's peak and a tenth of the
component's peak: 1e-5 for CGFDM (a copy), 1e-4 for ME-AWENO, whose conversion
runs in FP32 on the device with the moduli rounded to FP32 (a wrong index or
term is an O(1) error).
"""
import numpy as np
import pytest

from cgfdm_amd import model as M

pytestmark = pytest.mark.gpu
H = M.HALO


def planes(nz, ny, nx):
    """(axis, padded index, numpy index of the interior plane) for first, middle and last planes."""
    out = []
    for axis, n in ((0, nx), (1, ny), (2, nz)):
        for q in (0, n // 2, n - 1):
            idx = H + q
            sl = {0: np.s_[H:-H, H:-H, idx], 1: np.s_[H:-H, idx, H:-H], 2: np.s_[idx, H:-H, H:-H]}[axis]
            out.append((axis, idx, sl))
    return out


def physical_aweno(u, medium):
    vs, vp, rho = (medium[c].astype(np.float64) for c in range(3))
    mu = rho * vs * vs
    lam = rho * (vp * vp - 2.0 * vs * vs)
    exx, eyy, ezz, gyz, gxz, gxy = u[3], u[4], u[5], u[6], u[7], u[8]
    return np.stack([u[0] / rho, u[1] / rho, u[2] / rho,
                     (lam + 2 * mu) * exx + lam * (eyy + ezz),
                     (lam + 2 * mu) * eyy + lam * (exx + ezz),
                     (lam + 2 * mu) * ezz + lam * (exx + eyy),
                     mu * gxy, mu * gxz, mu * gyz])


@pytest.mark.parametrize("scheme", ["cgfdm", "aweno"])
def test_slices_match_field(scheme):
    from cgfdm_amd import solver
    m = M.small_config(nx=70, ny=20, nz=16, hills=2, layered=True, free_surface=True, nt=12)
    print(f"slices {scheme}: {m.nx}x{m.ny}x{m.nz}, FP32, 10 steps", flush=True)
    if scheme == "cgfdm":
        g = solver.Solver(m, precision=4)
    else:
        g = solver.Solver(m, precision=4, scheme=solver.SCHEME_AWENO)
    g.step(10)
    g.synchronize()
    u = g.wavefield()
    ref = u if scheme == "cgfdm" else physical_aweno(u, m.medium)
    tol = 1e-5 if scheme == "cgfdm" else 1e-4
    for c in range(9):
        assert np.abs(ref[c]).max() > 0, c
        for axis, idx, sl in planes(m.nz, m.ny, m.nx):
            want = ref[c][sl]
            got = g.slice(c, axis, idx)
            assert got.shape == want.shape, (c, axis, got.shape, want.shape)
            # planes where a component (nearly) vanishes, as on a traction-free
            # top, are measured against a tenth of the component's peak
            peak = max(float(np.abs(want).max()), 0.1 * float(np.abs(ref[c]).max()))
            err = float(np.abs(got - want).max()) / peak
            assert err < tol, (scheme, c, axis, idx, err)


def test_slice_rejects_out_of_range():
    from cgfdm_amd import solver
    m = M.small_config(nx=24, ny=20, nz=16, hills=0, layered=False, free_surface=True, nt=4)
    g = solver.Solver(m, precision=4)
    for comp, axis, idx in ((9, 0, H), (0, 3, H), (0, 0, H - 1), (0, 0, H + m.nx), (0, 2, H + m.nz)):
        with pytest.raises(Exception):
            g.slice(comp, axis, idx)
