"""ME-AWENO HIP path vs the CPU oracle on identical inputs (runs on the MI355X).

Tolerances: FP64 GPU vs FP64 oracle 1e-10 relative L2 (rounding order only);
FP32 GPU vs FP32 oracle 1e-5 (the GPU evaluates the WENO weights in FP32, the
oracle in FP64 as the reference does; the two FP32 runs then differ by a few
1e-7, the size of the FP32-vs-FP64 gap itself); FP32 GPU vs FP64 oracle 1e-4
on the station seismograms.
"""
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


@pytest.fixture(scope="module")
def solver_mod():
    from cgfdm_amd import solver
    return solver


CASES = {
    "aweno_curvilinear_fs": lambda nt: M.small_config(nx=24, ny=20, nz=16, hills=2, layered=True, free_surface=True,
                                                      nt=nt),
    "aweno_no_free_surface": lambda nt: M.small_config(nx=22, ny=18, nz=14, hills=0, layered=True,
                                                       free_surface=False, nt=nt),
    "aweno_basin": lambda nt: M.small_basin_config(nt=nt),
    "aweno_two_x_tiles": lambda nt: M.small_config(nx=70, ny=12, nz=14, hills=2, layered=True, free_surface=True,
                                                   nt=nt),
}


def run_pair(name, oracle_mod, solver_mod, prec, steps, oracle_prec=None):
    m = CASES[name](steps)
    m.name = name
    print(f"case {name} {m.nx}x{m.ny}x{m.nz} fs={m.free_surface} prec={prec} steps={steps}", flush=True)
    g = solver_mod.Solver(m, precision=prec, scheme=solver_mod.SCHEME_AWENO)
    o = oracle_mod.AwenoOracle(m, prec=oracle_prec or prec, threads=8)
    g.step(steps)
    g.synchronize()
    o.step(steps)
    return m, g, o


@pytest.mark.parametrize("name", list(CASES))
def test_aweno_fp64_matches_oracle(name, oracle_mod, solver_mod):
    m, g, o = run_pair(name, oracle_mod, solver_mod, 8, 12)
    assert g.it == o.it == 12
    wg, wo = g.wavefield(), o.wavefield()
    for v in range(9):
        assert rel_l2(wg[v], wo[v]) < 1e-10, (name, v, rel_l2(wg[v], wo[v]))
    np.testing.assert_allclose(g.stations(), o.stations(), rtol=1e-6, atol=1e-6 * np.abs(o.stations()).max())
    if m.free_surface:
        np.testing.assert_allclose(g.pgv(), o.pgv(), rtol=1e-6, atol=1e-6 * o.pgv().max())
        np.testing.assert_allclose(g.surface_displacement(), o.surface_dis(), rtol=1e-6,
                                   atol=1e-6 * np.abs(o.surface_dis()).max())


@pytest.mark.parametrize("name", list(CASES))
def test_aweno_fp32_matches_oracle(name, oracle_mod, solver_mod):
    m, g, o = run_pair(name, oracle_mod, solver_mod, 4, 12)
    wg, wo = g.wavefield(), o.wavefield()
    for v in range(9):
        assert rel_l2(wg[v], wo[v]) < 1e-5, (name, v, rel_l2(wg[v], wo[v]))


def test_aweno_setup_products_match_oracle(oracle_mod, solver_mod):
    m = CASES["aweno_curvilinear_fs"](4)
    g = solver_mod.Solver(m, precision=8, scheme=solver_mod.SCHEME_AWENO)
    o = oracle_mod.AwenoOracle(m, prec=8)
    for what in range(19):
        a, b = g.setup(what), o.setup(what)
        np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-12 * max(abs(b).max(), 1e-30), err_msg=str(what))


def test_aweno_fp32_seismograms_vs_fp64_oracle(oracle_mod, solver_mod):
    """The accuracy statement of the FP32 path: station seismograms within
    1e-4 relative L2 of the FP64 oracle on the basin case."""
    m, g, o = run_pair("aweno_basin", oracle_mod, solver_mod, 4, 30, oracle_prec=8)
    sg, so = g.stations(), o.stations()
    assert rel_l2(sg[..., :3], so[..., :3]) < 1e-4
    assert rel_l2(g.pgv(), o.pgv()) < 1e-4
