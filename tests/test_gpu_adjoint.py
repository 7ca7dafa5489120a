"""HIP adjoint of the CGFDM step against the oracle's discrete adjoint (MI355X).

The oracle adjoint is pinned by the dot-product identity and by central finite
differences (tests/test_adjoint_oracle.py). Here the GPU gradients of
J = 1/2 |station velocity|^2 with respect to mu, lambda, 1/rho and the
moment-rate table must match it: FP64 to 1e-10 of the largest entry, FP32 to
1e-3 (the FP32 forward/adjoint accumulate rounding over the run).
"""
import numpy as np
import pytest

from cgfdm_amd import model as M

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def solver_mod():
    from cgfdm_amd import solver
    return solver


CASES = {
    "pml_fs_curvilinear": dict(nx=20, ny=18, nz=14, npml=4, hills=2, free_surface=True),
    "pml_no_fs": dict(nx=18, ny=16, nz=12, npml=3, hills=0, free_surface=False),
    "no_pml_fs": dict(nx=16, ny=14, nz=12, npml=0, hills=2, free_surface=True),
}


def run(name, prec, oracle_mod, solver_mod, nst=12, ck=5):
    m = M.small_config(nt=nst, **CASES[name])
    print(f"adjoint case {name} prec={prec} steps={nst} checkpoint={ck}", flush=True)
    g = solver_mod.Solver(m, precision=prec)
    rec = g.forward_record(nst, ck)
    gr = g.adjoint(rec)
    o = oracle_mod.Oracle(m, prec=8, threads=8)
    r = oracle_mod.adjoint_misfit(o, nst, np.zeros((len(m.stations), nst, 3)), npts=len(m.src_cells),
                                  nt_src=m.src_rate.shape[1])
    return m, g, rec, gr, r


@pytest.mark.parametrize("name", list(CASES))
def test_fp64_adjoint_matches_oracle(name, oracle_mod, solver_mod):
    m, g, rec, gr, r = run(name, 8, oracle_mod, solver_mod)
    assert 0.5 * (rec ** 2).sum() == pytest.approx(r["J"], rel=1e-10)
    for k in ("gmu", "glam", "gb", "grate"):
        scale = np.abs(r[k]).max()
        assert scale > 0
        np.testing.assert_allclose(gr[k], r[k], rtol=0, atol=1e-10 * scale, err_msg=k)


def test_fp64_adjoint_with_stations_sharing_cells(oracle_mod, solver_mod):
    """Stations that share a cell (two pairs here) inject their adjoint sources
    into the same cell, in station order, as the oracle does."""
    nst, ck = 8, 3
    m = M.small_config(nt=nst, **CASES["pml_no_fs"])
    st = np.asarray(m.stations)
    m.stations = np.ascontiguousarray(np.concatenate([st, st[:1], st[2:3]]), np.int32)
    print(f"adjoint with {len(m.stations)} stations, two sharing a cell with others", flush=True)
    g = solver_mod.Solver(m, precision=8)
    rec = g.forward_record(nst, ck)
    gr = g.adjoint(rec)
    o = oracle_mod.Oracle(m, prec=8, threads=8)
    r = oracle_mod.adjoint_misfit(o, nst, np.zeros((len(m.stations), nst, 3)), npts=len(m.src_cells),
                                  nt_src=m.src_rate.shape[1])
    assert 0.5 * (rec ** 2).sum() == pytest.approx(r["J"], rel=1e-10)
    for k in ("gmu", "glam", "gb", "grate"):
        scale = np.abs(r[k]).max()
        assert scale > 0
        np.testing.assert_allclose(gr[k], r[k], rtol=0, atol=1e-10 * scale, err_msg=k)


def test_fp32_adjoint_matches_oracle(oracle_mod, solver_mod):
    m, g, rec, gr, r = run("pml_fs_curvilinear", 4, oracle_mod, solver_mod)
    for k in ("gmu", "glam", "gb", "grate"):
        err = np.linalg.norm(gr[k] - r[k]) / np.linalg.norm(r[k])
        assert err < 1e-3, (k, err)


def test_adjoint_leaves_the_forward_state(oracle_mod, solver_mod):
    """After forward_record + adjoint the solver is at the end of the forward run
    and continues exactly like a plain run."""
    m = M.small_config(nt=20, **CASES["pml_fs_curvilinear"])
    a = solver_mod.Solver(m, precision=8)
    rec = a.forward_record(10, 3)
    a.adjoint(rec)
    a.step(5)
    b = solver_mod.Solver(m, precision=8)
    b.step(15)
    assert a.it == b.it == 15
    np.testing.assert_array_equal(a.wavefield(), b.wavefield())


@pytest.mark.parametrize("nst,ck", [(10, 3), (6, 6)])
def test_stage_store_paths_agree(nst, ck, solver_mod, monkeypatch):
    """forward_record keeps the stage states of its last segment for the
    adjoint, which then recomputes only the other segments. The gradients equal
    those of a run without the stage store (every segment and stage
    recomputed), and a second adjoint call after the same forward_record gives
    the same gradients (the store then holds the first segment). The surface
    pass adds with atomics, so agreement is to rounding: 1e-12 of the largest
    entry in FP64."""
    m = M.small_config(nt=nst, **CASES["pml_fs_curvilinear"])
    print(f"stage store: {nst} steps, checkpoint {ck}", flush=True)
    g = solver_mod.Solver(m, precision=8)
    rec = g.forward_record(nst, ck)
    first = g.adjoint(rec)
    second = g.adjoint(rec)
    g.close()
    monkeypatch.setenv("CGFDM_ADJOINT_STORE_GB", "0")
    h = solver_mod.Solver(m, precision=8)
    rec0 = h.forward_record(nst, ck)
    plain = h.adjoint(rec0)
    h.close()
    np.testing.assert_array_equal(rec, rec0)
    for k in ("gmu", "glam", "gb", "grate"):
        scale = np.abs(plain[k]).max()
        assert scale > 0, k
        np.testing.assert_allclose(first[k], plain[k], rtol=0, atol=1e-12 * scale, err_msg=k)
        np.testing.assert_allclose(second[k], plain[k], rtol=0, atol=1e-12 * scale, err_msg=k)


def test_stage_store_reused_by_the_next_forward_record(solver_mod):
    """The stage store stays allocated from one forward_record to the next with
    the same checkpoint interval: the second forward_record + adjoint of a
    solver equals those of a fresh solver started at the same step."""
    nst, ck = 6, 4
    m = M.small_config(nt=3 * nst, **CASES["pml_fs_curvilinear"])
    print(f"stage store reuse: {nst} steps, checkpoint {ck}", flush=True)
    g = solver_mod.Solver(m, precision=8)
    g.adjoint(g.forward_record(nst, ck))
    rec = g.forward_record(nst, ck)
    grads = g.adjoint(rec)
    g.close()
    h = solver_mod.Solver(m, precision=8)
    h.step(nst)
    rec_h = h.forward_record(nst, ck)
    grads_h = h.adjoint(rec_h)
    h.close()
    np.testing.assert_array_equal(rec, rec_h)
    for k in ("gmu", "glam", "gb", "grate"):
        scale = np.abs(grads_h[k]).max()
        assert scale > 0, k
        np.testing.assert_allclose(grads[k], grads_h[k], rtol=0, atol=1e-12 * scale, err_msg=k)


@pytest.mark.parametrize("store_gb", [None, "0"])
def test_adjoint_buffers_reused_across_run_lengths(store_gb, solver_mod, monkeypatch):
    """The checkpoints, station records and work buffers of the adjoint stay
    allocated with the solver and grow when a later forward_record needs more
    (more steps, more checkpoints, a longer segment) or are reused when it needs
    less. Runs of 6, 10 and 4 steps with checkpoints every 4, 5 and 3 steps on
    one solver give the gradients of fresh solvers started at the same steps,
    with the stage store (None: sized from the free memory) and without it
    ("0": the segment states are recomputed)."""
    if store_gb is not None:
        monkeypatch.setenv("CGFDM_ADJOINT_STORE_GB", store_gb)
    runs = [(6, 4), (10, 5), (4, 3)]
    m = M.small_config(nt=sum(n for n, _ in runs) + 1, **CASES["pml_fs_curvilinear"])
    print(f"adjoint buffer reuse over runs {runs}, store budget {store_gb}", flush=True)
    g = solver_mod.Solver(m, precision=8)
    start = 0
    for nst, ck in runs:
        rec = g.forward_record(nst, ck)
        grads = g.adjoint(rec)
        h = solver_mod.Solver(m, precision=8)
        if start:
            h.step(start)
        rec_h = h.forward_record(nst, ck)
        grads_h = h.adjoint(rec_h)
        h.close()
        np.testing.assert_array_equal(rec, rec_h)
        for k in ("gmu", "glam", "gb", "grate"):
            scale = np.abs(grads_h[k]).max()
            assert scale > 0, (nst, ck, k)
            np.testing.assert_allclose(grads[k], grads_h[k], rtol=0, atol=1e-12 * scale, err_msg=f"{nst} {ck} {k}")
        start += nst
    g.close()


def test_device_bytes_drop_when_the_stage_store_is_freed(solver_mod, monkeypatch):
    """The solver's device-memory count goes down by what it frees: a
    forward_record under a zero stage-store budget frees the store that the
    previous one allocated (4 stage states of 9 fields per step of the 3-step
    segment, FP64: at least 36 x 3 x 8 bytes per cell)."""
    monkeypatch.delenv("CGFDM_ADJOINT_STORE_GB", raising=False)
    nst, ck = 6, 3
    m = M.small_config(nt=2 * nst + 1, **CASES["pml_fs_curvilinear"])
    print(f"device bytes with and without the stage store: {nst} steps, checkpoint {ck}", flush=True)
    g = solver_mod.Solver(m, precision=8)
    g.adjoint(g.forward_record(nst, ck))
    with_store = g.device_bytes()
    monkeypatch.setenv("CGFDM_ADJOINT_STORE_GB", "0")
    rec = g.forward_record(nst, ck)
    without = g.device_bytes()
    g.adjoint(rec)  # the recompute path then keeps the segment's step states
    assert g.device_bytes() > without
    cells = g.N
    g.close()
    assert without > 0
    assert with_store - without >= 36 * ck * 8 * cells, (with_store, without)


def test_autograd_wrapper_on_the_hip_backend(oracle_mod):
    """torch gradients through the HIP adjoint equal those through the oracle."""
    import torch
    from cgfdm_amd import autograd as A
    from test_autograd_wrapper import OracleBackend
    nst = 8
    m = M.small_config(nx=12, ny=11, nz=10, npml=3, hills=2, free_surface=True, nt=nst)
    w = torch.tensor(np.random.default_rng(3).standard_normal((len(m.stations), nst, 3)))
    grads = []
    for backend in (None, lambda mm, prec: OracleBackend(mm, prec, oracle_mod)):
        ts = [torch.tensor(m.medium[c], dtype=torch.float64, requires_grad=True) for c in range(3)]
        rate = torch.tensor(m.src_rate, dtype=torch.float64, requires_grad=True)
        loss = (A.seismograms(*ts, rate, m, nst, precision=8, checkpoint_every=3, backend=backend) * w).sum()
        loss.backward()
        grads.append([t.grad.numpy() for t in ts] + [rate.grad.numpy()])
    for a, b in zip(*grads):
        np.testing.assert_allclose(a, b, rtol=0, atol=1e-10 * np.abs(b).max())


def test_hip_gradients_match_finite_differences_of_the_hip_forward(solver_mod):
    """Central differences of the GPU forward run (FP64) in one random direction
    per parameter field (vs, vp, rho, moment-rate table) against the gradient of
    the GPU adjoint through the torch wrapper; relative error <= 1e-5."""
    import torch
    from cgfdm_amd import autograd as A
    nst = 8
    m = M.small_config(nx=12, ny=11, nz=10, npml=3, hills=2, free_surface=True, nt=nst)
    w = np.random.default_rng(5).standard_normal((len(m.stations), nst, 3))
    ts = [torch.tensor(m.medium[c], dtype=torch.float64, requires_grad=True) for c in range(3)]
    rate = torch.tensor(m.src_rate, dtype=torch.float64, requires_grad=True)
    loss = (A.seismograms(*ts, rate, m, nst, precision=8, checkpoint_every=3) * torch.tensor(w)).sum()
    loss.backward()
    grads = [t.grad.numpy() for t in ts] + [rate.grad.numpy()]
    base = [m.medium[0].astype(np.float64), m.medium[1].astype(np.float64), m.medium[2].astype(np.float64),
            m.src_rate.astype(np.float64)]

    def loss_at(args):
        mm = M.small_config(nx=12, ny=11, nz=10, npml=3, hills=2, free_surface=True, nt=nst)
        mm.medium = np.stack(args[:3]).astype(np.float32)
        mm.src_rate = args[3].astype(np.float32)
        g = solver_mod.Solver(mm, precision=8)
        return float((g.forward_record(nst, nst) * w).sum())

    rng = np.random.default_rng(9)
    for k in range(4):
        d = base[k] * rng.uniform(-0.5, 0.5, base[k].shape)
        eps = 1e-3
        vals, used = [], []
        for sg in (1, -1):
            args = [b.copy() for b in base]
            args[k] = (args[k] + sg * eps * d).astype(np.float32).astype(np.float64)
            used.append(args[k])
            vals.append(loss_at(args))
        fd = (vals[0] - vals[1]) / (2 * eps)
        an = float((grads[k] * (used[0] - used[1]) / (2 * eps)).sum())
        assert abs(fd - an) <= 1e-5 * abs(fd), (k, fd, an)


def test_adjoint_with_recomputed_stages_matches_oracle(oracle_mod, solver_mod, monkeypatch):
    """With no memory budget for the stage states of a segment the adjoint
    recomputes them per step; the gradients are the same."""
    monkeypatch.setenv("CGFDM_ADJOINT_STORE_GB", "0")
    m, g, rec, gr, r = run("pml_fs_curvilinear", 8, oracle_mod, solver_mod)
    for k in ("gmu", "glam", "gb", "grate"):
        scale = np.abs(r[k]).max()
        np.testing.assert_allclose(gr[k], r[k], rtol=0, atol=1e-10 * scale, err_msg=k)
