"""Interior-kernel variants and the tight device layout against the oracle.

The library picks its interior kernel and device layout from environment
knobs read at cgfdm_create (register z-marching with chunks of 5 planes and
of the whole column, x-paired, plane-tile and LDS z-marching interior
kernels; CGFDM_ROW_ALIGN=0 for the tight layout). Each variant steps the
curvilinear PML + free-surface case and must match the FP32 oracle as the
default does (3e-5 relative L2 per field).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

VARIANTS = [
    ("cgfdm", "CGFDM_INTERIOR_ZREG", "5"),
    ("cgfdm", "CGFDM_INTERIOR_ZREG", "64"),
    ("cgfdm", "CGFDM_INTERIOR_KC", "-1"),
    ("cgfdm", "CGFDM_INTERIOR_KC", "1"),
    ("cgfdm", "CGFDM_INTERIOR_KC", "16"),
    ("cgfdm", "CGFDM_ROW_ALIGN", "0"),
    ("aweno", "CGFDM_ROW_ALIGN", "0"),
]


def rel_l2(a, b):
    n = np.linalg.norm(b)
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / (n if n > 0 else 1.0))


@pytest.mark.parametrize("scheme,var,val", VARIANTS)
def test_variant_matches_oracle(scheme, var, val, monkeypatch, oracle_mod):
    from cgfdm_amd import solver
    import test_gpu_parity as P
    monkeypatch.setenv(var, val)
    m = P.make("curvilinear_pml_fs")
    print(f"variant {scheme} {var}={val}: {m.nx}x{m.ny}x{m.nz}, FP32, 12 steps", flush=True)
    if scheme == "cgfdm":
        g = solver.Solver(m, precision=4)
        o = oracle_mod.Oracle(m, prec=4, threads=8)
    else:
        g = solver.Solver(m, precision=4, scheme=solver.SCHEME_AWENO)
        o = oracle_mod.AwenoOracle(m, prec=4, threads=8)
    g.step(12)
    g.synchronize()
    o.step(12)
    for v in range(9):
        err = rel_l2(g.field(v), o.field(v))
        assert err < 3e-5, (scheme, var, val, v, err)
    sg, so = g.stations(), o.stations()
    assert rel_l2(sg, so) < 3e-5
