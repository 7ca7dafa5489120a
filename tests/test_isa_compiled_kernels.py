"""The library's compiled gfx950 kernels executed on the CPU.

Each case runs the real host library against a stand-in HIP runtime
(tests/isa/fakehip.c) that hands every kernel launch to an interpreter of the
kernel's gfx950 assembly (tests/isa/gcnemu.py, assembly from hipcc -S with the
library's flags). Every memory access is checked against the live
allocations and the result is compared with the oracle. This is how a
code-generation fault of the step kernels (a store through a stale base
register on one RK stage) was found, and it is the check every kernel passes
before it runs on the GPU. Slow (minutes per case): set CGFDM_ISA_TESTS=1.
"""
import os
import re
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
pytestmark = pytest.mark.skipif(os.environ.get("CGFDM_ISA_TESTS") != "1",
                                reason="slow (minutes per case): set CGFDM_ISA_TESTS=1")

# (scheme, case, precision, steps, extra, tolerance vs the oracle, environment);
# extra: coefficient storage (cgfdm) or checkpoint interval (adjoint)
CASES = [
    ("cgfdm", "curvilinear_pml_fs", 8, 2, 0, 1e-12, {}),
    ("cgfdm", "no_free_surface", 4, 2, 0, 1e-5, {}),
    ("cgfdm", "curvilinear_pml_fs", 4, 1, 1, 5e-3, {}),
    ("aweno", "aweno_curvilinear_fs", 8, 1, 0, 1e-12, {}),
    ("aweno", "aweno_two_x_tiles", 4, 1, 0, 1e-5, {}),
    # pipelined adjoint passes over two checkpoint segments, with the stage store
    # (the stand-in runtime reports no free memory, so the budget is given) and
    # without it (stages recomputed per step)
    ("adjoint", "pml_fs_curvilinear", 8, 3, 2, 1e-12, {"CGFDM_ADJOINT_STORE_GB": "1"}),
    ("adjoint", "pml_no_fs", 8, 2, 5, 1e-12, {}),
]


@pytest.mark.parametrize("scheme,name,prec,steps,extra,tol,env", CASES)
def test_compiled_kernels_run_in_bounds_and_match_the_oracle(scheme, name, prec, steps, extra, tol, env):
    r = subprocess.run([sys.executable, os.path.join(ROOT, "tests", "isa", "run_case.py"), scheme, name, str(prec),
                        str(steps), str(extra)], capture_output=True, text=True, timeout=3000,
                       env=dict(os.environ, **env))
    line = [l for l in r.stdout.splitlines() if l.startswith("RESULT")]
    assert r.returncode == 0 and line, r.stdout[-2000:] + r.stderr[-2000:]
    assert "fault None" in line[0], line[0]
    err = float(re.search(r"max rel error ([0-9.e+-]+)", line[0]).group(1))
    assert err < tol, line[0]
