"""One GPU test case through the library on the interpreted GPU, compared with
the oracle. Usage: run_case.py cgfdm|aweno|adjoint CASE PREC STEPS [COEF|CHECKPOINT]
Prints one line starting with RESULT. Test infrastructure only."""
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [HERE, os.path.join(ROOT, "cgfdm3d-eqr-sc_amd"), os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")]
import simgpu  # noqa: E402

sim = simgpu.Sim()
import numpy as np  # noqa: E402
from cgfdm_amd import model as M  # noqa: E402
from cgfdm_amd import solver as S  # noqa: E402
import oracle  # noqa: E402

scheme, name, prec, steps = sys.argv[1], sys.argv[2], int(sys.argv[3]), int(sys.argv[4])
extra = int(sys.argv[5]) if len(sys.argv) > 5 else 0
t = time.time()
if scheme == "adjoint":
    import test_gpu_adjoint as T
    m = M.small_config(nt=steps, **T.CASES[name])
    g = S.Solver(m, precision=prec)
    rec = g.forward_record(steps, max(extra, 1))
    gr = g.adjoint(rec)
    o = oracle.Oracle(m, prec=8, threads=4)
    r = oracle.adjoint_misfit(o, steps, np.zeros((len(m.stations), steps, 3)), npts=len(m.src_cells),
                              nt_src=m.src_rate.shape[1])
    err = max(float(np.abs(gr[k] - r[k]).max() / max(np.abs(r[k]).max(), 1e-300)) for k in ("gmu", "glam", "gb", "grate"))
    print(f"RESULT adjoint {name} prec {prec} steps {steps}: max rel error {err:.3e} fault {sim.fault} "
          f"launches {sim.nlaunch} {time.time() - t:.0f}s", flush=True)
    sys.exit(0)
if scheme == "cgfdm":
    import test_gpu_parity as T
    m = T.make(name, nt=max(12, steps))
    g = S.Solver(m, precision=prec, coef_storage=extra)
    o = oracle.Oracle(m, prec=prec if extra == 0 else 8, threads=4)
else:
    import test_gpu_aweno as T
    m = T.CASES[name](max(12, steps))
    g = S.Solver(m, precision=prec, scheme=S.SCHEME_AWENO)
    o = oracle.AwenoOracle(m, prec=prec, threads=4)
g.step(steps)
g.synchronize()
o.step(steps)
wg, wo = g.wavefield(), o.wavefield()
err = max(float(np.linalg.norm(wg[v] - wo[v]) / max(np.linalg.norm(wo[v]), 1e-300)) for v in range(9))
print(f"RESULT {scheme} {name} prec {prec} steps {steps}: max rel error {err:.3e} fault {sim.fault} "
      f"launches {sim.nlaunch} {time.time() - t:.0f}s", flush=True)
