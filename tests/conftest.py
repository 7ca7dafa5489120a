import os
import sys

import pytest

# The CGFDM step kernels are switched off in the library (the step returns
# CGFDM_ERR_DISABLED): their first MI355X run, the FP64 24x20x16 PML + free
# surface case, ended in a GPU memory fault whose cause has not been shown.
# Every GPU test that steps the CGFDM solver carries this mark.
CGFDM_STEP_OFF = pytest.mark.skip(
    reason="CGFDM step kernels switched off: unexplained GPU memory fault on their first run "
           "(test_fp64_wavefield_matches_oracle[curvilinear_pml_fs])")

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "cgfdm3d-eqr-sc_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU")
    config.addinivalue_line("markers", "slow: long CPU test")


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    oracle.build()
    return oracle
