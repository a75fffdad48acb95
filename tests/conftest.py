import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "hhg-solids_amd")
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)

HMCASE0 = os.path.join(GOLDEN, "hmcase0_tb.dat")
HALDANE = os.path.join(GOLDEN, "haldane_tb.dat")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def make_param(path_tb, nr=30, tmax=90.0, dt=0.02197, intensity=1e12, lam=3000.0, tmax_field=80.0,
               pol=(0.0, 1.0, 0.0), phi=(0.0, 0.0, 0.0), **extra):
    """The reference README / scripts/test_swe.py parameter dictionary."""
    import numpy as np
    d = {"path_tb": path_tb, "nr1": nr if isinstance(nr, int) else nr[0],
         "nr2": nr if isinstance(nr, int) else nr[1], "nr3": 1, "tmax": tmax, "dt": dt,
         "intensity": intensity, "lambda": lam, "tmax_field": tmax_field,
         "pol_vec": np.array(pol, dtype=float), "phi_vec": np.array(phi, dtype=float)}
    d.update(extra)
    return d


def rel_l2(a, b):
    import numpy as np
    a = np.asarray(a); b = np.asarray(b)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle
    if not os.path.exists(oracle.LIB_PATH):
        oracle.build()
    return oracle
