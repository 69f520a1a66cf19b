import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden", "rimp2_golden.npz")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU (run with -m gpu)")
    config.addinivalue_line("markers", "slow: long-running")


@pytest.fixture(scope="session")
def golden():
    z = np.load(GOLDEN, allow_pickle=False)
    cases = {}
    for name in z["names"]:
        name = str(name)
        d = z[f"{name}/dims"]
        case = {
            "b": z[f"{name}/b"], "eps": z[f"{name}/eps"], "nocc": int(d[0]), "nvir": int(d[1]),
            "first_occ": int(d[2]), "nocc_total": int(d[3]), "naux": int(d[4]),
            "energy": z[f"{name}/energy"], "energy_kappa": z[f"{name}/energy_kappa1.3"],
        }
        for tag in ("mp2", "scs"):
            k = f"{name}/gamma_{tag}"
            if k in z.files:
                case[f"gamma_{tag}"] = z[k]
                case[f"deps_{tag}"] = z[f"{name}/deps_{tag}"]
        cases[name] = case
    return cases


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle

    oracle.build()
    return oracle
