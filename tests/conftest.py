import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def _gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    """Initialised device handle for GPU tests; the library must be built."""
    if not _gpu_available():
        pytest.skip("no GPU")
    import torch
    import flashgemm_amd as fg
    fg.init(0)
    torch.cuda.init()
    return torch.device("cuda:0")


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle
    if not os.path.exists(oracle.LIB_PATH):
        oracle.build()
    return oracle


def golden_names(prefix: str = ""):
    return sorted(f[:-4] for f in os.listdir(GOLDEN) if f.endswith(".npz") and f.startswith(prefix))


def load_golden(name: str):
    import numpy as np
    z = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    d = {k: z[k] for k in z.files}
    n = sum(1 for k in d if k.startswith("A") and k[1:].isdigit())
    d["As"] = [d[f"A{i}"] for i in range(n)]
    d["dtype"] = str(d["dtype"])
    d["threads"] = int(d["threads"])
    return d
