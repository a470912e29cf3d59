import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

GOLDEN = Path(__file__).resolve().parent / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels through the C ABI)")


def load_golden(name):
    with np.load(GOLDEN / f"{name}.npz", allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


LBVH_CASES = ["soup2000", "clustered1500", "identical64", "flat300", "sparse_ids777", "tiny1", "tiny2",
              "tiny3", "tiny7", "rounding4", "signed_zero4000"]


@pytest.fixture(scope="session")
def gpu():
    """Initialise the HIP library on device 0 (fails loudly without a GPU)."""
    from lux_amd import lbvh

    lbvh.init(0)
    return lbvh
