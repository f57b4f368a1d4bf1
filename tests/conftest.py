import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: takes more than a few seconds on the CPU")
    # the oracle's C restatements are test infrastructure: build them when absent
    odir = os.path.join(ROOT, "oracle")
    if not (os.path.exists(os.path.join(odir, "libgpt_rng.so"))
            and os.path.exists(os.path.join(odir, "libwc_cpu.so"))):
        subprocess.run(["make", "-s", "-C", odir], check=True)


@pytest.fixture(scope="session")
def kats():
    import json

    with open(os.path.join(ROOT, "tests", "golden", "reference_kats.json")) as f:
        return json.load(f)
