import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); parity tests of the HIP path")


@pytest.fixture(scope="session")
def known():
    import json
    with open(os.path.join(ROOT, "tests", "golden", "known_answers.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def parser():
    """The HIP engine on device 0.  No CPU fallback: a missing build or device fails."""
    import oneline_amd as ol
    p = ol.GpuParser()
    print(f"[gpu] adapter: {p.adapter_name()}", flush=True)
    yield p
    p.close()
