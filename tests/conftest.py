import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")


@pytest.fixture(scope="session")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import tressfx_amd
    tressfx_amd.init(0)
    return torch


@pytest.fixture(scope="session")
def config3_workloads():
    """The batched workload at full size: 256 instances, strand counts
    log-uniform in [500, 50,000], 16/32/64 vertices per strand (76.4 M
    vertices), generated by spawned workers (safe after the GPU is in use)."""
    from tressfx_amd import synth
    workers = max(1, min(16, os.cpu_count() or 1))
    print(f"[config3] generating 256 instances ({workers} spawned workers)", flush=True)
    wls = synth.config3(workers=workers, start_method="spawn")
    assert sum(w.asset.num_total_vertices for w in wls) == 76392480
    return wls
