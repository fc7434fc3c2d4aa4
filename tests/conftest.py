import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))

H2 = ROOT / "hyperflow_amd" / "mechanisms" / "H2_Li.yaml"
DME = ROOT / "hyperflow_amd" / "mechanisms" / "dme_sk39.yaml"
GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); runs the HIP kernels")


@pytest.fixture(scope="session")
def oracle_h2():
    from oracle import Oracle
    return Oracle(H2)


@pytest.fixture(scope="session")
def oracle_dme():
    from oracle import Oracle
    return Oracle(DME)


@pytest.fixture(scope="session")
def mech_h2():
    import hyperflow_amd as hf
    return hf.Mechanism(H2)


@pytest.fixture(scope="session")
def mech_dme():
    import hyperflow_amd as hf
    return hf.Mechanism(DME)
