"""The C++ consumer of the ABI (tests/abi_consumer.cpp, built in-tree by
hyperflow_amd.build) on the GPU: the integration guide's ChemistryBatch
wrapper equals a direct hf_chem_advance bit for bit and keeps the
reference's error contract (chemistry_batch.cpp:21-50), and an Euler stage
over boxes in separate allocations at descending addresses
(hf_flux_stage_boxes, the AMReX FArrayBox layout of BlockAMR.cpp:210-243)
equals hf_flux_stage on the same boxes packed into one array."""
import subprocess

import pytest

from hyperflow_amd import build as B
from tests.conftest import H2

pytestmark = pytest.mark.gpu


def test_cxx_consumer_on_the_gpu():
    assert B.CONSUMER.exists(), "build the library first (python -m hyperflow_amd.build)"
    print(f"[gpu] {B.CONSUMER} {H2}", flush=True)
    r = subprocess.run([str(B.CONSUMER), str(H2)], capture_output=True, text=True, timeout=300)
    print(r.stdout, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = [ln for ln in r.stdout.splitlines() if ln.strip()]
    assert len(lines) == 8 and all(ln.startswith("PASS") for ln in lines), r.stdout
