"""The operation counts behind the chemistry roofline figure (CPU only)."""
import shutil
from pathlib import Path

import pytest

from hyperflow_amd import flops as F

HIPCC = "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not Path(HIPCC).exists() and shutil.which("hipcc") is None, reason="hipcc not installed")
def test_math_costs_match_generated_code():
    """exp, log, log10, division and sqrt are counted as the FP64 arithmetic
    of the code the compiler generates for them on gfx950."""
    hipcc = HIPCC if Path(HIPCC).exists() else shutil.which("hipcc")
    assert F.measure_math_flops(hipcc) == F.MATH_FLOPS


def test_counts_scale_with_the_operation_mix():
    """cell_flops is linear in the kernel's counters, and each term is the
    per-operation count of its counter."""
    from hyperflow_amd.build import GEN
    inc = GEN / "h2_li.inc"
    if not inc.exists():
        pytest.skip("topology headers are generated by the build")
    ns, rx = F.topology_summary(inc)
    assert ns == 9 and len(rx) == 21
    n = ns - 1
    one = dict(rhs=1, jac=0, lu=0, newton=0, steps=0)
    assert F.cell_flops(one, ns, rx) == F.rhs_flops(ns, rx)
    mix = dict(rhs=3, jac=2, lu=5, newton=7, steps=11)
    expect = (3 * F.rhs_flops(ns, rx) + 2 * F.jac_flops(ns, rx) + 5 * F.lu_flops(n) + 7 * F.newton_flops(n)
              + 11 * F.step_flops(n))
    assert F.cell_flops(mix, ns, rx) == pytest.approx(expect, rel=1e-15)
    # every reaction costs at least its forward exponential, every reversible one a second
    nrev = sum(r["rev"] for r in rx)
    assert F.rhs_flops(ns, rx) > (len(rx) + nrev) * F.EXP + F.LOG
