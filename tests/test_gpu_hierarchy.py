"""The multi-level, multi-box AMR step loop on the GPU (hyperflow_amd.hierarchy.
AmrDriver) against the oracle's restatement (oracle/hierarchy.py) and the
two-level driver.

Bars: states per level within 1e-8 of the oracle, normalised per component
by the level's largest magnitude, with chemistry converged (rtol 1e-12,
atol 1e-20) on both sides so that the integrator's own tolerance does not
enter (the step-loop bar of tests/test_gpu_driver.py); conservation of mass
and energy in a closed domain to 1e-13; bit-identical results where the
arithmetic is the same (one fine box against the two-level driver, a fine
level cut into several boxes against one box)."""
import numpy as np
import pytest

import hyperflow_amd as hf
from hyperflow_amd.driver import TwoLevelDriver, channel_2d
from hyperflow_amd.hierarchy import AmrDriver
from oracle.driver import REFLECTING, TRANSMISSIVE
from oracle.hierarchy import OracleAmrDriver
from tests.conftest import H2
from tests.test_oracle_amr import two_level_case
from tests.test_oracle_hierarchy import LEVEL1, LEVEL2, _split

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu
CONVERGED = dict(rtol=1e-12, atol=1e-20)


def levels_close(d, od, tol=1e-8):
    worst = 0.0
    for lev in range(d.finest + 1):
        Q, m = d.level_state(lev)
        R = od.interior(lev)
        assert (m == od.L[lev]["mask"]).all()
        v = m > 0
        a, b = Q[:, v], R[:, v]
        scale = np.abs(b).max(axis=1, keepdims=True) + 1e-300
        e = float((np.abs(a - b) / scale).max())
        worst = max(worst, e)
        print(f"[gpu]   level {lev}: {int(v.sum())} valid cells, normalised max difference {e:.3e}", flush=True)
    assert worst < tol
    return worst


def test_one_box_equals_two_level_driver(oracle_h2):
    """One interior fine box: the hierarchy driver and the two-level driver
    give the same bits."""
    print("[gpu] hierarchy vs two-level driver, one fine box, 2 steps", flush=True)
    mech = hf.Mechanism(H2, **CONVERGED)
    Q, dx, box = two_level_case(oracle_h2)
    two = TwoLevelDriver(mech, Q.shape[2], Q.shape[1], dx, dx, (REFLECTING,) * 4, box)
    two.set_state(Q)
    d = AmrDriver(mech, Q.shape[2], Q.shape[1], dx, dx, (REFLECTING,) * 4, max_level=1, blocking=4,
                  regrid_interval=0)
    d.set_grids([[], [box]])
    d.set_state(Q)
    for _ in range(2):
        assert two.step() == d.step()
    C, F = two.state()
    C2, F2 = d.level_state(0)[0], d.level_state(1)[0][:, box[1]:box[3] + 1, box[0]:box[2] + 1]
    for a, b, name in ((C2, C, "coarse"), (F2, F, "fine")):
        scale = np.abs(b).reshape(b.shape[0], -1).max(1)[:, None, None] + 1e-300
        e = np.abs(a - b) / scale
        k = np.unravel_index(np.argmax(e), e.shape)
        print(f"[gpu]   {name}: max normalised difference {e.max():.3e} at {k}; repairs two-level {two.repairs} "
              f"hierarchy {d.repairs}; first-order stages {two.first_order_stages} {d.first_order_stages}",
              flush=True)
        assert np.array_equal(a, b)


def test_three_levels_match_oracle_and_conserve(oracle_h2):
    """Two fine levels of two boxes each (boxes sharing sides, level 2
    nested in both level-1 boxes) in a closed domain: two Strang steps
    against the oracle, mass and energy kept to roundoff."""
    print("[gpu] three levels, closed domain, 2 steps vs oracle", flush=True)
    mech = hf.Mechanism(H2, **CONVERGED)
    Q, dx, _ = two_level_case(oracle_h2)
    d = AmrDriver(mech, Q.shape[2], Q.shape[1], dx, dx, (REFLECTING,) * 4, max_level=2, blocking=4,
                  regrid_interval=0)
    d.set_grids([[], LEVEL1, LEVEL2])
    d.set_state(Q)
    od = OracleAmrDriver(oracle_h2, Q.shape[2], Q.shape[1], dx, dx, (REFLECTING,) * 4, max_level=2)
    od.set_grids([[], LEVEL1, LEVEL2])
    od.set_state(Q)
    ns = mech.ns
    t0 = d.totals()
    for _ in range(2):
        h, oh = d.step(), od.step()
        assert abs(h - oh) <= 1e-12 * oh
    levels_close(d, od)
    t1 = d.totals()
    for a, b in ((t0[:ns].sum(), t1[:ns].sum()), (t0[ns + 2], t1[ns + 2])):
        assert abs(b - a) <= 1e-13 * abs(a)


def test_split_boxes_bit_identical():
    """The same fine levels as one set of boxes and cut into twice as many:
    identical states (ghost cells from the neighbouring box, no flux
    register across shared sides)."""
    print("[gpu] split boxes, 1 step", flush=True)
    from oracle import Oracle
    o = Oracle(H2)
    mech = hf.Mechanism(H2, **CONVERGED)
    Q, dx, _ = two_level_case(o)
    outs = []
    for l1, l2 in ((LEVEL1, LEVEL2), (_split(LEVEL1), _split(LEVEL2))):
        d = AmrDriver(mech, Q.shape[2], Q.shape[1], dx, dx, (REFLECTING,) * 4, max_level=2, blocking=4,
                      regrid_interval=0)
        d.set_grids([[], l1, l2])
        d.set_state(Q)
        d.step()
        outs.append([d.level_state(lev) for lev in range(3)])
    for (a, ma), (b, mb) in zip(*outs):
        assert (ma == mb).all()
        assert np.array_equal(a[:, ma > 0], b[:, mb > 0])


def small_channel(mech, nx=128, ny=32):
    """The channel workload's physics at the same cell size on a smaller
    domain: 128 x 32 base cells of 1.953125e-4 m, hot strip x < 0.01 m."""
    dx = 0.2 / 1024
    Q, _ = channel_2d(mech, nx=nx, ny=ny, length=nx * dx)
    return Q, dx


def test_adaptive_channel_matches_oracle(oracle_h2):
    """The channel case (reflecting left, bottom and top, transmissive
    right) with two refinement levels chosen from ErrorEst's tags (blocking
    8, max_grid_size 32, grid_eff 0.7) and regridded every 2 steps: 4 Strang
    steps on the GPU; the oracle follows the same boxes (its remap from the
    GPU's regrid) and matches to 1e-8 per level."""
    print("[gpu] adaptive channel 128x32, 2 refinement levels, 4 steps, regrid every 2", flush=True)
    mech = hf.Mechanism(H2, **CONVERGED)
    Q, dx = small_channel(mech)
    bc = (hf.HF_BC_REFLECTING, hf.HF_BC_TRANSMISSIVE, hf.HF_BC_REFLECTING, hf.HF_BC_REFLECTING)
    d = AmrDriver(mech, Q.shape[2], Q.shape[1], dx, dx, bc, max_level=2, blocking=8, max_grid_size=32,
                  regrid_interval=2)
    d.initial_grids(Q)
    grids = d.grids()
    print(f"[gpu]   initial boxes per level: {[len(g) for g in grids]}; finest {d.finest}", flush=True)
    assert d.finest == 2 and len(grids[1]) >= 2
    od = OracleAmrDriver(oracle_h2, Q.shape[2], Q.shape[1], dx, dx, bc, max_level=2)
    od.set_grids(grids)
    od.set_state(Q)
    levels_close(d, od, 1e-13)
    for k in range(4):
        h, oh = d.step(), od.step()
        assert abs(h - oh) <= 1e-10 * oh
        if d.steps % 2 == 0 and d.grids() != od_grids(od):
            defect = od.remap(d.grids())
            print(f"[gpu]   regrid after step {d.steps}: boxes {[len(g) for g in d.grids()]}, oracle defect "
                  f"{defect:.2e}", flush=True)
        levels_close(d, od)
    # a regrid that moves boxes: finer temperature criterion, new boxes, remap
    d.tag["temperature_tag"] = d.tag["pressure_tag"] = 0.01
    before = d.grids()
    d.regrid()
    assert d.grids() != before
    defect = od.remap(d.grids())
    print(f"[gpu]   forced regrid: boxes {[len(g) for g in before]} -> {[len(g) for g in d.grids()]}, defect GPU "
          f"{d.max_remap_defect:.2e} oracle {defect:.2e}", flush=True)
    levels_close(d, od)
    h, oh = d.step(), od.step()
    levels_close(d, od)
    assert d.max_remap_defect < 1e-12


def od_grids(od):
    return [[]] + [list(od.L[lev]["boxes"]) for lev in range(1, od.finest + 1)]


def test_channel_workload_hierarchy():
    """The channel workload itself (1024 x 256 base, two refinement levels,
    blocking 16, max_grid_size 64): the tags give ragged boxes of 16 to 64
    cells per side on both fine levels; two Strang steps with a regrid keep
    the hierarchy admissible and nested, every remap conservative."""
    print("[gpu] channel 1024x256, 2 refinement levels, 2 steps + regrid", flush=True)
    mech = hf.Mechanism(H2)
    Q, dx = channel_2d(mech)
    bc = (hf.HF_BC_REFLECTING, hf.HF_BC_TRANSMISSIVE, hf.HF_BC_REFLECTING, hf.HF_BC_REFLECTING)
    d = AmrDriver(mech, 1024, 256, dx, dx, bc, max_level=2, blocking=16, max_grid_size=64, regrid_interval=2)
    d.initial_grids(Q)
    for _ in range(2):
        d.step()
    g = d.grids()
    sizes = [(b[2] - b[0] + 1, b[3] - b[1] + 1) for lev in (1, 2) for b in g[lev]]
    print(f"[gpu]   boxes per level {[len(x) for x in g]}, leaves {d.leaf_count()}, sizes "
          f"{sorted(set(sizes))}, regrids {d.regrids} (changed {d.grid_changes}), defect "
          f"{d.max_remap_defect:.2e}", flush=True)
    assert d.finest == 2 and d.regrids == 1 and len(g[1]) >= 2 and len(g[2]) >= 2
    assert all(16 <= w <= 64 and 16 <= h <= 64 for w, h in sizes)
    assert d.max_remap_defect < 1e-10


def _graph_vs_host(mech, Q, dx, bc, steps, cfl, **kw):
    """The same steps with each step one HIP graph (hf_amr_strang_step) and
    call by call: (graph driver, host driver, dts of both)."""
    runs = []
    for graph in (True, False):
        d = AmrDriver(mech, Q.shape[2], Q.shape[1], dx, dx, bc, cfl=cfl, graph=graph, **kw)
        d.initial_grids(Q)
        dts = [d.step() for _ in range(steps)]
        dts.append(d.step(target=d.time + 0.25 * dts[-1]))
        runs.append((d, dts))
    (g, dg), (e, de) = runs
    print(f"[gpu]   graph steps {g.graph_steps} of {g.steps}; host driver: first-order stages "
          f"{e.first_order_stages}, retries {e.retries}, repairs {e.repairs}; grids {[len(b) for b in g.grids()]}",
          flush=True)
    return g, e, dg, de


def _assert_same_bits(g, e):
    assert g.grids() == e.grids()
    assert g.repairs == e.repairs and g.regrids == e.regrids
    for lev in range(g.finest + 1):
        (a, ma), (b, mb) = g.level_state(lev), e.level_state(lev)
        assert (ma == mb).all()
        assert np.array_equal(a[:, ma > 0], b[:, mb > 0]), f"level {lev}"


def test_graph_steps_equal_host_steps():
    """Whole hierarchy steps as replayed HIP graphs (recaptured when a regrid
    changes the boxes) give the same bits as the same steps call by call: the
    adaptive channel with two refinement levels, regrids every 2 steps, and a
    step to a target time."""
    print("[gpu] hierarchy graph step vs host step, channel 128x32, 2 refinement levels, 7 steps", flush=True)
    mech = hf.Mechanism(H2, **CONVERGED)
    Q, dx = small_channel(mech)
    bc = (hf.HF_BC_REFLECTING, hf.HF_BC_TRANSMISSIVE, hf.HF_BC_REFLECTING, hf.HF_BC_REFLECTING)
    g, e, dg, de = _graph_vs_host(mech, Q, dx, bc, 6, 0.5, max_level=2, blocking=8, max_grid_size=32,
                                  regrid_interval=2)
    assert dg == de
    _assert_same_bits(g, e)
    assert e.graph_steps == 0
    if e.first_order_stages == 0 and e.retries == 0:
        assert g.graph_steps == g.steps == 7


def test_graph_step_hands_back_failed_steps():
    """A chemistry batch that fails inside the graph (max_steps half the
    step count the first step needs): the graph restores every level and
    the call-by-call step (dt halved until the batch converges) takes it,
    with the same bits as the host driver throughout."""
    print("[gpu] hierarchy graph step with a failing chemistry batch: hand-back to the host step", flush=True)
    mech = hf.Mechanism(H2, **CONVERGED)
    Q, dx = small_channel(mech)
    bc = (hf.HF_BC_REFLECTING, hf.HF_BC_TRANSMISSIVE, hf.HF_BC_REFLECTING, hf.HF_BC_REFLECTING)
    kw = dict(max_level=2, blocking=8, max_grid_size=32, regrid_interval=0)
    probe = AmrDriver(mech, Q.shape[2], Q.shape[1], dx, dx, bc, graph=False, **kw)
    probe.initial_grids(Q)
    probe.step()
    most = probe.last_chem_stats["max_cell_steps"]
    print(f"[gpu]   first step: at most {most} integrator steps per cell; max_steps {max(most // 2, 8)}", flush=True)
    tight = hf.Mechanism(H2, max_steps=max(most // 2, 8), **CONVERGED)
    g, e, dg, de = _graph_vs_host(tight, Q, dx, bc, 2, 0.5, **kw)
    assert dg == de
    _assert_same_bits(g, e)
    assert e.retries > 0
    assert g.graph_steps < g.steps
    assert g.first_order_stages == e.first_order_stages and g.retries == e.retries


def test_totals_over_leaves():
    """hf_totals_cells (BlockAMR::totals) over every box of three levels:
    the leaf cells' values times the cell area, plain and absolute, against
    numpy sums of the same cells to 1e-13."""
    print("[gpu] hierarchy totals over the leaves, three levels", flush=True)
    mech = hf.Mechanism(H2, **CONVERGED)
    Q, dx = small_channel(mech)
    bc = (hf.HF_BC_REFLECTING, hf.HF_BC_TRANSMISSIVE, hf.HF_BC_REFLECTING, hf.HF_BC_REFLECTING)
    d = AmrDriver(mech, Q.shape[2], Q.shape[1], dx, dx, bc, max_level=2, blocking=8, max_grid_size=32)
    d.initial_grids(Q)
    d.step()
    assert d.finest == 2
    ref = {False: np.zeros(mech.nv), True: np.zeros(mech.nv)}
    for lev in range(d.finest + 1):
        A, m = d.level_state(lev)
        L = d.levels[lev]
        for absolute in (False, True):
            v = np.abs(A) if absolute else A
            ref[absolute] += v[:, m == 1].sum(axis=1) * (L.dx * L.dy)
    for absolute in (False, True):
        got = d.totals(absolute)
        # relative to the integral of |value| (the momenta sum to nearly zero)
        err = np.abs(got - ref[absolute]) / (ref[True] + 1e-300)
        print(f"[gpu]   absolute={absolute}: worst relative difference {err.max():.2e}", flush=True)
        assert err.max() < 1e-13
