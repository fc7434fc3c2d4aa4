"""GPU single-box Strang driver (hyperflow_amd.driver.BoxDriver over the C
ABI) against the oracle's restatement of the same loop (oracle/driver.py),
and its building blocks one by one.

Tolerances: ghost fill exact; roundoff repair to 1e-15 of the donor; the
driven state after a few Strang steps to 1e-8 of each component's scale,
with both sides' chemistry run converged (rtol 1e-12, atol 1e-20) so that the
comparison tests the flow, splitting and boundary logic rather than two
integrators' truncation errors.
"""
import numpy as np
import pytest

import hyperflow_amd as hf
from hyperflow_amd.driver import (HF_BC_PERIODIC, HF_BC_REFLECTING, HF_BC_TRANSMISSIVE, BoxDriver,
                                  shock_tube_1d)
from oracle.driver import OracleBoxDriver, fill_ghosts, repair_roundoff

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("bc,one_d", [((HF_BC_REFLECTING, HF_BC_TRANSMISSIVE, -1, -1), True),
                                      ((HF_BC_TRANSMISSIVE, HF_BC_REFLECTING, HF_BC_REFLECTING, HF_BC_PERIODIC), False),
                                      ((HF_BC_PERIODIC, HF_BC_PERIODIC, HF_BC_TRANSMISSIVE, HF_BC_REFLECTING), False)])
def test_ghost_fill_matches_oracle(mech_h2, bc, one_d):
    print(f"[gpu] ghost fill bc={bc} one_d={one_d}", flush=True)
    nx, ny = 11, (1 if one_d else 7)
    d = BoxDriver(mech_h2, nx, ny, 1e-4, 0.0 if one_d else 1e-4, bc)
    G = np.random.default_rng(4).standard_normal(tuple(d.U.shape))
    d.U.copy_(torch.from_numpy(G))
    d._fill(d.U)
    fill_ghosts(G, mech_h2.ns, bc, one_d)
    out = d.U.cpu().numpy()
    rows = slice(2, ny + 2) if one_d else slice(None)
    np.testing.assert_array_equal(out[:, rows, :], G[:, rows, :])


def test_repair_roundoff_matches_oracle(mech_h2):
    print("[gpu] repair roundoff", flush=True)
    nx, ny, ns = 9, 6, mech_h2.ns
    d = BoxDriver(mech_h2, nx, ny, 1e-4, 1e-4, (HF_BC_TRANSMISSIVE,) * 4)
    Q = np.abs(np.random.default_rng(5).standard_normal((mech_h2.nv, ny, nx)))
    eps = np.finfo(float).eps
    for n, cells in ((2, [(1, 1), (4, 7)]), (5, [(0, 0)]), (7, [(5, 8), (2, 3), (3, 3)])):
        for (j, i) in cells:
            Q[n, j, i] = -(3 + j) * eps * Q[n].max()
    d.set_state(Q)
    from ctypes import byref, c_int64
    from hyperflow_amd._lib import load
    box = d._box()
    rep, ff = c_int64(0), c_int64(-1)
    assert load().hf_repair_roundoff(d._h, d.U.data_ptr(), byref(box), byref(rep), byref(ff)) == 0
    n_or = repair_roundoff(Q, ns)
    assert rep.value == n_or == 6
    np.testing.assert_allclose(d.state(), Q, rtol=1e-15, atol=0)
    Q2 = d.state()
    Q2[1, 2, 2] = -1e-9 * Q2[1].max()
    d.set_state(Q2)
    r = load().hf_repair_roundoff(d._h, d.U.data_ptr(), byref(box), byref(rep), byref(ff))
    assert r == hf._lib.HF_EDOMAIN and ff.value == 2 * nx + 2


def _compare(a, b, rel, ns):
    """Per-component error over the component's scale; species scales are
    floored at 1e-6 of the largest density (values below the chemistry's
    atol carry no information, as in the chemistry parity tests)."""
    scale = np.abs(b).max(axis=tuple(range(1, b.ndim)), keepdims=True)
    rho = b[:ns].sum(0).max()
    scale[:ns] = np.maximum(scale[:ns], 1e-6 * rho)
    err = np.abs(a - b) / np.maximum(scale, 1e-300)
    assert err.max() <= rel, (err.max(), np.unravel_index(err.argmax(), err.shape))


def test_shock_tube_1d_matches_oracle(mech_h2, oracle_h2):
    """The 1-D workload (2000 cells, 0.1 m, hot equilibrium slab at 2000 K /
    20 atm next to a reflecting wall, transmissive outflow): 3 Strang steps."""
    print("[gpu] 1-D Strang driver: 2000 cells, 3 steps", flush=True)
    Q, dx = shock_tube_1d(mech_h2)
    m = hf.Mechanism(hf.H2_LI, rtol=1e-12, atol=1e-20)
    g = BoxDriver(m, 2000, 1, dx, 0.0, (HF_BC_REFLECTING, HF_BC_TRANSMISSIVE, -1, -1), cfl=0.5)
    o = OracleBoxDriver(oracle_h2, 2000, 1, dx, 0.0, (HF_BC_REFLECTING, HF_BC_TRANSMISSIVE, -1, -1), cfl=0.5)
    g.set_state(Q)
    o.Q[:] = Q
    for _ in range(3):
        dtg, dto = g.step(), o.step()
        assert dtg == pytest.approx(dto, rel=1e-12)
    print(f"[gpu] 1-D driver: t={g.time:.4e} first-order stages {g.first_order_stages}/{o.first_order_stages} "
          f"repairs {g.repairs} retries {g.retries}/{o.retries}", flush=True)
    _compare(g.state(), o.Q, 1e-8, mech_h2.ns)


def test_box_2d_matches_oracle(mech_h2, oracle_h2):
    """A 2-D box: hot high-pressure square in a cold mixture, reflecting and
    transmissive x sides, reflecting / periodic y: 2 Strang steps."""
    print("[gpu] 2-D Strang driver: 40 x 24 box, 2 steps", flush=True)
    nx, ny, dx = 40, 24, 1e-4
    bc = (HF_BC_TRANSMISSIVE, HF_BC_REFLECTING, HF_BC_REFLECTING, HF_BC_PERIODIC)
    Q1, _ = shock_tube_1d(mech_h2, nx=4, length=0.04)   # cell 0 hot (x < 0.01 m), cell 3 cold
    cold, hot = Q1[:, 0, 3], Q1[:, 0, 0]
    assert hot[:mech_h2.ns].sum() != cold[:mech_h2.ns].sum()
    Q = np.repeat(np.repeat(cold[:, None, None], ny, 1), nx, 2)
    Q[:, 8:16, 10:20] = hot[:, None, None]
    m = hf.Mechanism(hf.H2_LI, rtol=1e-12, atol=1e-20)
    g = BoxDriver(m, nx, ny, dx, dx, bc, cfl=0.5)
    o = OracleBoxDriver(oracle_h2, nx, ny, dx, dx, bc, cfl=0.5)
    g.set_state(Q)
    o.Q[:] = Q
    for _ in range(2):
        dtg, dto = g.step(), o.step()
        assert dtg == pytest.approx(dto, rel=1e-12)
    _compare(g.state(), o.Q, 1e-8, mech_h2.ns)


@pytest.mark.parametrize("one_d", [True, False])
def test_fused_euler_stage_equals_separate_calls(mech_h2, one_d):
    """hf_euler_stage (fill, flux stage, repair, validation in one call) gives
    the same bits as the four calls made one by one."""
    from ctypes import byref, c_int32, c_int64
    from hyperflow_amd._lib import load
    print(f"[gpu] fused Euler stage, one_d={one_d}", flush=True)
    if one_d:
        Q, dx = shock_tube_1d(mech_h2, nx=300, length=0.015)
        nx, ny, dy, bc = 300, 1, 0.0, (HF_BC_REFLECTING, HF_BC_TRANSMISSIVE, -1, -1)
    else:
        Q1, dx = shock_tube_1d(mech_h2, nx=4, length=0.04)
        cold, hot = Q1[:, 0, 3], Q1[:, 0, 0]
        nx, ny, dy = 30, 20, dx
        Q = np.repeat(np.repeat(cold[:, None, None], ny, 1), nx, 2)
        Q[:, 5:12, 8:19] = hot[:, None, None]
        bc = (HF_BC_TRANSMISSIVE, HF_BC_REFLECTING, HF_BC_REFLECTING, HF_BC_PERIODIC)
    d = BoxDriver(mech_h2, nx, ny, dx, dy, bc)
    d.set_state(Q)
    h = 0.4 * d.stable_dt() / d.cfl
    A = d.U.clone()
    d.euler(d.U, d.U1, h)                                  # fused
    fused = d.U1.clone()
    # the same stage call by call
    box = d._box(h)
    L, hdl = load(), d._h
    assert L.hf_fill_ghosts(hdl, A.data_ptr(), byref(box), d.bc) == 0
    out = torch.zeros_like(A)
    from hyperflow_amd._lib import hf_flux_stats
    st, ff = hf_flux_stats(), c_int64(-1)
    L.hf_flux_stage(hdl, byref(box), 1, A.data_ptr(), out.data_ptr(), 0, byref(st), byref(ff))
    rep = c_int64(0)
    assert L.hf_repair_roundoff(hdl, out.data_ptr(), byref(box), byref(rep), byref(ff)) == 0
    assert L.hf_validate_box(hdl, out.data_ptr(), byref(box), dx, dy, None, byref(ff)) == 0
    assert torch.equal(fused[:, 2:ny + 2, 2:nx + 2], out[:, 2:ny + 2, 2:nx + 2])
    # a materially negative species: both attempts are rejected, the cell is named
    B = d.U.clone()
    B[1, 2 + (ny // 2 if not one_d else 0), 2 + 7] = -1e-3 * float(B[:mech_h2.ns, 2, 2].sum())
    fo = c_int32(0)
    r = L.hf_euler_stage(hdl, B.data_ptr(), d.U2.data_ptr(), byref(box), d.bc, None, dx, dy, byref(rep), byref(fo),
                         byref(ff))
    assert r == hf._lib.HF_EDOMAIN and fo.value == 1
    assert ff.value == (ny // 2 if not one_d else 0) * nx + 7


@pytest.mark.parametrize("case", ["shock_tube", "channel"])
def test_graph_steps_equal_host_steps(mech_h2, case):
    """Whole Strang steps replayed as one HIP graph (hf_box_strang_step) give
    the same bits as the same steps taken call by call, including a step to a
    target time; after the first capture every step is a replay."""
    print(f"[gpu] graph step vs host step, {case}", flush=True)
    if case == "shock_tube":
        Q, dx = shock_tube_1d(mech_h2, nx=400, length=0.02, hot_x=0.002)
        args = (Q.shape[2], 1, dx, 0.0, (HF_BC_REFLECTING, HF_BC_TRANSMISSIVE, -1, -1))
    else:
        from hyperflow_amd.driver import channel_2d
        Q, dx = channel_2d(mech_h2, nx=96, ny=32, length=96 * 0.2 / 1024)
        args = (96, 32, dx, dx, (HF_BC_REFLECTING, HF_BC_TRANSMISSIVE, HF_BC_REFLECTING, HF_BC_REFLECTING))
    runs = []
    for graph in (False, True):
        d = BoxDriver(mech_h2, *args, graph=graph)
        d.set_state(Q)
        dts = [d.step() for _ in range(6)]
        dts.append(d.step(target=d.time + 0.25 * dts[-1]))
        runs.append((d.state(), dts, d))
    (a, da, ha), (b, db, g) = runs
    print(f"[gpu]   graph steps {g.graph_steps}, host steps {g.host_steps}, last {getattr(g, 'last_graph', None)}",
          flush=True)
    assert da == db
    assert np.array_equal(a, b)
    assert g.graph_steps == 7 and g.host_steps == 0 and g.last_graph == 1
    assert ha.graph_steps == 0
