"""Output files in the reference's formats (hyperflow_amd/output.py): the
headers are the reference's own (HyperFlow_V3/src/BlockAMR.cpp:80-81,
376-379), one row per active cell with 17-digit numbers, box extents, the
active-cell integrals, and the summary's key set and order
(BlockAMR.cpp:426-436)."""
import csv
import json

import numpy as np
import pytest

from hyperflow_amd import output as O

SPECIES = ["H2", "O2", "H2O", "H", "O", "OH", "HO2", "H2O2", "N2"]


def level(active=None):
    ns = len(SPECIES)
    Q = np.zeros((ns + 3, 2, 3))
    for j in range(2):
        for i in range(3):
            Q[:ns, j, i] = 0.125 * (1 + i + 2 * j) / ns     # rho = (1 + i + 2j) / 8
            Q[ns, j, i] = 2.0 * Q[:ns, j, i].sum()           # u = 2
            Q[ns + 1, j, i] = -1.0 * Q[:ns, j, i].sum()      # v = -1
            Q[ns + 2, j, i] = 1000.0 + i
    p = np.full((2, 3), 2 * O.ATM)
    T = np.full((2, 3), 300.0)
    return O.Level(0, Q, p, T, (4, 6), 0.5, 0.25, (-1.0, 0.0), active)


def test_headers_are_the_reference_ones(tmp_path):
    O.write_snapshot(tmp_path, 3, 1.5e-6, [level()], SPECIES)
    head = (tmp_path / "state_3.csv").read_text().splitlines()[0]
    assert head == "t,level,i,j,x,y,dx,dy,rho,u,v,p,p_atm,T,rhoE," + ",".join("Y_" + s for s in SPECIES)
    assert (tmp_path / "boxes_3.csv").read_text().splitlines()[0] == "level,x_left,x_right,y_bottom,y_top,dx,dy"
    h = O.History(tmp_path)
    h.close()
    assert (tmp_path / "history.csv").read_text().strip() == (
        "step,t,dt,finest_level,leaf_cells,stored_cells,dx_min,dy_min,mass,momentum_x,momentum_y,total_energy,"
        "regrids,retries")


def test_state_rows_and_boxes(tmp_path):
    act = np.array([[True, False, True], [True, True, True]])
    O.write_snapshot(tmp_path, 0, 0.0, [level(act)], SPECIES)
    rows = list(csv.DictReader(open(tmp_path / "state_0.csv")))
    assert len(rows) == 5                                   # the inactive cell is left out
    r = rows[1]                                             # (i, j) = (2, 0) in box order -> level (6, 6)
    assert (r["i"], r["j"], r["level"]) == ("6", "6", "0")
    assert float(r["x"]) == -1.0 + 6.5 * 0.5 and float(r["y"]) == 6.5 * 0.25
    assert float(r["rho"]) == 0.375 and float(r["u"]) == 2.0 and float(r["v"]) == -1.0
    assert float(r["p_atm"]) == 2.0 and float(r["T"]) == 300.0 and float(r["rhoE"]) == 1002.0
    assert abs(sum(float(r["Y_" + s]) for s in SPECIES) - 1.0) < 1e-15
    box = list(csv.DictReader(open(tmp_path / "boxes_0.csv")))[0]
    assert [float(box[k]) for k in ("x_left", "x_right", "y_bottom", "y_top")] == [1.0, 2.5, 1.5, 2.0]
    # 17 significant digits
    assert r["x"] == format(-1.0 + 6.5 * 0.5, ".17g") and "e" not in r["i"]


def test_totals_and_history(tmp_path):
    L = level(np.array([[True, False, True], [True, True, True]]))
    s = O.totals([L])
    ns = len(SPECIES)
    rho = np.array([[1, 2, 3], [3, 4, 5]]) / 8.0
    act = L.active
    assert np.isclose(s[:ns].sum(), (rho[act]).sum() * 0.125, rtol=1e-15)
    assert np.isclose(s[ns], 2 * (rho[act]).sum() * 0.125, rtol=1e-15)
    h = O.History(tmp_path)
    h.log(7, 1e-6, 2e-9, [L], ns, retries=1)
    h.close()
    line = (tmp_path / "history.csv").read_text().splitlines()[1].split(",")
    assert line[0] == "7" and line[4] == "5" and line[5] == "6" and line[-1] == "1"
    assert float(line[8]) == pytest.approx(s[:ns].sum(), rel=1e-15)


def test_summary_keys(tmp_path):
    r = O.write_summary(tmp_path, completed=True, final_time=1e-6, steps=3, retries=0,
                        initial_conserved_integrals=[1.0, 2.0], final_conserved_integrals=[1.0, 2.0],
                        time_subcycling=False, refinement_ratio=[2, 2])
    data = json.loads((tmp_path / "summary.json").read_text())
    assert list(data) == ["completed", "final_time", "steps", "retries", "initial_conserved_integrals",
                          "final_conserved_integrals", "time_subcycling", "refinement_ratio"] and data == r
    with pytest.raises(ValueError):
        O.write_summary(tmp_path, not_a_key=1)


def test_wedge_case_setup_matches_the_reference_case_file(mech_h2, oracle_h2):
    """teng_osw_base against HyperFlow_V3/cases/teng_case1's stored values:
    u0 = 2588.5462044182495, v0 = -891.3079364573576 (the reference computes
    them from 2737.7 m/s at 19 degrees, BlockAMR.cpp:61-66), the wedge tip at
    x = 0 on the 144-cell base grid from x = -0.128 mm (column 8), inflow left
    and top, wedge bottom, outflow right (the case's boundary_lo / _hi)."""
    import hyperflow_amd as hf
    from hyperflow_amd.driver import teng_osw_base
    w = teng_osw_base(mech_h2)
    ns = mech_h2.ns
    q = w["inlet"]
    rho = q[:ns].sum()
    assert q[ns] / rho == pytest.approx(2588.5462044182495, rel=1e-15)
    assert q[ns + 1] / rho == pytest.approx(-891.3079364573576, rel=1e-15)
    assert w["wedge_i"] == 8 and w["dx"] == pytest.approx(1.6e-5, rel=1e-14)
    assert w["bc"] == (hf.HF_BC_INFLOW, hf.HF_BC_TRANSMISSIVE, hf.HF_BC_WEDGE, hf.HF_BC_INFLOW)
    p, _, T, st = oracle_h2.properties(q.reshape(1, -1).copy())
    assert st[0] == 0 and T[0] == pytest.approx(958.1, rel=1e-10) and p[0] == pytest.approx(1202900.0, rel=1e-10)
