"""The oracle's multi-level step loop (oracle/hierarchy.py, restating
BlockAMR::run / flow / euler / react / repairRoundoff over several levels
and boxes), pinned by invariants of the composite scheme:

* a closed domain (all walls) with two fine levels of several boxes each
  keeps mass and energy to roundoff through Strang steps (reflux and
  average-down over every box side that faces a coarse cell, none across
  sides two boxes share);
* a fine level cut into several boxes gives the same numbers as one box
  (ghost cells read from the neighbouring box);
* one fine box reproduces the two-level restatement (OracleTwoLevelDriver);
* a uniform gas at rest stays uniform on every level.
"""
import numpy as np

from oracle.driver import REFLECTING, TRANSMISSIVE, OracleTwoLevelDriver
from oracle.hierarchy import OracleAmrDriver
from tests.test_oracle_amr import fresh_Y, state, two_level_case

LEVEL1 = [(16, 12, 31, 35), (32, 12, 55, 27)]
LEVEL2 = [(40, 32, 79, 47), (40, 48, 55, 63)]


def totals(d):
    s = d.totals()
    ns = d.ns
    return np.array([s[:ns].sum(), s[ns + 2]])


def test_closed_domain_three_levels_conserve(oracle_h2):
    o = oracle_h2
    Q, dx, _ = two_level_case(o)
    d = OracleAmrDriver(o, Q.shape[2], Q.shape[1], dx, dx, (REFLECTING,) * 4, max_level=2)
    d.set_grids([[], LEVEL1, LEVEL2])
    assert d.finest == 2
    assert (d.L[1]["mask"] == 2).sum() == sum((b[2] - b[0] + 1) * (b[3] - b[1] + 1) for b in LEVEL2) // 4
    d.set_state(Q)
    before = totals(d)
    d.step()
    d.step()
    after = totals(d)
    assert d.time > 0
    assert (np.abs(after - before) / np.abs(before)).max() < 1e-13
    # the finest level moved on its own
    assert not np.allclose(d.interior(2)[o.ns + 2][32:48, 40:80], 0)


def _split(boxes):
    out = []
    for i0, j0, i1, j1 in boxes:
        im = (i0 + i1 + 1) // 2
        im -= im % 2
        out += [(i0, j0, im - 1, j1), (im, j0, i1, j1)]
    return out


def test_split_boxes_give_the_same_numbers(oracle_h2):
    o = oracle_h2
    Q, dx, _ = two_level_case(o)
    runs = []
    for l1, l2 in ((LEVEL1, LEVEL2), (_split(LEVEL1), _split(LEVEL2))):
        d = OracleAmrDriver(o, Q.shape[2], Q.shape[1], dx, dx, (REFLECTING,) * 4, max_level=2)
        d.set_grids([[], l1, l2])
        d.set_state(Q)
        d.step()
        runs.append([d.interior(lev).copy() for lev in range(3)])
    for lev in range(3):
        m = d.L[lev]["mask"] > 0
        a, b = runs[0][lev][:, m], runs[1][lev][:, m]
        scale = np.abs(a).max(axis=1, keepdims=True) + 1e-300
        assert (np.abs(a - b) / scale).max() < 1e-14


def test_one_box_reduces_to_two_level_restatement(oracle_h2):
    o = oracle_h2
    Q, dx, box = two_level_case(o)
    two = OracleTwoLevelDriver(o, Q.shape[2], Q.shape[1], dx, dx, (REFLECTING,) * 4, box, cfl=0.5)
    two.set_state(Q)
    multi = OracleAmrDriver(o, Q.shape[2], Q.shape[1], dx, dx, (REFLECTING,) * 4, max_level=1)
    multi.set_grids([[], [box]])
    multi.set_state(Q)
    for _ in range(2):
        assert two.step() == multi.step()
    C, F = multi.interior(0), multi.interior(1)[:, box[1]:box[3] + 1, box[0]:box[2] + 1]
    assert np.abs(C - two.Qc).max() <= 1e-15 * np.abs(two.Qc).max()
    assert np.abs(F - two.Qf).max() <= 1e-15 * np.abs(two.Qf).max()


def test_uniform_gas_stays_uniform(oracle_h2):
    o = oracle_h2
    q = state(o, 300.0, 101325.0, fresh_Y(o))
    Q = np.repeat(np.repeat(q[:, None, None], 24, 1), 40, 2)
    d = OracleAmrDriver(o, 40, 24, 1e-4, 1e-4, (TRANSMISSIVE, TRANSMISSIVE, REFLECTING, REFLECTING), max_level=2)
    d.set_grids([[], LEVEL1, LEVEL2])
    d.set_state(Q)
    d.step()
    # (the chemistry step makes radicals of order 1e-41 in the cold mixture,
    # the same in every cell)
    floor = 1e-12 * q[:o.ns].sum()
    first = None
    for lev in range(3):
        m = d.L[lev]["mask"] > 0
        V = d.interior(lev)[:, m]
        assert (np.abs(V - q[:, None]) <= 1e-12 * np.maximum(np.abs(q[:, None]), floor)).all()
        assert (V == V[:, :1]).all()
        first = V[:, 0] if first is None else first
        assert (V[:, 0] == first).all()
