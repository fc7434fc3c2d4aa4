"""ChemistryBatch semantics of the oracle (chemistry_batch.cpp:21-50)."""
import numpy as np

from tests.conftest import GOLDEN


def cells(n=12):
    g = np.load(GOLDEN / "synth_h2_seed2.npz")
    return g["U"][:, :n].copy(), g["dt"][:n].copy()


def test_count_zero_and_invalid_dt(oracle_h2):
    U, dt = cells()
    r, *_ = oracle_h2.chem_advance(U[:, :0].copy(), 1e-7)
    assert r == 0
    r, st, ff, msg = oracle_h2.chem_advance(U, -1.0)
    assert r == 1 and "invalid" in msg
    r, st, ff, msg = oracle_h2.chem_advance(U, np.nan)
    assert r == 1


def test_all_or_nothing_lowest_index(oracle_h2):
    U, dt = cells()
    U[0, 7] = np.nan
    U[3, 5] = -1.0
    before = U.copy()
    r, st, ff, msg = oracle_h2.chem_advance(U, dt)
    assert r == 2 and ff == 5 and msg.startswith("chemistry cell 5")
    assert np.array_equal(U, before, equal_nan=True)


def test_max_steps_failure(oracle_h2):
    U, dt = cells()
    before = U.copy()
    r, st, ff, msg = oracle_h2.chem_advance(U, dt, maxsteps=2)
    assert r == 3 and ff >= 0
    assert np.array_equal(U, before)


def test_element_and_mass_conservation(oracle_h2):
    U, dt = cells()
    m = oracle_h2.m
    before = U.copy()
    r, *_ = oracle_h2.chem_advance(U, dt, rtol=1e-10, atol=1e-14)
    assert r == 0
    el0 = (before[:9] / m["W"][:, None]).T @ m["comp"]
    el1 = (U[:9] / m["W"][:, None]).T @ m["comp"]
    assert np.all(np.abs(el1 - el0) <= 1e-12 * np.abs(el0))
    assert np.allclose(U[:9].sum(0), before[:9].sum(0), rtol=1e-14)
    assert np.array_equal(U[9:], before[9:])          # momenta and energy untouched


def test_thread_count_invariance(oracle_h2):
    U1, dt = cells(24)
    U8 = U1.copy()
    assert oracle_h2.chem_advance(U1, dt, threads=1)[0] == 0
    assert oracle_h2.chem_advance(U8, dt, threads=8)[0] == 0
    assert np.array_equal(U1, U8)
