"""The oracle's fixed-step backward-Euler objective (the function the adjoint
differentiates; no reference counterpart, so it is pinned by the mathematics):

* energy identity: at fixed rho and e, rho cv dT/dt = -sum_k u_k wdot_k = qdot,
  so one tiny step satisfies h qdot(y_1) = rho cv (T_1 - T_0) to first order;
* first-order convergence of T(t_end) to the converged BDF solution of the
  same cell as the step count doubles;
* the ln A shift enters the rates: a zero shift reproduces the unshifted
  objective exactly, a shift of H + O2 <=> O + OH moves it in every cell.
"""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def cells(mech_h2, oracle_h2):
    U, _ = mech_h2.synth_cells(2, 5, 16, with_dt=False)
    ns = mech_h2.ns
    rho = U[:ns].sum(0)
    Y0 = (U[:ns] / rho).T.copy()
    _, _, T0, _ = oracle_h2.properties(U.T.copy())
    return U, rho, T0, Y0


def test_single_step_heat_release_identity(oracle_h2, cells):
    U, rho, T0, Y0 = cells
    # move to reacting states first
    r, _, Ta, Ya = oracle_h2.chem_fixed_be(rho, T0, Y0, 1e-8, 150)
    assert r == 0
    r, J1, T1, Y1 = oracle_h2.chem_fixed_be(rho, Ta, Ya, 1e-9, 1)
    assert r == 0
    cv = np.array([oracle_h2.e_cv(T1[c], Y1[c])[1] for c in range(len(rho))])
    sel = np.abs(T1 - Ta) > 1e-6      # cells with measurable heat release in the step
    assert sel.sum() >= 8
    ratio = (J1 - T1)[sel] / (rho * cv * (T1 - Ta))[sel]
    np.testing.assert_allclose(ratio, 1.0, rtol=1e-4)


def test_backward_euler_converges_first_order(oracle_h2, cells):
    U, rho, T0, Y0 = cells
    t_end = 2e-6
    ref = U.copy()
    r, *_ = oracle_h2.chem_advance(ref, np.full(len(rho), t_end), rtol=1e-12, atol=1e-20)
    assert r == 0
    _, _, Tref, _ = oracle_h2.properties(ref.T.copy())
    err = []
    for S in (100, 200, 400):
        r, _, Te, _ = oracle_h2.chem_fixed_be(rho, T0, Y0, t_end / S, S)
        assert r == 0
        err.append(np.abs(Te - Tref))
    sel = err[0] > 1e-3
    assert sel.sum() >= 4
    assert np.all(err[1][sel] < err[0][sel]) and np.all(err[2][sel] < err[1][sel])
    rate = err[0][sel] / err[2][sel]          # ~4 for first order over two halvings
    assert np.all((rate > 2.5) & (rate < 6.5)), rate


def test_lnA_shift_enters_the_rates(oracle_h2, mech_h2, cells):
    _, rho, T0, Y0 = cells
    base = oracle_h2.chem_fixed_be(rho, T0, Y0, 1e-8, 20)[1]
    zero = oracle_h2.chem_fixed_be(rho, T0, Y0, 1e-8, 20, lnA_shift=np.zeros(mech_h2.nr))[1]
    np.testing.assert_array_equal(zero, base)
    shifted = np.zeros(mech_h2.nr)
    shifted[0] = 1e-3
    moved = oracle_h2.chem_fixed_be(rho, T0, Y0, 1e-8, 20, lnA_shift=shifted)[1]
    assert np.all(np.isfinite(moved))
    # reaction 0 (H + O2 <=> O + OH) fires in every hot cell: J must move
    assert np.all(moved != base)
