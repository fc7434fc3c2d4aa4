"""Pins of the oracle's CVODE-equivalent BDF integrator and chemistry step."""
import numpy as np
import pytest

import oracle
from tests.conftest import GOLDEN


def test_linear_decoupled_closed_form():
    A = np.diag([-1.0, -50.0, -3000.0])
    for rtol in (1e-6, 1e-9):
        r, y, st = oracle.bdf_linear(A, [1.0, 2.0, 3.0], 1.0, rtol, rtol * 1e-6)
        assert r == 0
        exact = np.array([1.0, 2.0, 3.0]) * np.exp(np.diag(A))
        assert abs(y[0] - exact[0]) / exact[0] < 100 * rtol


def test_linear_coupled_closed_form():
    # y' = [[-2, 1], [1, -2]] y: eigenvalues -1, -3
    A = np.array([[-2.0, 1.0], [1.0, -2.0]])
    y0 = np.array([1.0, 0.0])
    t = 0.7
    exact = 0.5 * np.array([np.exp(-t) + np.exp(-3 * t), np.exp(-t) - np.exp(-3 * t)])
    r, y, st = oracle.bdf_linear(A, y0, t, 1e-10, 1e-14)
    assert r == 0 and np.allclose(y, exact, rtol=1e-8)


def test_robertson_reference_values():
    # Hairer & Wanner, Solving ODEs II: Robertson at t = 40
    ref = np.array([0.7158270687193135, 9.185534764557e-06, 0.2841637457372])
    r, y, st = oracle.bdf_robertson([1.0, 0.0, 0.0], 40.0, 1e-10, 1e-16)
    assert r == 0
    assert np.allclose(y, ref, rtol=1e-7)
    assert abs(y.sum() - 1.0) < 1e-12        # linear invariant of the system


def test_error_shrinks_with_tolerance():
    ref = np.array([0.7158270687193135, 9.185534764557e-06, 0.2841637457372])
    errs = []
    for rtol in (1e-5, 1e-7, 1e-9):
        r, y, st = oracle.bdf_robertson([1.0, 0.0, 0.0], 40.0, rtol, rtol * 1e-6)
        errs.append(np.abs(y - ref).max() / ref.max())
    assert errs[0] > errs[1] > errs[2]


def test_max_steps_is_a_convergence_failure():
    r, y, st = oracle.bdf_robertson([1.0, 0.0, 0.0], 40.0, 1e-10, 1e-16, maxsteps=10)
    assert r == 3


def test_chemistry_matches_radau_golden(oracle_h2):
    """Oracle (rtol 1e-10, atol 1e-14) vs Radau IIA at rtol 1e-12 on the
    0-D validation scenario and seeded workload cells: T within 1e-8
    relative; Y within 1e-6 relative above a floor of 1e-6."""
    g = np.load(GOLDEN / "chem_radau.npz")
    U = g["U0"].copy()
    r, st, ff, msg = oracle_h2.chem_advance(U, g["dt"], rtol=1e-10, atol=1e-14)
    assert r == 0, msg
    rho = U[:9].sum(0)
    Y = U[:9] / rho
    p, a, T, s = oracle_h2.properties(U.T.copy())
    assert np.max(np.abs(T - g["Tend"]) / g["Tend"]) < 1e-8
    assert np.max(np.abs(Y - g["Yend"]) / np.maximum(np.abs(g["Yend"]), 1e-6)) < 1e-6
    # the reference's 0-D check tolerances (validate.py:154): T < 1e-7, Y abs < 1e-8
    assert abs(T[0] - g["Tend"][0]) / g["Tend"][0] < 1e-7
    assert np.abs(Y[:, 0] - g["Yend"][:, 0]).max() < 1e-8


def test_equilibrium_is_a_fixed_point(oracle_h2, mech_h2):
    m = oracle_h2.m
    X = np.zeros(9)
    X[0], X[1], X[8] = 2, 1, 3.76
    X /= X.sum()
    Y = mech_h2.equilibrate_tp(2200.0, 5 * 101325.0, X)
    U = mech_h2.state_from_TPY(2200.0, 5 * 101325.0, Y)[:, None].copy()
    r, st, ff, msg = oracle_h2.chem_advance(U, 1e-5, rtol=1e-10, atol=1e-14)
    assert r == 0
    assert np.allclose(U[:9, 0] / U[:9, 0].sum(), Y, rtol=1e-6, atol=1e-12)
