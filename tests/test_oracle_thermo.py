"""Pins of the oracle's thermodynamics (NASA7, e/cv, T(e), properties)."""
import numpy as np
import pytest

from oracle import thermo_np as tn
from oracle.mech import R_GAS

pytestmark = []


def _species_h_cp(m, T):
    cp, h, s = tn._nasa(m, np.atleast_1d(T))
    return cp[0], h[0], s[0]


@pytest.mark.parametrize("T", [400.0, 900.0, 1500.0, 2500.0])
def test_cp_is_dh_dT_and_ds_dT(oracle_h2, T):
    m = oracle_h2.m
    d = 1e-3
    cp, h, s = _species_h_cp(m, T)
    _, hp, sp = _species_h_cp(m, T + d)
    _, hm, sm = _species_h_cp(m, T - d)
    dh = ((hp * (T + d)) - (hm * (T - d))) / (2 * d)      # d(h/R)/dT
    ds = (sp - sm) / (2 * d)                               # d(s/R)/dT
    assert np.allclose(dh, cp, rtol=1e-8)
    assert np.allclose(ds, cp / T, rtol=1e-7)


def test_nasa_ranges_continuous_at_tmid(oracle_h2, oracle_dme):
    for o in (oracle_h2, oracle_dme):
        m = o.m
        for k in range(m["ns"]):
            T = m["Tmid"][k]
            lo, hi = m["low"][k], m["high"][k]
            cpl = sum(lo[i] * T ** i for i in range(5))
            cph = sum(hi[i] * T ** i for i in range(5))
            assert abs(cpl - cph) / cph < 2e-3, m["species"][k]


def test_known_heat_capacities(oracle_h2):
    # JANAF: N2 cp(300 K) = 29.12 J/mol/K, H2O cp(1000 K) = 41.27, H2 cp(2000 K) = 34.28
    m = oracle_h2.m
    for name, T, cp_ref in (("N2", 300.0, 29.12), ("H2O", 1000.0, 41.27), ("H2", 2000.0, 34.28)):
        cp, _, _ = _species_h_cp(m, T)
        k = m["species"].index(name)
        assert abs(cp[k] * R_GAS / 1000.0 - cp_ref) / cp_ref < 5e-3


def test_gamma_of_nitrogen_at_300K(oracle_h2):
    m = oracle_h2.m
    ns = m["ns"]
    Y = np.zeros(ns)
    Y[m["species"].index("N2")] = 1.0
    T, p = 300.0, 101325.0
    rho = p * m["W"][ns - 1] / (R_GAS * T)
    e, _ = oracle_h2.e_cv(T, Y)
    U = np.zeros(ns + 3)
    U[:ns] = rho * Y
    U[ns + 2] = rho * e
    P, a, TT, st = oracle_h2.properties(U)
    assert st[0] == 0
    gamma = a[0] ** 2 * rho / P[0]
    assert abs(gamma - 1.400) < 2e-3
    assert abs(TT[0] - T) < 1e-9 and abs(P[0] - p) / p < 1e-12


def test_energy_temperature_round_trip(oracle_h2, oracle_dme):
    rng = np.random.default_rng(1)
    for o in (oracle_h2, oracle_dme):
        ns = o.ns
        for _ in range(40):
            Y = rng.random(ns) ** 4
            Y /= Y.sum()
            T = rng.uniform(250, 3500)
            e, cv = o.e_cv(T, Y)
            assert abs(o.temperature(e, Y, T0=1000.0) - T) <= 1e-12 * T
            # also unnormalised mass fractions (the reference does not renormalise)
            Y2 = Y * 1.001
            e2, _ = o.e_cv(T, Y2)
            assert abs(o.temperature(e2, Y2) - T) <= 1e-12 * T


def test_numpy_and_c_restatements_agree(oracle_h2):
    m = oracle_h2.m
    rng = np.random.default_rng(2)
    n = 64
    Y = rng.random((n, m["ns"]))
    Y /= Y.sum(1, keepdims=True)
    T = rng.uniform(300, 3000, n)
    e, cv = tn.e_cv(m, T, Y)
    for i in range(n):
        ec, cvc = oracle_h2.e_cv(T[i], Y[i])
        assert abs(ec - e[i]) <= 1e-13 * max(abs(e[i]), cv[i] * T[i])
        assert abs(cvc - cv[i]) <= 1e-14 * cv[i]
    rho = rng.uniform(0.05, 20, n)
    U = np.zeros((n, m["ns"] + 3))
    U[:, : m["ns"]] = rho[:, None] * Y
    U[:, m["ns"]] = rho * rng.uniform(-800, 800, n)
    U[:, m["ns"] + 1] = rho * rng.uniform(-800, 800, n)
    U[:, m["ns"] + 2] = rho * (e + 0.5 * ((U[:, m["ns"]] / rho) ** 2 + (U[:, m["ns"] + 1] / rho) ** 2))
    p1, a1, T1 = tn.properties(m, U)
    p2, a2, T2, st = oracle_h2.properties(U)
    assert (st == 0).all()
    assert np.allclose(T1, T, rtol=1e-12) and np.allclose(T2, T, rtol=1e-12)
    assert np.allclose(p1, p2, rtol=1e-13) and np.allclose(a1, a2, rtol=1e-13)


def test_properties_rejects_inadmissible(oracle_h2):
    ns = oracle_h2.ns
    U = np.zeros((4, ns + 3))
    U[:, ns - 1] = 1.0
    U[:, ns + 2] = 2e5
    U[0, 0] = -1e-30          # negative partial density
    U[1, ns] = np.nan         # non-finite momentum
    U[2, :ns] = 0             # zero density
    U[3, ns + 2] = -1e12      # energy below any temperature
    _, _, _, st = oracle_h2.properties(U)
    assert (st == 2).all()
