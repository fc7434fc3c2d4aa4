"""Pins of the oracle's kinetics (net production rates)."""
import numpy as np
import pytest

from oracle import thermo_np as tn
from oracle.mech import R_GAS


def _random_states(m, n, seed):
    rng = np.random.default_rng(seed)
    T = rng.uniform(600, 3200, n)
    C = rng.random((n, m["ns"])) ** 3 * rng.uniform(0.001, 5.0, (n, 1))
    return T, C


@pytest.mark.parametrize("which", ["h2", "dme"])
def test_element_conservation(which, oracle_h2, oracle_dme):
    o = oracle_h2 if which == "h2" else oracle_dme
    m = o.m
    T, C = _random_states(m, 32, 3)
    for i in range(32):
        w = o.wdot(T[i], C[i])
        flux = np.abs(w)[:, None] * m["comp"]
        assert np.all(np.abs(w @ m["comp"]) <= 1e-13 * flux.sum(0).clip(min=1e-300))


@pytest.mark.parametrize("which", ["h2", "dme"])
def test_numpy_and_c_kinetics_agree(which, oracle_h2, oracle_dme):
    o = oracle_h2 if which == "h2" else oracle_dme
    m = o.m
    T, C = _random_states(m, 16, 4)
    w = tn.wdot(m, T, C)
    for i in range(16):
        wc = o.wdot(T[i], C[i])
        scale = np.abs(tn.rates_of_progress(m, T[i:i + 1], C[i:i + 1])).max()
        assert np.abs(wc - w[i]).max() <= 1e-12 * scale


def test_unit_conversion_known_answer(oracle_h2, oracle_dme):
    # R1 H + O2 <=> O + OH: 3.547e15 cm3/mol/s in the DME file == 3.547e12 m3/kmol/s in H2_Li
    assert oracle_dme.m["A"][0] == pytest.approx(3.547e12, rel=1e-15)
    assert oracle_h2.m["A"][0] == pytest.approx(3.547e12, rel=1e-15)
    # three-body R5 H2 + M <=> H + H + M: order 1 + M = 2, cm3/mol -> m3/kmol is 1e-3,
    # and the H2_Li file states the same rate in SI-kmol units
    assert oracle_dme.m["A"][4] == pytest.approx(4.577e19 * 1e-3, rel=1e-15)
    assert oracle_h2.m["A"][4] == pytest.approx(4.577e16, rel=1e-15)
    # Ea in cal/mol -> K
    assert oracle_h2.m["EaR"][0] == pytest.approx(1.6599e4 * 4184.0 / R_GAS, rel=1e-15)


def test_detailed_balance_at_equilibrium(oracle_h2, oracle_dme, mech_h2, mech_dme):
    """At a TP equilibrium (product solver, element potentials) every reaction's
    forward and reverse rates balance: pins Kc, the standard state and units."""
    cases = [(oracle_h2, mech_h2, {"H2": 2, "O2": 1, "N2": 3.76}, 2000.0, 20.0),
             (oracle_h2, mech_h2, {"H2": 4, "O2": 1, "N2": 3.76}, 2600.0, 1.0),
             (oracle_dme, mech_dme, {"CH3OCH3": 1, "O2": 3, "N2": 11.28}, 2200.0, 10.0)]
    for o, mech, comp, T, patm in cases:
        m = o.m
        X = np.zeros(m["ns"])
        for k, v in comp.items():
            X[m["species"].index(k)] = v
        X /= X.sum()
        Y = mech.equilibrate_tp(T, patm * 101325.0, X)
        wbar = 1.0 / np.sum(Y / m["W"])
        rho = patm * 101325.0 * wbar / (R_GAS * T)
        C = rho * Y / m["W"]
        q = tn.rates_of_progress(m, np.array([T]), C[None])[0]
        m_fwd = dict(m, rev=np.zeros_like(m["rev"]))
        fwd = tn.rates_of_progress(m_fwd, np.array([T]), C[None])[0]
        assert np.all(np.abs(q) <= 1e-9 * fwd + 1e-300), (np.abs(q) / fwd).max()


def _k_of(m, i, T, C, a, b):
    """Forward rate constant of reaction i with reactants a, b at concentrations C."""
    m1 = dict(m, rev=np.zeros_like(m["rev"]))
    q = tn.rates_of_progress(m1, np.array([T]), C[None])[0, i]
    return q / (C[a] * C[b])


def test_lindemann_falloff_limits(oracle_dme):
    """CO + O (+M) <=> CO2 (+M) has no Troe block: k -> k0 [M] as Pr -> 0 and
    k -> k_inf as Pr -> inf (Lindemann blending Pr/(1+Pr))."""
    m = oracle_dme.m
    i = next(j for j, e in enumerate(m["equations"]) if e.startswith("CO + O (+M)"))
    assert m["kind"][i] == 2 and m["troe"][i] == 0
    T = 1500.0
    kinf = m["A"][i] * np.exp(m["b"][i] * np.log(T) - m["EaR"][i] / T)
    k0 = m["A0"][i] * np.exp(m["b0"][i] * np.log(T) - m["Ea0R"][i] / T)
    a, b, n2 = (m["species"].index(s) for s in ("CO", "O", "N2"))
    for Mc, expect in ((1e-14, k0 * 1e-14), (1e14, kinf)):
        C = np.zeros(m["ns"])
        C[a] = C[b] = 1e-20
        C[n2] = Mc
        Meff = C @ m["eff"][i]
        exp_k = k0 * Meff if Mc < 1 else kinf
        assert _k_of(m, i, T, C, a, b) == pytest.approx(exp_k, rel=1e-9)


def test_troe_center_of_falloff(oracle_h2):
    """At the falloff centre, log10 Pr = 0.4 + 0.67 log10 Fcent, the Troe
    broadening equals Fcent: k = k_inf Pr/(1+Pr) Fcent (Troe 1983)."""
    m = oracle_h2.m
    i = int(np.nonzero(m["kind"] == 2)[0][0])          # H + O2 (+M) <=> HO2 (+M)
    T = 1200.0
    kinf = m["A"][i] * np.exp(m["b"][i] * np.log(T) - m["EaR"][i] / T)
    k0 = m["A0"][i] * np.exp(m["b0"][i] * np.log(T) - m["Ea0R"][i] / T)
    A, T3, T1 = m["troeA"][i], m["troeT3"][i], m["troeT1"][i]
    fcent = (1 - A) * np.exp(-T / T3) + A * np.exp(-T / T1)
    pr = 10 ** (0.4 + 0.67 * np.log10(fcent))
    a, b, n2 = (m["species"].index(s) for s in ("H", "O2", "N2"))
    C = np.zeros(m["ns"])
    C[a] = C[b] = 1e-12
    C[n2] = pr * kinf / k0 - (C @ m["eff"][i])
    assert _k_of(m, i, T, C, a, b) == pytest.approx(kinf * pr / (1 + pr) * fcent, rel=1e-12)
