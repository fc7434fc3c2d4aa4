"""The oracle's single-level step loop (oracle/driver.py, restating
BlockAMR::run / flow / euler / repairRoundoff / fill): its composition is
pinned by invariants of the discrete scheme:

* ghost fill: reflecting mirrors with the normal momentum negated,
  transmissive copies the edge cell, periodic wraps (checked cell by cell);
* repairRoundoff: roundoff negatives are zeroed and taken from the peak cell
  (each species total preserved exactly), material negatives raise;
* a periodic box conserves every species mass, both momenta and energy;
* a uniform gas at rest stays uniform next to walls and outflows;
* a mirror-symmetric state between two walls stays mirror-symmetric.
"""
import numpy as np
import pytest

from oracle.driver import PERIODIC, REFLECTING, TRANSMISSIVE, DomainError, OracleBoxDriver, fill_ghosts, repair_roundoff


def test_fill_ghosts_cell_by_cell():
    rng = np.random.default_rng(0)
    ns, nx, ny = 9, 6, 5
    G = rng.standard_normal((ns + 3, ny + 4, nx + 4))
    fill_ghosts(G, ns, (REFLECTING, TRANSMISSIVE, PERIODIC, REFLECTING), one_d=False)
    V = lambda n, i, j: G[n, j + 2, i + 2]
    for n in range(ns + 3):
        for j in range(ny):
            sgn = -1.0 if n == ns else 1.0
            assert V(n, -1, j) == sgn * V(n, 0, j) and V(n, -2, j) == sgn * V(n, 1, j)
            assert V(n, nx, j) == V(n, nx - 1, j) and V(n, nx + 1, j) == V(n, nx - 1, j)
        for i in range(-2, nx + 2):
            assert V(n, i, -1) == V(n, i, ny - 1) and V(n, i, -2) == V(n, i, ny - 2)
            sgn = -1.0 if n == ns + 1 else 1.0
            assert V(n, i, ny) == sgn * V(n, i, ny - 1) and V(n, i, ny + 1) == sgn * V(n, i, ny - 2)


def test_repair_roundoff_semantics():
    ns = 9
    Q = np.abs(np.random.default_rng(1).standard_normal((ns + 3, 4, 5)))
    Q[3, 0, 0] = Q[3].max() * 1.5            # the donor of species 3
    peak = Q[3].max()
    total = Q[3].sum()
    Q[3, 2, 1] = -10 * np.finfo(float).eps * peak
    Q[3, 3, 4] = -5 * np.finfo(float).eps * peak
    before = Q[3].sum()
    n = repair_roundoff(Q, ns)
    assert n == 2 and Q[3].min() >= 0
    assert Q[3].sum() == pytest.approx(before, rel=1e-15)
    assert Q[3, 0, 0] == pytest.approx(peak - 15 * np.finfo(float).eps * peak, rel=1e-15)
    assert total > 0
    Q[4, 1, 1] = -1e-6 * Q[4].max()
    with pytest.raises(DomainError):
        repair_roundoff(Q, ns)


def _mixture_state(mech, T, p_atm, u=0.0):
    X = np.zeros(mech.ns)
    for name, x in (("H2", 2.0), ("O2", 1.0), ("N2", 3.76)):
        X[mech.species.index(name)] = x
    Y = X * mech.W / (X * mech.W).sum()
    return mech.state_from_TPY(T, p_atm * 101325.0, Y, u)


def test_periodic_box_conserves(oracle_h2, mech_h2):
    nx, ny = 12, 8
    d = OracleBoxDriver(oracle_h2, nx, ny, 1e-4, 1e-4, (PERIODIC,) * 4)
    q = _mixture_state(mech_h2, 300.0, 1.0)
    d.Q[:] = q[:, None, None]
    d.Q[:, 3:6, 4:8] = _mixture_state(mech_h2, 600.0, 3.0, u=50.0)[:, None, None]
    ns = mech_h2.ns
    m = oracle_h2.m

    def totals(Q):
        # total mass, element amounts, both momenta, energy (chemistry moves
        # mass between species, so species totals are not invariants)
        mass = Q[:ns].sum()
        el = (Q[:ns].reshape(ns, -1).sum(1) / m["W"]) @ m["comp"]
        return np.concatenate([[mass], el, Q[ns:].sum(axis=(1, 2))])

    tot0 = totals(d.Q)
    for _ in range(2):
        d.step()
    tot1 = totals(d.Q)
    k = len(tot0) - 3
    scale = np.concatenate([np.abs(tot0[:k]), np.abs(d.Q[ns:]).max(axis=(1, 2)) * nx * ny])
    assert np.all(np.abs(tot1 - tot0) <= 1e-12 * scale)


def test_uniform_gas_stays_uniform(oracle_h2, mech_h2):
    d = OracleBoxDriver(oracle_h2, 40, 1, 1e-4, 0.0, (REFLECTING, TRANSMISSIVE, -1, -1))
    d.Q[:] = _mixture_state(mech_h2, 300.0, 1.0)[:, None, None]
    q0 = d.Q.copy()
    for _ in range(3):
        d.step()
    assert np.allclose(d.Q, q0, rtol=1e-12, atol=1e-12 * np.abs(q0).max())


def test_wall_mirror_symmetry(oracle_h2, mech_h2):
    nx = 40
    d = OracleBoxDriver(oracle_h2, nx, 1, 1e-4, 0.0, (REFLECTING, REFLECTING, -1, -1))
    d.Q[:] = _mixture_state(mech_h2, 300.0, 1.0)[:, None, None]
    d.Q[:, :, 16:24] = _mixture_state(mech_h2, 1500.0, 5.0)[:, None, None]
    for _ in range(3):
        d.step()
    mirror = d.Q[:, :, ::-1].copy()
    mirror[mech_h2.ns] *= -1.0
    scale = np.abs(d.Q).max(axis=(1, 2), keepdims=True)
    assert np.all(np.abs(d.Q - mirror) <= 1e-11 * scale)
