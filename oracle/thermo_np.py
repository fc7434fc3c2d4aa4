"""numpy restatement of the reference's thermo/kinetics (TEST INFRASTRUCTURE ONLY).

An independent second restatement, vectorised over cells, used to cross-check
the C++ oracle (oracle/oracle.cpp) and, through it, the HIP kernels. It follows:
  * ``properties`` (HyperFlow_V3/src/physics.cpp:4-15): Y = rhoY/rho without
    renormalisation, T from e by Newton, p = rho R T / Wbar, a = sqrt(cp/cv R T / Wbar);
  * ``density`` / ``internalEnergy`` (HyperFlow_V3/src/solver.hpp:21-25);
  * ``hllc`` / ``minmod`` (HyperFlow_V3/src/physics.cpp:17-44);
  * net production rates as Cantera's GasKinetics computes them
    (called at HyperFlow_V3/src/chemistry.cpp:45): Arrhenius, three-body [M],
    Lindemann/Troe falloff, reverse rates from Kc with p0 = 1 atm;
  * the chemistry right-hand side of HyperFlow_V3/src/chemistry.cpp:36-53.
"""
from __future__ import annotations

import numpy as np

from .mech import R_GAS, ONE_ATM, THREE_BODY, FALLOFF

SMALL = 1e-300


def _nasa(m, T):
    """cp/R, h/RT, s/R per species; T shape (n,) -> (n, ns)."""
    T = np.asarray(T, float)[:, None]
    lowmask = T <= m["Tmid"][None, :]
    a = np.where(lowmask[..., None], m["low"][None], m["high"][None])
    a1, a2, a3, a4, a5, a6, a7 = (a[..., i] for i in range(7))
    cp = a1 + T * (a2 + T * (a3 + T * (a4 + T * a5)))
    h = a1 + T * (a2 / 2 + T * (a3 / 3 + T * (a4 / 4 + T * a5 / 5))) + a6 / T
    s = a1 * np.log(T) + T * (a2 + T * (a3 / 2 + T * (a4 / 3 + T * a5 / 4))) + a7
    return cp, h, s


def e_cv(m, T, Y):
    """Mass-specific internal energy and cv (J/kg, J/kg/K) for unnormalised Y."""
    cp, h, _ = _nasa(m, T)
    T = np.asarray(T, float)[:, None]
    yw = Y / m["W"][None, :]
    e = np.sum(yw * (h - 1.0) * R_GAS * T, axis=1)
    cv = np.sum(yw * (cp - 1.0) * R_GAS, axis=1)
    return e, cv


def temperature(m, e, Y, T0=None, tol=1e-14, itmax=100):
    T = np.full(len(e), 1000.0) if T0 is None else np.array(T0, float)
    for _ in range(itmax):
        ee, cv = e_cv(m, T, Y)
        dT = (ee - e) / cv
        T = T - dT
        if np.all(np.abs(dT) <= tol * T):
            break
    return T


def properties(m, U):
    """U: (n, ns+3) conserved [rhoY..., rhou, rhov, rhoE] -> p, a, T."""
    ns = m["ns"]
    rho = U[:, :ns].sum(axis=1)
    Y = U[:, :ns] / rho[:, None]
    e = U[:, ns + 2] / rho - 0.5 * ((U[:, ns] / rho) ** 2 + (U[:, ns + 1] / rho) ** 2)
    T = temperature(m, e, Y)
    wbar = 1.0 / np.sum(Y / m["W"][None, :], axis=1)
    cp, _, _ = _nasa(m, T)
    cpm = np.sum(Y / m["W"] * cp * R_GAS, axis=1)
    _, cvm = e_cv(m, T, Y)
    p = rho * R_GAS * T / wbar
    a = np.sqrt(cpm / cvm * R_GAS * T / wbar)
    return p, a, T


def minmod(a, b):
    return np.where(a * b > 0, np.copysign(np.minimum(np.abs(a), np.abs(b)), a), 0.0)


def hllc(m, L, R, n, pl, pr):
    """Vectorised HLLC with Davis wave speeds, V3 form; n = normal momentum index."""
    ns = m["ns"]
    E = ns + 2
    rl, rr = L[:, :ns].sum(1), R[:, :ns].sum(1)
    ul, ur = L[:, n] / rl, R[:, n] / rr
    (Pl, al), (Pr, ar) = pl, pr
    sl = np.minimum(ul - al, ur - ar)
    sr = np.maximum(ul + al, ur + ar)

    def flux(q, v, p):
        f = v[:, None] * q
        f[:, n] += p
        f[:, E] += p * v
        return f

    fl, fr = flux(L, ul, Pl), flux(R, ur, Pr)
    dl, dr = rl * (sl - ul), rr * (sr - ur)
    with np.errstate(divide="ignore", invalid="ignore"):
        sm = (Pr - Pl + dl * ul - dr * ur) / (dl - dr)
    left = sm >= 0
    q = np.where(left[:, None], L, R)
    v = np.where(left, ul, ur)
    s = np.where(left, sl, sr)
    p = np.where(left, Pl, Pr)
    with np.errstate(divide="ignore", invalid="ignore"):
        star = q * ((s - v) / (s - sm))[:, None]
        rs = star[:, :ns].sum(1)
        star[:, n] = rs * sm
        star[:, E] += rs * (sm - v) * (sm + p / (q[:, :ns].sum(1) * (s - v)))
        f = flux(q, v, p) + s[:, None] * (star - q)
    out = np.where((sl >= 0)[:, None], fl, np.where((sr <= 0)[:, None], fr, f))
    return out


def rates_of_progress(m, T, C, lnA_shift=None):
    """Net rate of progress q (n, nr) in kmol/m^3/s at temperature T, concentrations C."""
    T = np.asarray(T, float)[:, None]
    logT, invT = np.log(T), 1.0 / T
    shift = 0.0 if lnA_shift is None else lnA_shift[None, :]
    kf = m["A"][None] * np.exp(m["b"][None] * logT - m["EaR"][None] * invT + shift)
    M = C @ m["eff"].T                                 # (n, nr) third-body concentrations
    tb = m["kind"] == THREE_BODY
    fo = m["kind"] == FALLOFF
    kf = np.where(tb[None], kf * M, kf)
    if fo.any():
        k0 = m["A0"][None] * np.exp(m["b0"][None] * logT - m["Ea0R"][None] * invT + shift)
        with np.errstate(divide="ignore", invalid="ignore"):
            pr = M * k0 / (kf + SMALL)
        lpr = np.log10(np.maximum(pr, SMALL))
        A, T3, T1, T2 = m["troeA"], m["troeT3"], m["troeT1"], m["troeT2"]
        rt3 = np.where(np.abs(T3) < SMALL, np.inf, 1.0 / np.where(T3 == 0, 1, T3))
        rt1 = np.where(np.abs(T1) < SMALL, np.inf, 1.0 / np.where(T1 == 0, 1, T1))
        with np.errstate(over="ignore"):
            fcent = (1 - A)[None] * np.exp(-T * rt3[None]) + A[None] * np.exp(-T * rt1[None])
        fcent = fcent + np.where((T2 != 0)[None], np.exp(-T2[None] / T), 0.0)
        lfc = np.log10(np.maximum(fcent, SMALL))
        c = -0.4 - 0.67 * lfc
        nn = 0.75 - 1.27 * lfc
        f1 = (lpr + c) / (nn - 0.14 * (lpr + c))
        lgf = lfc / (1.0 + f1 * f1)
        F = np.where((m["troe"] > 0)[None], 10.0 ** lgf, 1.0)
        kfo = kf * pr / (1.0 + pr) * F
        kf = np.where(fo[None], kfo, kf)
    _, h, s = _nasa(m, T[:, 0])
    g = h - s                                          # g/RT
    nu = (m["nu_p"] - m["nu_r"]).astype(float)
    dG = g @ nu.T                                      # (n, nr)
    dn = nu.sum(1)[None]
    logc0 = np.log(ONE_ATM / (R_GAS * T))
    rkc = np.exp(dG - dn * logc0)                      # 1/Kc
    kr = np.where(m["rev"][None] > 0, kf * rkc, 0.0)
    fwd = kf * np.prod(np.where(m["nu_r"][None] > 0, C[:, None, :] ** m["nu_r"][None], 1.0), axis=2)
    rev = kr * np.prod(np.where(m["nu_p"][None] > 0, C[:, None, :] ** m["nu_p"][None], 1.0), axis=2)
    return fwd - rev


def wdot(m, T, C, lnA_shift=None):
    q = rates_of_progress(m, T, C, lnA_shift)
    nu = (m["nu_p"] - m["nu_r"]).astype(float)
    return q @ nu


def chem_rhs(m, y, rho, e, T0=None):
    """dy/dt for y = Y[0..ns-2] at fixed rho and e (HyperFlow_V3/src/chemistry.cpp:36-53)."""
    ns = m["ns"]
    Y = np.concatenate([y, 1.0 - y.sum(1, keepdims=True)], axis=1)
    T = temperature(m, e, Y, T0)
    C = rho[:, None] * Y / m["W"][None]
    w = wdot(m, T, C)
    return (w * m["W"][None] / rho[:, None])[:, : ns - 1], T
