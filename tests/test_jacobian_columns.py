"""The analytic Jacobian of the one-cell-per-wave chemistry kernel
(hf_wave.hpp WaveGas::jacobian, column lists built in topo_instance.hip
fill()) restated in numpy and checked against central differences of the
oracle's right-hand side (chemistry.cpp:36-53; the reference's CVODE forms
the same derivative by difference quotients, chemistry.cpp:15-25).

The restatement follows the kernel's algorithm step for step: per reaction,
the derivative of its net rate through each reactant and product slot,
through [M] and through T at fixed concentrations (falloff with Troe
broadening included; a zero Troe T1 or T3 is an infinite rate whose
derivative term vanishes); d wdot / d C assembled by concentration columns
from slot lists, the default third-body efficiency added per row; the chain
rule to the mass-fraction unknowns with the closing species and
dT/dy_s = -(e_s - e_last)/cv. A dropped term, a wrong sign or a transposed
index shows as a difference far above the difference quotient's
truncation (relative 1e-8 of the largest entry)."""
import numpy as np
import pytest

from oracle import mech, thermo_np as tn
from tests.conftest import DME, H2

R_GAS = 8314.462618
SMALL = 1e-300


def slots(m):
    """Per reaction: reactant and product slots (a species once per unit of
    its stoichiometric coefficient) and the listed efficiencies."""
    out = []
    for i in range(m["nr"]):
        reac = [k for k in range(m["ns"]) for _ in range(int(m["nu_r"][i, k]))]
        prod = [k for k in range(m["ns"]) for _ in range(int(m["nu_p"][i, k]))]
        out.append((reac, prod))
    return out


def efficiencies(m):
    """Default efficiency and (species, correction) pairs per third-body reaction."""
    dflt, listed = np.ones(m["nr"]), [[] for _ in range(m["nr"])]
    for i in range(m["nr"]):
        if m["kind"][i] == mech.ELEMENTARY:
            continue
        row = m["eff"][i]
        vals, counts = np.unique(row, return_counts=True)
        dflt[i] = vals[np.argmax(counts)]
        listed[i] = [(k, row[k] - dflt[i]) for k in range(m["ns"]) if row[k] != dflt[i]]
    return dflt, listed


def jacobian(m, y, rho, e):
    ns, nr = m["ns"], m["nr"]
    N = ns - 1
    W = m["W"]
    sl = slots(m)
    dflt, listed = efficiencies(m)
    Y = np.concatenate([y, [1 - y.sum()]])
    T = tn.temperature(m, np.array([e]), Y[None])[0]
    C = rho * Y / W
    logt, it = np.log(T), 1 / T
    logc0 = np.log(101325.0 / R_GAS) - logt
    cp, h, s = tn._nasa(m, np.array([T]))
    hrt, grt, cpr = h[0], h[0] - s[0], cp[0]
    ctot = C.sum()
    dslot = [None] * nr
    dM, dT = np.zeros(nr), np.zeros(nr)
    for i in range(nr):
        reac, prod = sl[i]
        kinf = m["A"][i] * np.exp(m["b"][i] * logt - m["EaR"][i] * it)
        dlnkinf = (m["b"][i] + m["EaR"][i] * it) * it
        kind = m["kind"][i]
        Mc = dflt[i] * ctot + sum(c * C[k] for k, c in listed[i]) if kind else 0.0
        kf, dlnkf, dkfdM = kinf, dlnkinf, 0.0
        if kind == mech.THREE_BODY:
            kf, dkfdM = kinf * Mc, kinf
        elif kind == mech.FALLOFF:
            k0 = m["A0"][i] * np.exp(m["b0"][i] * logt - m["Ea0R"][i] * it)
            pr = Mc * k0 / (kinf + SMALL)
            F = g_lpr = g_lc = dlc = 0.0
            F = 1.0
            if m["troe"][i] > 0:
                A, T3, T1, T2 = m["troeA"][i], m["troeT3"][i], m["troeT1"][i], m["troeT2"][i]
                r3 = np.inf if abs(T3) < SMALL else 1 / T3
                r1 = np.inf if abs(T1) < SMALL else 1 / T1
                e3, e1 = np.exp(-T * r3), np.exp(-T * r1)
                fc = (1 - A) * e3 + A * e1
                dfc = -(1 - A) * (r3 * e3 if e3 > 0 else 0.0) - A * (r1 * e1 if e1 > 0 else 0.0)
                if T2 != 0:
                    e2 = np.exp(-T2 * it)
                    fc, dfc = fc + e2, dfc + T2 * it * it * e2
                fcc = max(fc, SMALL)
                lc, lpr = np.log10(fcc), np.log10(max(pr, SMALL))
                c, n = -0.4 - 0.67 * lc, 0.75 - 1.27 * lc
                u = lpr + c
                den = n - 0.14 * u
                f1 = u / den
                qq = 1 / (1 + f1 * f1)
                F = 10.0 ** (lc * qq)
                dlgf = -2 * lc * f1 * qq * qq
                g_lpr = dlgf * n / den ** 2 if pr > SMALL else 0.0
                g_lc = qq + dlgf * (n / den ** 2 * -0.67 + -u / den ** 2 * -1.27)
                dlc = dfc / (fcc * np.log(10)) if fc > SMALL else 0.0
            opr = 1 / (1 + pr)
            kf = kinf * pr * opr * F
            dlnpr = (m["b0"][i] + m["Ea0R"][i] * it) * it - dlnkinf
            dlnkf = dlnkinf + opr * dlnpr + g_lpr * dlnpr + np.log(10) * g_lc * dlc
            dkfdM = k0 * F * opr * (opr + g_lpr)
        pf, pb = np.prod(C[reac]), np.prod(C[prod])
        dn = len(prod) - len(reac)
        dg = grt[prod].sum() - grt[reac].sum()
        dh = hrt[prod].sum() - hrt[reac].sum()
        rkc = np.exp(dg - dn * logc0) if m["rev"][i] else 0.0
        kr = kf * rkc
        dlnrkc = (dn - dh) * it if m["rev"][i] else 0.0
        dslot[i] = ([kf * np.prod(np.delete(C[reac], j)) for j in range(len(reac))],
                    [-kr * np.prod(np.delete(C[prod], j)) for j in range(len(prod))])
        dM[i] = dkfdM * (pf - rkc * pb)
        dT[i] = kf * dlnkf * pf - kr * (dlnkf + dlnrkc) * pb
    # d wdot / d C by columns: a slot's derivative passes to every slot's species
    D = np.zeros((ns, ns))
    for i in range(nr):
        reac, prod = sl[i]
        signed = [(k, -1.0) for k in reac] + [(k, 1.0) for k in prod]
        for sp_, der in list(zip(reac, dslot[i][0])) + list(zip(prod, dslot[i][1])):
            for k, sg in signed:
                D[k, sp_] += sg * der
        for sp_, corr in listed[i]:
            for k, sg in signed:
                D[k, sp_] += sg * corr * dM[i]
    nu = (m["nu_p"] - m["nu_r"]).astype(float)
    a = nu.T @ dT
    u = nu.T @ np.where(np.asarray(m["kind"]) != mech.ELEMENTARY, dflt * dM, 0.0)
    ek = (hrt - 1) * R_GAS * T / W
    cv = R_GAS * np.sum(Y / W * (cpr - 1))
    dTdy = -(ek[:N] - ek[N]) / cv
    J = (W[:N, None] / W[None, :N]) * (D[:N, :N] + u[:N, None]) \
        - (W[:N] / W[N] * (D[:N, N] + u[:N]))[:, None] + (W[:N] / rho * a[:N])[:, None] * dTdy[None, :]
    return J


def central_differences(m, y, rho, e):
    N = len(y)
    J = np.zeros((N, N))
    for s in range(N):
        hs = 1e-7 * max(abs(y[s]), 1e-10)
        yp, ym = y.copy(), y.copy()
        yp[s] += hs
        ym[s] -= hs
        fp = tn.chem_rhs(m, yp[None], np.array([rho]), np.array([e]))[0][0]
        fm = tn.chem_rhs(m, ym[None], np.array([rho]), np.array([e]))[0][0]
        J[:, s] = (fp - fm) / (2 * hs)
    return J


@pytest.mark.parametrize("path,seed", [(DME, 4), (H2, 2)], ids=["dme", "h2"])
def test_column_jacobian_matches_central_differences(path, seed):
    m = mech.load(str(path))
    ns = m["ns"]
    rng = np.random.default_rng(seed)
    W = m["W"]
    names = m["species"]
    fuel = "CH3OCH3" if "CH3OCH3" in names else "H2"
    for T0 in (800.0, 1300.0, 2100.0):
        # fresh mixture plus a seeded sprinkle of every species (radicals present)
        X = np.full(ns, 1e-4) * rng.uniform(0.5, 1.5, ns)
        X[names.index(fuel)] += 0.1
        X[names.index("O2")] += 0.2
        X[names.index("N2")] += 0.7
        Y = X * W / np.sum(X * W)
        rho = 2e5 / (R_GAS * T0 * np.sum(Y / W))
        e = tn.e_cv(m, np.array([T0]), Y[None])[0][0]
        y = Y[:-1].copy()
        Ja = jacobian(m, y, rho, e)
        Jf = central_differences(m, y, rho, e)
        err = np.abs(Ja - Jf).max() / np.abs(Jf).max()
        assert err < 1e-6, (T0, err)
