"""Write tests/golden/chem_radau.npz: reference trajectories for the chemistry
oracle, from an independent method (scipy Radau IIA, order 5) on the numpy
restatement of the right-hand side (oracle/thermo_np.py), at rtol 1e-12.

Cells: the uniform stoichiometric H2/air scenario of the reference's
zero-dimensional check (HyperFlow_V3/tests/validate.py:140-161: 1500 K,
5 atm, 2e-7 s), plus the first seeded cells of the batched-chemistry
workload (seed 2). Run from the repository root:

    python tests/golden/make_chem_golden.py
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
from scipy.integrate import solve_ivp

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle import mech as mech_py  # noqa: E402
from oracle import thermo_np as tn  # noqa: E402

MECH = ROOT / "hyperflow_amd" / "mechanisms" / "H2_Li.yaml"


def conserved_from_TPX(m, T, p, X):
    Y = X * m["W"] / np.sum(X * m["W"])
    wbar = 1.0 / np.sum(Y / m["W"])
    rho = p * wbar / (mech_py.R_GAS * T)
    e, _ = tn.e_cv(m, np.array([T]), Y[None])
    U = np.zeros(m["ns"] + 3)
    U[: m["ns"]] = rho * Y
    U[m["ns"] + 2] = rho * e[0]
    return U


def integrate(m, U, dt):
    ns = m["ns"]
    rho = U[:ns].sum()
    e = U[ns + 2] / rho
    y0 = U[: ns - 1] / rho
    state = {"T": None}

    def f(t, y):
        dy, T = tn.chem_rhs(m, y[None], np.array([rho]), np.array([e]), state["T"])
        state["T"] = T
        return dy[0]

    sol = solve_ivp(f, (0.0, dt), y0, method="Radau", rtol=1e-12, atol=1e-20)
    assert sol.success, sol.message
    y = sol.y[:, -1]
    Y = np.concatenate([y, [1.0 - y.sum()]])
    T = tn.temperature(m, np.array([e]), Y[None])[0]
    return Y, T


def main(ncells=24):
    m = mech_py.load(MECH)
    ns = m["ns"]
    X = np.zeros(ns)
    X[m["species"].index("H2")], X[m["species"].index("O2")], X[m["species"].index("N2")] = 2, 1, 3.76
    X /= X.sum()
    cells = [conserved_from_TPX(m, 1500.0, 5 * 101325.0, X)]
    dts = [2e-7]
    syn = np.load(ROOT / "tests" / "golden" / "synth_h2_seed2.npz")
    for i in range(ncells):
        cells.append(syn["U"][:, i])
        dts.append(syn["dt"][i])
    U0 = np.array(cells).T.copy()
    dts = np.array(dts)
    Yend = np.zeros((ns, U0.shape[1]))
    Tend = np.zeros(U0.shape[1])
    for i in range(U0.shape[1]):
        Yend[:, i], Tend[i] = integrate(m, U0[:, i], dts[i])
        print(i, dts[i], Tend[i], flush=True)
    np.savez(ROOT / "tests" / "golden" / "chem_radau.npz", U0=U0, dt=dts, Yend=Yend, Tend=Tend)


if __name__ == "__main__":
    main()
