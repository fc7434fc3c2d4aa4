"""Algorithmic operation counts of the chemistry integrator, for roofline
accounting. A transcendental (exp, log, log10, pow, sqrt) or a division is
counted as one operation; fused multiply-adds as two. The counts follow the
arithmetic of the kernels (hf_device.hpp, hf_bdf.hpp) term by term and depend
only on the mechanism's topology.
"""
from __future__ import annotations


def rhs_flops(ns, reactions, newton_iters=2.0):
    """One right-hand-side evaluation: T from e, species g/RT, rates, dY/dt."""
    n = ns - 1
    f = 0.0
    f += 2 * ns                       # closure Y_last = 1 - sum(y), C_k = rho Y_k / W_k
    f += 24 * ns + ns                 # energy polynomial of the mixture (two ranges, 6 coefficients)
    f += newton_iters * 22            # e(T), cv(T) Horner (16), step division, update, test
    f += 26 * ns + 2                  # h/RT and s/R per species, log T
    f += 3 + ns                       # 1/T, log(p0/RT), total concentration
    for r in reactions:
        nr_, np_ = r["nreac"], r["nprod"]
        f += 5                        # k = exp(lnA + b lnT - Ea/RT)
        if r["kind"] == 1:
            f += 2 * r["neff"] + 2    # [M] and k [M]
        if r["kind"] == 2:
            f += 2 * r["neff"] + 2 + 5 + 3      # [M], k0, Pr
            if r["troe"]:
                f += 3 + (2 if r["troe"] == 4 else 0) + 2 + 16   # Fcent, log10 Fcent, log10 Pr, F
            f += 5                    # Pr/(1+Pr) F blend
        if r["rev"]:
            f += 2 * (nr_ + np_) + 4  # sum nu g/RT, dn, exp, k_r
        f += nr_ + np_ + 1            # concentration products, q
        f += nr_ + np_                # wdot accumulation
    f += 3 * n                        # dY/dt = wdot W / rho
    return f


def jac_flops(ns, reactions):
    """Analytic Jacobian: rate constants with T derivatives, d q / d C and
    d q / d T per reaction, chain rule to the mass-fraction unknowns."""
    n = ns - 1
    f = rhs_flops(ns, reactions, newton_iters=0.0)
    for r in reactions:
        nr_, np_ = r["nreac"], r["nprod"]
        f += 4                                    # d ln k / dT
        if r["kind"] == 2:
            f += 12 + (14 if r["troe"] else 0)    # falloff derivatives
        f += 6                                    # d q / d T
        f += (nr_ + np_) * (max(nr_, np_) + 2 * (nr_ + np_))   # d q / d C_s into the touched rows
        if r["kind"] != 0:
            f += ns * (2 + 2 * (nr_ + np_))       # third-body column for every species
    f += ns * 14 + 4                              # cv, species energies, 1/cv, 1/rho
    f += n * (4 + 5 * n)                          # chain rule into J
    return f


def step_flops(n, q=3):
    """BDF bookkeeping per step: prediction, coefficients, error test, update."""
    pascal = n * q * (q + 1) // 2
    return 2 * pascal + 60 + 8 * n + 2 * n * (q + 1)


def lu_flops(n):
    return (2.0 / 3.0) * n ** 3 + n * n + n          # M = I - gamma J, factor


def newton_flops(n):
    return 2.0 * n * n + 8 * n                        # residual, triangular solves, update, WRMS


def topology_summary(topo_inc_path):
    """Reactions summary parsed from a generated topology header."""
    import re
    text = open(topo_inc_path).read()
    def arr(name):
        m = re.search(r"static constexpr int %s\[NR\] = \{([^}]*)\}" % name, text)
        return [int(x) for x in m.group(1).split(",")]
    ns = int(re.search(r"NS = (\d+)", text).group(1))
    keys = ("kind", "rev", "troe", "nreac", "nprod", "neff")
    cols = {k: arr(k) for k in keys}
    reactions = [dict(zip(keys, vals)) for vals in zip(*(cols[k] for k in keys))]
    return ns, reactions


def cell_flops(stats, ns, reactions):
    """Total algorithmic operations of a chemistry launch from its counters."""
    n = ns - 1
    return (stats["rhs"] * rhs_flops(ns, reactions) + stats["jac"] * jac_flops(ns, reactions)
            + stats["lu"] * lu_flops(n) + stats["newton"] * newton_flops(n) + stats["steps"] * step_flops(n))
