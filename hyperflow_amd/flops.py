"""Algorithmic operation counts of the chemistry integrator, for roofline
accounting. The counts follow the arithmetic of the kernels (hf_device.hpp,
hf_bdf.hpp) term by term and depend only on the mechanism's topology.

Fused multiply-adds count as two operations, other FP64 adds and multiplies
as one. A transcendental or a division counts as the FP64 arithmetic of the
code hipcc -O3 generates for it on gfx950 (an FMA as two, an add or multiply
as one; compares, selects, conversions and exponent manipulation are not
counted): exp 27, log 78 (a compensated evaluation), log10 88, division 13,
sqrt 16. `measure_math_flops` recompiles the functions and counts them.
"""
from __future__ import annotations

EXP, LOG, LOG10, DIV, SQRT = 27, 78, 88, 13, 16
MATH_FLOPS = {"exp": EXP, "log": LOG, "log10": LOG10, "div": DIV, "sqrt": SQRT}


def measure_math_flops(hipcc="/opt/rocm/bin/hipcc", arch="gfx950"):
    """FP64 arithmetic of exp, log, log10, 1/x and sqrt in gfx950 code:
    compiles one tiny kernel per function and counts v_fma/v_fmac (x2) and
    v_add/v_mul (x1) on f64 in its assembly."""
    import re
    import subprocess
    import tempfile
    from pathlib import Path
    exprs = {"exp": "exp(a[i])", "log": "log(a[i])", "log10": "log10(a[i])", "div": "1.0 / a[i]",
             "sqrt": "sqrt(a[i])"}
    out = {}
    with tempfile.TemporaryDirectory() as d:
        for name, e in exprs.items():
            src = Path(d) / f"{name}.hip"
            asm = Path(d) / f"{name}.s"
            src.write_text("#include <hip/hip_runtime.h>\n"
                           f"__global__ void k(const double* a, double* o) {{ const int i = threadIdx.x; o[i] = {e}; }}\n")
            subprocess.run([hipcc, "-O3", f"--offload-arch={arch}", "-Xclang", "-target-feature", "-Xclang",
                            "-packed-fp32-ops", "--cuda-device-only", "-S", "-o", str(asm), str(src)],
                           check=True, capture_output=True)
            text = asm.read_text()
            body = text[text.index("_Z1kPKdPd:"):]
            body = body[:body.index("s_endpgm")]
            fma = len(re.findall(r"^\s+v_(fma|fmac|div_fmas)_f64", body, re.M))
            addmul = len(re.findall(r"^\s+v_(add|mul)_f64", body, re.M))
            out[name] = 2 * fma + addmul
    return out


def rhs_flops(ns, reactions, newton_iters=2.0):
    """One right-hand-side evaluation: T from e, species g/RT, rates, dY/dt."""
    n = ns - 1
    f = 0.0
    f += 2 * ns                       # closure Y_last = 1 - sum(y), C_k = rho Y_k / W_k
    f += 24 * ns + ns                 # energy polynomial of the mixture (two ranges, 6 coefficients)
    f += newton_iters * (21 + DIV)    # e(T), cv(T) Horner, step (e - e*)/cv, update, test
    f += 26 * ns + LOG + DIV          # h/RT and s/R per species, log T, 1/T
    f += 1 + ns                       # ln(p0/RT) = ln(p0/R) - ln T, total concentration
    for r in reactions:
        nr_, np_ = r["nreac"], r["nprod"]
        f += 4 + EXP                  # k = exp(lnA + b lnT - Ea/RT)
        if r["kind"] == 1:
            f += 2 * r["neff"] + 2    # [M] and k [M]
        if r["kind"] == 2:
            f += 2 * r["neff"] + 2 + 4 + EXP + 2 + DIV      # [M], k0, Pr = [M] k0 / k
            if r["troe"]:
                f += 2 * EXP + 5 + ((EXP + 2) if r["troe"] == 4 else 0)   # Fcent
                f += 2 * LOG10 + 12 + 2 * DIV + EXP + 1                   # log10 Fcent, log10 Pr, F
            f += 1 + DIV + 3          # Pr/(1+Pr) F blend
        if r["rev"]:
            f += (nr_ + np_) + 2 + EXP + 1   # sum nu g/RT, 1/Kc = exp(.), k_r
        f += nr_ + np_ + 1            # concentration products, q
        f += nr_ + np_                # wdot accumulation
    f += 2 * n + DIV                  # dY/dt = wdot W / rho
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
    f += ns * 14 + 2 + 2 * DIV                    # cv, species energies, 1/cv, 1/rho
    f += n * (4 + 5 * n)                          # chain rule into J
    return f


def step_flops(n, q=3):
    """BDF bookkeeping per step: prediction, coefficients, error test (one
    WRMS norm), update, and the step-size ratio eta = (6 dsm)^(-1/L) (a root,
    counted as a log and an exp; 1.5 per step on average with the order
    checks)."""
    pascal = n * q * (q + 1) // 2
    return 2 * pascal + 60 + 8 * n + 2 * n * (q + 1) + SQRT + 1.5 * (LOG + EXP + DIV)


def lu_flops(n):
    return (2.0 / 3.0) * n ** 3 + n * n + n * DIV    # M = I - gamma J, factor (reciprocal pivots)


def newton_flops(n):
    return 2.0 * n * n + 8 * n + SQRT                 # residual, triangular solves, update, WRMS


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


def dq_jac_flops(n):
    """Difference-quotient Jacobian beyond its n rate evaluations (which the
    kernels count as right-hand sides): the increments (a WRMS norm, n
    maxima and reciprocals) and n columns of n scaled differences."""
    return 2 * n + SQRT + n * (4 + DIV) + 2 * n * n


def cell_flops(stats, ns, reactions, analytic_jac=True):
    """Total algorithmic operations of a chemistry launch from its counters.
    analytic_jac False: the Jacobian is CVODE's difference quotient (the
    large-mechanism kernels), whose rate evaluations are in stats["rhs"]."""
    n = ns - 1
    jac = jac_flops(ns, reactions) if analytic_jac else dq_jac_flops(n)
    return (stats["rhs"] * rhs_flops(ns, reactions) + stats["jac"] * jac
            + stats["lu"] * lu_flops(n) + stats["newton"] * newton_flops(n) + stats["steps"] * step_flops(n))
