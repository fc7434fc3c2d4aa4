"""Oracle-side mechanism loader (TEST INFRASTRUCTURE ONLY).

Reads a Cantera-YAML mechanism with PyYAML into flat numpy tables. This is a
second, independent reading of the same files the product's C++ loader parses
(hyperflow_amd/csrc/mechanism.cpp); tests compare the two table by table.

Semantics restated from Cantera 3.2.0 (the reference loads mechanisms through
``Cantera::newSolution(mech, "", "none")``, HyperFlow_V3/src/chemistry_batch.cpp:17-18):
  * molecular weights from IUPAC conventional atomic weights;
  * NASA7 with one or two temperature ranges (the low range is used for T <= Tmid);
  * Arrhenius A converted to SI kmol units by the reaction order
    (three-body and falloff-k0 count M as one more reactant);
  * activation energy with an explicit unit ("cal/mol") or the file default.
"""
from __future__ import annotations

import re

import numpy as np
import yaml

R_GAS = 8314.462618          # J/kmol/K
ONE_ATM = 101325.0
ATOMIC_WEIGHT = {"H": 1.008, "C": 12.011, "N": 14.007, "O": 15.999}

ELEMENTARY, THREE_BODY, FALLOFF = 0, 1, 2


def _energy_K(value, default_unit):
    if isinstance(value, (int, float)):
        number, unit = float(value), default_unit
    else:
        parts = str(value).split()
        number, unit = float(parts[0]), (" ".join(parts[1:]) or default_unit)
    per_kmol = {"cal/mol": 4184.0, "kcal/mol": 4.184e6, "J/mol": 1000.0,
                "kJ/mol": 1.0e6, "J/kmol": 1.0}
    if unit == "K":
        return number
    return number * per_kmol[unit] / R_GAS


def _side(text, index):
    """Parse one side of an equation -> (species counts, has M, has (+M))."""
    falloff = False
    if re.search(r"\(\+\s*M\s*\)", text):
        falloff = True
        text = re.sub(r"\(\+\s*M\s*\)", " ", text)
    counts = {}
    third = False
    for term in [t.strip() for t in text.split(" + ")]:
        if not term:
            continue
        if term == "M":
            third = True
            continue
        m = re.match(r"^(\d+)\s+(\S+)$", term)
        coef, name = (int(m.group(1)), m.group(2)) if m else (1, term)
        k = index[name]
        counts[k] = counts.get(k, 0) + coef
    return counts, third, falloff


def load(path):
    with open(path, "r", newline=None) as f:
        doc = yaml.safe_load(f)
    units = doc.get("units", {}) or {}
    length = {"m": 1.0, "cm": 0.01, "mm": 1e-3}[units.get("length", "m")]
    quantity = {"kmol": 1.0, "mol": 1e-3}[units.get("quantity", "kmol")]
    ea_unit = units.get("activation-energy", "J/kmol")
    conc = quantity / length ** 3

    phase = doc["phases"][0]
    elements = list(phase["elements"])
    species = [str(s) for s in phase["species"]]
    index = {s: k for k, s in enumerate(species)}
    ns, ne = len(species), len(elements)
    W = np.zeros(ns)
    comp = np.zeros((ns, ne))
    Tmin, Tmid, Tmax = np.zeros(ns), np.zeros(ns), np.zeros(ns)
    low, high = np.zeros((ns, 7)), np.zeros((ns, 7))
    for sp in doc["species"]:
        k = index.get(str(sp["name"]))
        if k is None:
            continue
        for el, n in sp["composition"].items():
            comp[k, elements.index(el)] = n
            W[k] += n * ATOMIC_WEIGHT[el]
        th = sp["thermo"]
        assert th["model"] == "NASA7"
        ranges, data = th["temperature-ranges"], th["data"]
        if len(ranges) == 3:
            Tmin[k], Tmid[k], Tmax[k] = ranges
            low[k], high[k] = data[0], data[1]
        else:
            Tmin[k], Tmax[k] = ranges
            Tmid[k] = Tmax[k]
            low[k] = high[k] = data[0]

    rx = doc.get("reactions", [])
    nr = len(rx)
    out = dict(kind=np.zeros(nr, np.int32), rev=np.ones(nr, np.int32),
               A=np.zeros(nr), b=np.zeros(nr), EaR=np.zeros(nr),
               A0=np.zeros(nr), b0=np.zeros(nr), Ea0R=np.zeros(nr),
               troe=np.zeros(nr, np.int32), troeA=np.zeros(nr), troeT3=np.zeros(nr),
               troeT1=np.zeros(nr), troeT2=np.zeros(nr),
               nu_r=np.zeros((nr, ns), np.int32), nu_p=np.zeros((nr, ns), np.int32),
               eff=np.ones((nr, ns)))
    equations = []
    for i, r in enumerate(rx):
        eq = r["equation"]
        equations.append(eq)
        if "<=>" in eq:
            lhs, rhs = eq.split("<=>")
        elif "=>" in eq:
            lhs, rhs = eq.split("=>")
            out["rev"][i] = 0
        else:
            lhs, rhs = eq.split(" = ")
        cr, tbl, fol = _side(lhs, index)
        cp, tbr, forr = _side(rhs, index)
        for k, v in cr.items():
            out["nu_r"][i, k] = v
        for k, v in cp.items():
            out["nu_p"][i, k] = v
        order = sum(cr.values())
        typ = r.get("type", "elementary")

        def arr(rc, order_):
            return (float(rc["A"]) * conc ** (1 - order_), float(rc["b"]),
                    _energy_K(rc["Ea"], ea_unit))

        if typ == "elementary":
            out["kind"][i] = ELEMENTARY
            out["A"][i], out["b"][i], out["EaR"][i] = arr(r["rate-constant"], order)
        elif typ == "three-body":
            assert tbl and tbr
            out["kind"][i] = THREE_BODY
            out["A"][i], out["b"][i], out["EaR"][i] = arr(r["rate-constant"], order + 1)
        elif typ == "falloff":
            assert fol and forr
            out["kind"][i] = FALLOFF
            out["A"][i], out["b"][i], out["EaR"][i] = arr(r["high-P-rate-constant"], order)
            out["A0"][i], out["b0"][i], out["Ea0R"][i] = arr(r["low-P-rate-constant"], order + 1)
            if "Troe" in r:
                t = r["Troe"]
                out["troe"][i] = 4 if "T2" in t else 3
                out["troeA"][i], out["troeT3"][i], out["troeT1"][i] = t["A"], t["T3"], t["T1"]
                out["troeT2"][i] = t.get("T2", 0.0)
        else:
            raise ValueError(f"unsupported reaction type {typ}")
        for name, e in (r.get("efficiencies") or {}).items():
            if name in index:
                out["eff"][i, index[name]] = e
    return dict(path=str(path), species=species, elements=elements, ns=ns, nr=nr, ne=ne,
                W=W, comp=comp, Tmin=Tmin, Tmid=Tmid, Tmax=Tmax, low=low, high=high,
                equations=equations, **out)
