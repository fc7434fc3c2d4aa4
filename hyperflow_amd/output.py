"""Output files of a run in the reference's formats, the ones its own test
scripts read (HyperFlow_V3/tests/validate.py:30-61, check_cases.py):

  state_<k>.csv   one row per active cell (BlockAMR::snapshot, BlockAMR.cpp:374-388):
                  t,level,i,j,x,y,dx,dy,rho,u,v,p,p_atm,T,rhoE,Y_<species>...
  boxes_<k>.csv   one row per box (BlockAMR.cpp:379-383):
                  level,x_left,x_right,y_bottom,y_top,dx,dy
  history.csv     one row per step (BlockAMR::log, BlockAMR.cpp:360-366; header :80-81)
  summary.json    the run's totals (BlockAMR::run, BlockAMR.cpp:426-438)

Numbers are written with 17 significant digits (std::setprecision(17)).
Conserved integrals are sums over the active cells of value x dx x dy
(BlockAMR::totals, BlockAMR.cpp:347-353). The writers take plain arrays; the
drivers' ``snapshot_levels`` gathers them from the device, with p and T from
the library's properties kernel.
"""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np

STATE_HEAD = "t,level,i,j,x,y,dx,dy,rho,u,v,p,p_atm,T,rhoE"
BOXES_HEAD = "level,x_left,x_right,y_bottom,y_top,dx,dy"
HISTORY_HEAD = ("step,t,dt,finest_level,leaf_cells,stored_cells,dx_min,dy_min,mass,momentum_x,momentum_y,"
                "total_energy,regrids,retries")
ATM = 101325.0


def _g(v):
    """std::ostream << double at setprecision(17) (the %g form with 17 digits)."""
    return format(float(v), ".17g")


class Level:
    """One level's output data: Q (nv, ny, nx) conserved values of a box whose
    cell (0, 0) has level indices lo, its active mask (ny, nx), p and T per
    cell, cell sizes and the level's physical origin."""

    def __init__(self, level, Q, p, T, lo, dx, dy, prob_lo, active=None):
        self.level, self.Q, self.p, self.T = level, Q, p, T
        self.lo, self.dx, self.dy, self.prob_lo = tuple(lo), dx, dy, tuple(prob_lo)
        self.active = np.ones(Q.shape[1:], dtype=bool) if active is None else active


def write_snapshot(out, frame, time, levels, species):
    """state_<frame>.csv and boxes_<frame>.csv for the given Levels."""
    out = Path(out)
    out.mkdir(parents=True, exist_ok=True)
    ns = len(species)
    with open(out / f"state_{frame}.csv", "w") as f, open(out / f"boxes_{frame}.csv", "w") as b:
        f.write(STATE_HEAD + "".join(f",Y_{s}" for s in species) + "\n")
        b.write(BOXES_HEAD + "\n")
        for L in levels:
            ny, nx = L.Q.shape[1:]
            i0, j0 = L.lo
            x0, y0 = L.prob_lo
            b.write(",".join([str(L.level), _g(x0 + i0 * L.dx), _g(x0 + (i0 + nx) * L.dx), _g(y0 + j0 * L.dy),
                              _g(y0 + (j0 + ny) * L.dy), _g(L.dx), _g(L.dy)]) + "\n")
            for j in range(ny):                       # BoxIterator order: i fastest
                for i in range(nx):
                    if not L.active[j, i]:
                        continue
                    q = L.Q[:, j, i]
                    rho = q[:ns].sum()
                    I, J = i0 + i, j0 + j
                    row = [_g(time), str(L.level), str(I), str(J), _g(x0 + (I + 0.5) * L.dx),
                           _g(y0 + (J + 0.5) * L.dy), _g(L.dx), _g(L.dy), _g(rho), _g(q[ns] / rho), _g(q[ns + 1] / rho),
                           _g(L.p[j, i]), _g(L.p[j, i] / ATM), _g(L.T[j, i]), _g(q[ns + 2])]
                    row += [_g(q[n] / rho) for n in range(ns)]
                    f.write(",".join(row) + "\n")


def totals(levels, absolute=False):
    """Active-cell integrals of every conserved component (BlockAMR::totals)."""
    nv = levels[0].Q.shape[0]
    s = np.zeros(nv, dtype=np.longdouble)
    for L in levels:
        Q = np.abs(L.Q) if absolute else L.Q
        s += (Q[:, L.active].astype(np.longdouble) * L.dx * L.dy).sum(axis=1)
    return s.astype(float)


class History:
    """history.csv, one line per step (BlockAMR::log)."""

    def __init__(self, out):
        Path(out).mkdir(parents=True, exist_ok=True)
        self.f = open(Path(out) / "history.csv", "w")
        self.f.write(HISTORY_HEAD + "\n")

    def log(self, step, t, dt, levels, ns, regrids=0, retries=0):
        sums = totals(levels)
        leaf = sum(int(L.active.sum()) for L in levels)
        stored = sum(int(L.active.size) for L in levels)
        fin = levels[-1]
        self.f.write(",".join([str(step), _g(t), _g(dt), str(fin.level), str(leaf), str(stored), _g(fin.dx),
                               _g(fin.dy), _g(sums[:ns].sum()), _g(sums[ns]), _g(sums[ns + 1]), _g(sums[ns + 2]),
                               str(regrids), str(retries)]) + "\n")
        self.f.flush()

    def close(self):
        self.f.close()


def write_summary(out, **fields):
    """summary.json with the reference's keys (BlockAMR.cpp:426-436); the
    caller passes the values it has, lists for the integrals."""
    keys = ("completed", "final_time", "cfl", "steps", "retries", "finest_level", "leaf_cells", "max_leaf_cells",
            "full_finest_cells", "dx_min", "dy_min", "regrids", "max_remap_relative_defect",
            "reflux_stage_interfaces", "positivity_first_order_stages", "roundoff_redistributions",
            "max_repaired_negative_Y", "redistributed_species_integrals", "chemistry_cell_calls",
            "chemistry_threads", "last_chemistry_team", "chemistry_seconds", "flow_seconds", "regrid_seconds",
            "wall_seconds", "min_dt", "max_dt", "initial_conserved_integrals", "final_conserved_integrals",
            "time_subcycling", "refinement_ratio")
    unknown = set(fields) - set(keys)
    if unknown:
        raise ValueError(f"not a summary key: {sorted(unknown)}")
    result = {k: fields[k] for k in keys if k in fields}
    Path(out).mkdir(parents=True, exist_ok=True)
    (Path(out) / "summary.json").write_text(json.dumps(result, indent=2) + "\n")
    return result


def snapshot_levels(driver, prob_lo=(0.0, 0.0)):
    """Levels of a BoxDriver or TwoLevelDriver, p and T from the library's
    properties kernel (device), for the writers above."""
    mech = driver.mech
    nv = mech.nv

    def props(Qd):
        flat = Qd.reshape(nv, -1).contiguous()
        p, _, T = mech.properties(flat)
        return p.cpu().numpy().reshape(Qd.shape[1:]), T.cpu().numpy().reshape(Qd.shape[1:])
    if hasattr(driver, "coarse"):
        c = driver.coarse
        Qc = c.interior().contiguous()
        pc, Tc = props(Qc)
        active = driver.active.reshape(c.ny, c.nx).cpu().numpy()
        Qf = driver.fine_interior().contiguous()
        pf, Tf = props(Qf)
        FI0, FJ0, _, _ = driver.fine_box
        return [Level(0, Qc.cpu().numpy(), pc, Tc, (0, 0), c.dx, c.dy, prob_lo, active),
                Level(1, Qf.cpu().numpy(), pf, Tf, (FI0, FJ0), driver.dxf, driver.dyf, prob_lo)]
    Q = driver.interior().contiguous()
    p, T = props(Q)
    return [Level(0, Q.cpu().numpy(), p, T, (0, 0), driver.dx, driver.dy if not driver.one_d else driver.dx,
                  prob_lo)]
