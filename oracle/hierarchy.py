"""CPU restatement of the reference's multi-level step loop — TEST
INFRASTRUCTURE ONLY (the checker of hyperflow_amd.hierarchy.AmrDriver).

Follows HyperFlow_V3/src/BlockAMR.cpp over a hierarchy of any number of
levels at ratio 2 with several boxes per fine level:
  fill            :137-152 (FillPatchTwoLevels: the ghost cells of a level's
                  boxes that no box of the level holds come from the coarser
                  level by ConservativeInterp, :18-51 of its file; then the
                  physical boundary rules, FlowBoundary.hpp:7-37)
  euler           :202-265 (every level's stage, reflux finest to coarsest
                  with ClearInternalBorders, average, repairRoundoff over the
                  leaf cells of all levels, validate; first-order redo)
  flow, react     :266-285 (one chemistry batch over every leaf cell)
  repairRoundoff  :299-334, stableDt :335-346, totals :347-353
  regridChecked   :367-373 (the new boxes are given: they come from the
                  product's box bookkeeping, whose rules are pinned on their
                  own in tests/test_amr_grid.py)
Each level is held as one array over its whole domain with two ghost layers
and a mask (0 no box, 1 leaf, 2 covered), the cells of level l at
[:, j + 2, i + 2]. Every arithmetic step is the oracle's own (flux stage,
properties, chemistry, interpolation, average-down, tags) or numpy below.

Parity status: the composition is pinned by composite conservation on a
closed domain with several boxes on two fine levels, by a uniform state
staying uniform, by a fine level cut into several boxes giving the same
result as one box, and by reducing to OracleTwoLevelDriver for one fine box
(tests/test_oracle_hierarchy.py). The donor of the roundoff repair is the
first leaf in (level, j, i) order holding the largest mass; the reference's
is first in (level, box, j, i) order, which differs only when two cells hold
exactly the same largest mass in different boxes.
"""
from __future__ import annotations

import numpy as np

from .amr import average_down, interp, tag_cells
from .driver import EPS, PERIODIC, REFLECTING, WEDGE, DomainError, fill_ghosts


def _strips(box):
    i0, j0, i1, j1 = box
    return [(i0 - 2, j0 - 2, i1 + 2, j0 - 1), (i0 - 2, j1 + 1, i1 + 2, j1 + 2), (i0 - 2, j0, i0 - 1, j1),
            (i1 + 1, j0, i1 + 2, j1)]


def _reg(mc, fc, mf, f0, f1):
    reg = mc * fc
    reg = reg + mf * f0
    return reg + mf * f1


class OracleAmrDriver:
    def __init__(self, oracle, nx0, ny0, dx0, dy0, bc, max_level=2, cfl=0.5, rtol=1e-12, atol=1e-20,
                 max_attempts=18, inlet=None, wedge_i=0):
        self.o = oracle
        self.ns, self.nv = oracle.ns, oracle.ns + 3
        self.nx0, self.ny0, self.dx0, self.dy0 = nx0, ny0, dx0, dy0
        self.bc = list(bc)
        if PERIODIC in self.bc:
            raise ValueError("fine levels on periodic domains are not restated")
        self.max_level = max_level
        self.cfl, self.rtol, self.atol, self.max_attempts = cfl, rtol, atol, max_attempts
        self.inlet = None if inlet is None else np.asarray(inlet, dtype=float)
        self.wedge_i0 = int(wedge_i)
        self.L = []
        for lev in range(max_level + 1):
            nx, ny = nx0 << lev, ny0 << lev
            self.L.append(dict(lev=lev, nx=nx, ny=ny, dx=dx0 / (1 << lev), dy=dy0 / (1 << lev),
                               Q=np.zeros((self.nv, ny + 4, nx + 4)), mask=np.zeros((ny, nx), dtype=np.uint8),
                               boxes=[]))
        self.finest = 0
        self.time, self.steps, self.retries, self.first_order_stages, self.repairs = 0.0, 0, 0, 0, 0
        self.set_grids([[]])

    # ------------------------------------------------------------ grids and state
    def set_grids(self, boxes):
        nlev = 1
        while nlev < len(boxes) and nlev <= self.max_level and boxes[nlev]:
            nlev += 1
        self.finest = nlev - 1
        for lev in range(self.max_level + 1):
            L = self.L[lev]
            L["boxes"] = [(0, 0, L["nx"] - 1, L["ny"] - 1)] if lev == 0 else \
                ([tuple(b) for b in boxes[lev]] if lev <= self.finest else [])
        for lev in range(self.finest + 1):
            L = self.L[lev]
            valid = np.zeros((L["ny"], L["nx"]), dtype=bool)
            for i0, j0, i1, j1 in L["boxes"]:
                valid[j0:j1 + 1, i0:i1 + 1] = True
            m = valid.astype(np.uint8)
            if lev < self.finest:
                for i0, j0, i1, j1 in self.L[lev + 1]["boxes"]:
                    m[j0 // 2:j1 // 2 + 1, i0 // 2:i1 // 2 + 1] = 2
            L["mask"] = m

    def interior(self, lev, G=None):
        L = self.L[lev]
        G = L["Q"] if G is None else G
        return G[:, 2:L["ny"] + 2, 2:L["nx"] + 2]

    def set_state(self, Q0):
        self.interior(0)[:] = Q0
        for lev in range(1, self.finest + 1):
            self.fill([L["Q"] for L in self.L], lev - 1)
            for b in self.L[lev]["boxes"]:
                self._interp_into(self.L[lev - 1]["Q"], self.L[lev]["Q"], lev, b, None)
        self.average([L["Q"] for L in self.L])

    # ------------------------------------------------------------ pieces
    def _interp_into(self, C, G, lev, region, mask):
        F, _, failed = interp(self.o, C, (-2, -2), region)
        if failed.size:
            raise DomainError("inadmissible coarse parent in interpolation")
        L = self.L[lev]
        i0, j0, i1, j1 = region
        for jj in range(j0, j1 + 1):
            for ii in range(i0, i1 + 1):
                if mask is not None:
                    if not (0 <= ii < L["nx"] and 0 <= jj < L["ny"]) or mask[jj, ii] != 0:
                        continue
                G[:, jj + 2, ii + 2] = F[:, jj - j0, ii - i0]

    def fill(self, A, lev):
        L = self.L[lev]
        if lev > 0:
            for b in L["boxes"]:
                for r in _strips(b):
                    self._interp_into(A[lev - 1], A[lev], lev, r, L["mask"])
        fill_ghosts(A[lev], self.ns, self.bc, False, self.inlet, self.wedge_i0 << lev)

    def average(self, A):
        for lev in range(self.finest, 0, -1):
            C = self.L[lev - 1]
            cov = C["mask"] == 2
            if not cov.any():
                continue
            Fq = self.interior(lev, A[lev])
            avg = average_down(Fq, (0, 0), (0, 0, C["nx"] - 1, C["ny"] - 1))
            Cq = self.interior(lev - 1, A[lev - 1])
            Cq[:, cov] = avg[:, cov]

    def reflux(self, Cq, cfx, cfy, ffx, ffy, lev, h):
        """Flux-register corrections of level lev-1 from level lev, per leaf
        cell and face in the order x-lo, x-hi, y-lo, y-hi."""
        C = self.L[lev - 1]
        m = C["mask"]
        ny, nx = m.shape
        leaf = m == 1
        cov = m == 2
        pad = np.zeros((ny + 2, nx + 2), dtype=bool)
        pad[1:-1, 1:-1] = cov
        xlo, xhi = leaf & pad[1:-1, :-2], leaf & pad[1:-1, 2:]
        ylo, yhi = leaf & pad[:-2, 1:-1], leaf & pad[2:, 1:-1]
        dxc, dyc = C["dx"], C["dy"]
        vol = dxc * dyc
        for n in range(self.nv):
            u = Cq[n]
            # x-lo face of cell (i, j): coarse face i, fine face 2i at rows 2j, 2j+1
            r = _reg(-h * dyc, cfx[n, :, :nx], h * (0.5 * dyc), ffx[n, 0::2, 0:2 * nx:2], ffx[n, 1::2, 0:2 * nx:2])
            u[xlo] = u[xlo] + r[xlo] / vol
            r = _reg(-h * dyc, cfx[n, :, 1:], h * (0.5 * dyc), ffx[n, 0::2, 2::2], ffx[n, 1::2, 2::2])
            u[xhi] = u[xhi] + (-r[xhi]) / vol
            r = _reg(-h * dxc, cfy[n, :ny, :], h * (0.5 * dxc), ffy[n, 0:2 * ny:2, 0::2], ffy[n, 0:2 * ny:2, 1::2])
            u[ylo] = u[ylo] + r[ylo] / vol
            r = _reg(-h * dxc, cfy[n, 1:, :], h * (0.5 * dxc), ffy[n, 2::2, 0::2], ffy[n, 2::2, 1::2])
            u[yhi] = u[yhi] + (-r[yhi]) / vol

    def repair(self, A):
        """repairRoundoff over the leaf cells of every level."""
        ns = self.ns
        peak = np.zeros(ns)
        donor = [None] * ns
        mass = np.zeros(ns)
        for lev in range(self.finest + 1):
            L = self.L[lev]
            vol = L["dx"] * L["dy"]
            Q = self.interior(lev, A[lev])
            leaf = L["mask"] == 1
            for n in range(ns):
                v = np.where(leaf, Q[n], 0.0)
                peak[n] = max(peak[n], v.max())
                mv = v * vol
                k = int(np.argmax(mv))
                if mv.flat[k] > mass[n]:
                    mass[n] = mv.flat[k]
                    donor[n] = (lev, k // L["nx"], k % L["nx"])
        deficit = np.zeros(ns)
        repairs = 0
        for lev in range(self.finest + 1):
            L = self.L[lev]
            vol = L["dx"] * L["dy"]
            Q = self.interior(lev, A[lev])
            leaf = L["mask"] == 1
            for n in range(ns):
                bad = leaf & ~(Q[n] >= 0)
                for j, i in zip(*np.nonzero(bad)):
                    v = Q[n, j, i]
                    if not np.isfinite(v) or -v > 64 * EPS * peak[n]:
                        raise DomainError("materially negative species after composite flow stage")
                    deficit[n] += -v * vol
                    Q[n, j, i] = 0.0
                    repairs += 1
        for n in range(ns):
            if deficit[n] > 0:
                if donor[n] is None or not (mass[n] > deficit[n]):
                    raise DomainError("insufficient positive donor for roundoff redistribution")
                lev, j, i = donor[n]
                L = self.L[lev]
                self.interior(lev, A[lev])[n, j, i] -= deficit[n] / (L["dx"] * L["dy"])
        return repairs

    def validate(self, A):
        """validate() over every valid cell; the CFL rate over the leaves."""
        rate = 0.0
        for lev in range(self.finest + 1):
            L = self.L[lev]
            Q = self.interior(lev, A[lev])
            valid = L["mask"] > 0
            cells = Q[:, valid]
            p, a, T, st = self.o.properties(np.ascontiguousarray(cells.T))
            if np.any(st != 0):
                raise DomainError("inadmissible state")
            leaf = (L["mask"] == 1)[valid]
            rho = cells[:self.ns].sum(0)
            r = (np.abs(cells[self.ns] / rho) + a) / L["dx"] + (np.abs(cells[self.ns + 1] / rho) + a) / L["dy"]
            if leaf.any():
                rate = max(rate, float(r[leaf].max()))
        return rate

    def _walls(self, lev, box):
        L = self.L[lev]
        i0, j0, i1, j1 = box
        w = 0
        for k, touch in enumerate((i0 == 0, i1 == L["nx"] - 1, j0 == 0, j1 == L["ny"] - 1)):
            if touch and (self.bc[k] == REFLECTING or (k == 2 and self.bc[k] == WEDGE)):
                w |= 1 << k
        return w

    def euler(self, A, h):
        for lev in range(self.finest + 1):
            self.fill(A, lev)
        for first_order in (False, True):
            out = [G.copy() for G in A[:self.finest + 1]]
            fx, fy = [], []
            for lev in range(self.finest + 1):
                L = self.L[lev]
                nx, ny = L["nx"], L["ny"]
                FX, FY = np.zeros((self.nv, ny, nx + 1)), np.zeros((self.nv, ny + 1, nx))
                for b in L["boxes"]:
                    i0, j0, i1, j1 = b
                    G = np.ascontiguousarray(A[lev][:, j0:j1 + 5, i0:i1 + 5])
                    _, ob, _, _, bfx, bfy = self.o.flux_stage(G, h / L["dx"], h / L["dy"], first_order=first_order,
                                                              walls=self._walls(lev, b),
                                                              wedge_i=(self.wedge_i0 << lev) - i0, fluxes=True)
                    out[lev][:, j0 + 2:j1 + 3, i0 + 2:i1 + 3] = ob
                    FX[:, j0:j1 + 1, i0:i1 + 2] = bfx
                    FY[:, j0:j1 + 2, i0:i1 + 1] = bfy
                fx.append(FX)
                fy.append(FY)
            for lev in range(self.finest, 0, -1):
                self.reflux(self.interior(lev - 1, out[lev - 1]), fx[lev - 1], fy[lev - 1], fx[lev], fy[lev], lev, h)
            self.average(out)
            try:
                if self.repair(out):
                    self.average(out)
                self.validate(out)
                return out
            except DomainError:
                if first_order:
                    raise
                self.first_order_stages += 1
        raise DomainError("unreachable")

    def flow(self, h):
        A = [L["Q"] for L in self.L[:self.finest + 1]]
        old = [G.copy() for G in A]
        A1 = self.euler(A, h)
        A2 = self.euler(A1, h)
        for lev in range(self.finest + 1):
            self.L[lev]["Q"] = 0.5 * A2[lev] + 0.5 * old[lev]
        A = [L["Q"] for L in self.L[:self.finest + 1]]
        self.average(A)
        self.validate(A)

    def react(self, h):
        parts = []
        for lev in range(self.finest + 1):
            leaf = self.L[lev]["mask"] == 1
            parts.append(self.interior(lev)[:, leaf])
        cells = np.ascontiguousarray(np.concatenate(parts, axis=1))
        r, st, ff, msg = self.o.chem_advance(cells, np.full(cells.shape[1], h), rtol=self.rtol, atol=self.atol)
        if r != 0:
            raise DomainError(msg)
        k = 0
        for lev in range(self.finest + 1):
            leaf = self.L[lev]["mask"] == 1
            n = int(leaf.sum())
            self.interior(lev)[:, leaf] = cells[:, k:k + n]
            k += n
        A = [L["Q"] for L in self.L[:self.finest + 1]]
        self.average(A)
        self.validate(A)

    def stable_dt(self):
        return self.cfl / self.validate([L["Q"] for L in self.L[:self.finest + 1]])

    def step(self):
        dt = self.stable_dt()
        backup = [L["Q"].copy() for L in self.L[:self.finest + 1]]
        for _ in range(self.max_attempts):
            try:
                self.flow(0.5 * dt)
                self.react(dt)
                self.flow(0.5 * dt)
                break
            except DomainError:
                for lev, G in enumerate(backup):
                    self.L[lev]["Q"] = G.copy()
                self.retries += 1
                dt *= 0.5
        else:
            raise RuntimeError("time advancement failed")
        self.time += dt
        self.steps += 1
        return dt

    # ------------------------------------------------------------ totals, tags, remap
    def totals(self, absolute=False):
        s = np.zeros(self.nv)
        for lev in range(self.finest + 1):
            L = self.L[lev]
            Q = self.interior(lev)
            Q = np.abs(Q) if absolute else Q
            s += Q[:, L["mask"] == 1].sum(axis=1) * L["dx"] * L["dy"]
        return s

    def tags(self, lev, **kw):
        """ErrorEst's tags of level lev's boxes (oracle.amr.tag_cells)."""
        L = self.L[lev]
        A = [X["Q"] for X in self.L[:self.finest + 1]]
        self.fill(A, lev)
        out = np.zeros((L["ny"], L["nx"]), dtype=bool)
        for i0, j0, i1, j1 in L["boxes"]:
            G = np.ascontiguousarray(L["Q"][:, j0:j1 + 5, i0:i1 + 5])
            out[j0:j1 + 1, i0:i1 + 1] = tag_cells(self.o, G, L["dx"], L["dy"], x0=i0 * L["dx"], y0=j0 * L["dy"],
                                                  time=self.time, **kw)
        return out

    def remap(self, new_boxes):
        """regridChecked with the given boxes: new fine data by interpolation,
        the old fine cells kept where old and new boxes overlap; average,
        validate and the defect check. Returns the relative defect."""
        A = [L["Q"] for L in self.L[:self.finest + 1]]
        self.average(A)
        before, scale = self.totals(), self.totals(True)
        old_valid = [self.L[lev]["mask"] > 0 for lev in range(self.finest + 1)]
        old_q = [L["Q"].copy() for L in self.L[:self.finest + 1]]
        old_finest = self.finest
        self.set_grids(new_boxes)
        for lev in range(1, self.finest + 1):
            A = [L["Q"] for L in self.L]
            self.fill(A, lev - 1)
            G = np.zeros_like(self.L[lev]["Q"])
            for b in self.L[lev]["boxes"]:
                self._interp_into(self.L[lev - 1]["Q"], G, lev, b, None)
            if lev <= old_finest:
                keep = old_valid[lev] & (self.L[lev]["mask"] > 0)
                gi = G[:, 2:-2, 2:-2]
                gi[:, keep] = old_q[lev][:, 2:-2, 2:-2][:, keep]
            self.L[lev]["Q"] = G
        A = [L["Q"] for L in self.L[:self.finest + 1]]
        self.average(A)
        self.validate(A)
        after = self.totals()
        defect = float(np.max(np.abs(after - before) / np.maximum(scale, 1e-300)))
        if defect > 1e-8:
            raise RuntimeError("nonconservative remap")
        return defect
