"""Strang-split reacting-flow steps on one single-level box, on the GPU.

Mirrors the reference's step loop for a single level
(HyperFlow_V3/src/BlockAMR.cpp):
  run      :400-425  dt = min(target - t, stableDt()); flow(dt/2); react(dt);
                     flow(dt/2); on any failure restore the state, halve dt
                     and retry (18 attempts)
  flow     :266-270  Heun / SSP-RK2: U1 = E(U), U2 = E(U1), U = U/2 + U2/2,
                     then validate
  euler    :202-265  ghost fill, NND + HLLC stage, repairRoundoff, validate;
                     on domain_error redo the stage with first-order faces
  react    :271-285  ChemistryBatch::advance over every valid cell
  stableDt :335-346  cfl / max((|u|+a)/dx + (|v|+a)/dy)
Every arithmetic step runs in libhyperflow_hip.so (hf_euler_stage: ghost
fill, flux stage, repair and validation in one call; hf_validate_box,
hf_lincomb, hf_chem_advance); torch only holds the device arrays.

A box with ny = 1 and dy = 0 is one-dimensional: x faces only and an x-only
CFL rate (the reference itself needs ny >= 2 with reflecting y walls).
"""
from __future__ import annotations

from ctypes import byref, c_double, c_int32, c_int64

import numpy as np

from ._lib import (HF_BC_INFLOW, HF_BC_KEEP, HF_BC_PERIODIC, HF_BC_REFLECTING, HF_BC_TRANSMISSIVE,  # noqa: F401
                   HF_BC_WEDGE,
                   HF_CHEM_BATCHED, HF_EDOMAIN, HF_OK,
                   HyperflowError, check, hf_box, hf_cells, hf_flux_stats, load)
from .chemistry import ChemistryBatch


class BoxDriver:
    """Advances the conserved state of one ghosted box.

    bc: (x_lo, x_hi, y_lo, y_hi) boundary codes; y codes are ignored for a
    1-D box. Reflecting sides are walls: their face fluxes of species,
    tangential momentum and energy are zeroed (BlockAMR.cpp:224-229).
    HF_BC_INFLOW sides hold the fixed state ``inlet`` (nv conserved values);
    an HF_BC_WEDGE bottom is an outflow before column ``wedge_i`` and a wall
    from it on, in the ghost fill and in the face fluxes alike
    (FlowBoundary.hpp:25-34, BlockAMR.cpp:225-229).
    """

    def __init__(self, mechanism, nx, ny, dx, dy, bc, cfl=0.5, chem_mode=HF_CHEM_BATCHED, max_attempts=18,
                 inlet=None, wedge_i=0, graph=True):
        import torch
        self.torch = torch
        self.mech = mechanism
        self.nx, self.ny, self.dx, self.dy = int(nx), int(ny), float(dx), float(dy)
        self.one_d = self.ny == 1 and self.dy == 0.0
        self.cfl = float(cfl)
        self.max_attempts = max_attempts
        nv = mechanism.nv
        shape = (nv, self.ny + 4, self.nx + 4)
        self.U = torch.zeros(shape, dtype=torch.float64, device="cuda")
        self.U1 = torch.zeros_like(self.U)
        self.U2 = torch.zeros_like(self.U)
        self.old = torch.zeros_like(self.U)
        self.backup = torch.zeros_like(self.U)
        bc = list(bc)
        if self.one_d:
            bc[2] = bc[3] = -1
        self.bc = (c_int32 * 4)(*bc)
        walls = 0
        for k in range(4):
            if bc[k] == HF_BC_REFLECTING or (k == 2 and bc[k] == HF_BC_WEDGE):
                walls |= 1 << k
        self.wedge_i = int(wedge_i)
        self.inlet = None
        if inlet is not None:
            vals = [float(v) for v in inlet]
            if len(vals) != nv:
                raise ValueError("inlet must hold ns+3 conserved values")
            self.inlet = (c_double * nv)(*vals)
        self.walls = walls
        self.chem = ChemistryBatch(mechanism, mode=chem_mode)
        self._rate = None                  # CFL rate of the current state, when known
        self.time = 0.0
        self.steps = 0
        self.retries = 0
        self.first_order_stages = 0
        self.repairs = 0
        self._h = mechanism.handle
        # whole steps as one replayed HIP graph (hf_box_strang_step); a step the
        # graph hands back (a stage to redo first-order, a failure) is taken here
        self.graph = bool(graph)
        self.graph_steps = 0
        self.host_steps = 0

    # ------------------------------------------------------------ state
    def interior(self, A=None):
        A = self.U if A is None else A
        return A[:, 2:self.ny + 2, 2:self.nx + 2]

    def set_state(self, Q):
        """Q: (nv, ny, nx) conserved values (numpy or torch)."""
        t = self.torch
        Qt = Q if isinstance(Q, t.Tensor) else t.from_numpy(Q)
        self.interior().copy_(Qt.reshape(self.interior().shape))
        self._rate = None

    def state(self):
        return self.interior().cpu().numpy().copy()

    def _box(self, h=0.0):
        gx, gy = self.nx + 4, self.ny + 4
        plane = gx * gy
        return hf_box(self.nx, self.ny, self.walls, self.wedge_i, 0, gx, plane, 2 * gx + 2, gx, plane, h / self.dx,
                      0.0 if self.one_d else h / self.dy)

    # ------------------------------------------------------------ stages
    def _fill(self, A):
        box = self._box()
        check(self._h, load().hf_fill_ghosts_flow(self._h, A.data_ptr(), byref(box), self.bc, self.inlet))

    def _validate(self, A):
        box = self._box()
        rate, ff = c_double(0.0), c_int64(-1)
        r = load().hf_validate_box(self._h, A.data_ptr(), byref(box), self.dx, 0.0 if self.one_d else self.dy,
                                   byref(rate), byref(ff))
        check(self._h, r, ff.value)
        return rate.value

    def euler(self, src, dst, h):
        """One Euler stage src -> dst (interior), with the first-order redo:
        ghost fill, flux stage, repair and validation in one library call."""
        box = self._box(h)
        rep, fo, ff = c_int64(0), c_int32(0), c_int64(-1)
        r = load().hf_euler_stage(self._h, src.data_ptr(), dst.data_ptr(), byref(box), self.bc, self.inlet, self.dx,
                                  0.0 if self.one_d else self.dy, byref(rep), byref(fo), byref(ff))
        self.repairs += rep.value
        self.first_order_stages += fo.value
        check(self._h, r, ff.value)

    def flow(self, h):
        self.old.copy_(self.U)
        self.euler(self.U, self.U1, h)
        self.euler(self.U1, self.U2, h)
        n = self.U.numel()
        check(self._h, load().hf_lincomb(self._h, self.U2.data_ptr(), self.old.data_ptr(), 0.5, 0.5, n))
        self.U, self.U2 = self.U2, self.U
        self._rate = self._validate(self.U)

    def react(self, h):
        self._rate = None
        if self.one_d:
            view = self.U[:, 2, 2:self.nx + 2]          # (nv, nx), row-contiguous, ld = plane
            self.chem.advance(view, h)
        else:
            cells = self.interior().reshape(self.mech.nv, -1).contiguous()
            self.chem.advance(cells, h)
            self.interior().copy_(cells.reshape(self.interior().shape))

    def stable_dt(self):
        """cfl / max rate (BlockAMR::stableDt); the rate of the state a step
        ended on is the one its final validation computed."""
        rate = self._rate if self._rate is not None else self._validate(self.U)
        return self.cfl / rate

    def _graph_step(self, target):
        """hf_box_strang_step: (dt, taken) with taken 0 when the step must be
        taken with host decisions (the state is back at the step's start)."""
        box = self._box()
        dt, taken, rep = c_double(0.0), c_int32(0), c_int64(0)
        left = (target - self.time) if target is not None else -1.0
        r = load().hf_box_strang_step(self._h, self.U.data_ptr(), self.U1.data_ptr(), self.U2.data_ptr(),
                                      self.old.data_ptr(), self.backup.data_ptr(), byref(box), self.bc, self.inlet,
                                      self.dx, 0.0 if self.one_d else self.dy, self.cfl, self.chem.mode, left,
                                      byref(dt), byref(taken), byref(rep))
        check(self._h, r)
        if taken.value:
            self.repairs += rep.value
        return dt.value, taken.value

    def step(self, target=None):
        """One Strang step; returns the dt taken."""
        if self.graph:
            torch = self.torch
            torch.cuda.current_stream().synchronize()     # torch's writes to the arrays are done
            dt, taken = self._graph_step(target)
            if taken:
                self._rate = None
                self.last_graph = taken
                self.graph_steps += 1
                self.time = target if target is not None and dt >= target - self.time else self.time + dt
                self.steps += 1
                return dt
        self.host_steps += 1
        dt = self.stable_dt()
        if target is not None:
            dt = min(target - self.time, dt)
        self.backup.copy_(self.U)
        for _ in range(self.max_attempts):
            try:
                self.flow(0.5 * dt)
                self.react(dt)
                self.flow(0.5 * dt)
                break
            except HyperflowError:
                self.U.copy_(self.backup)
                self._rate = None
                self.retries += 1
                dt *= 0.5
        else:
            raise RuntimeError("time advancement failed")
        self.time = target if target is not None and dt >= target - self.time else self.time + dt
        self.steps += 1
        return dt


def shock_tube_1d(mech, nx=2000, length=0.1, hot_x=0.01, hot=(2000.0, 20.0), cold=(300.0, 1.0)):
    """The 1-D plain workload: stoichiometric H2/air (X H2:O2:N2 = 2:1:3.76,
    BlockAMR.cpp:67); x < hot_x holds the TP-equilibrium composition at
    (T, p atm) = hot, the rest the unburnt mixture at cold; at rest.
    Returns (Q (nv, 1, nx), dx)."""
    import numpy as np
    atm = 101325.0
    X = np.zeros(mech.ns)
    for name, x in (("H2", 2.0), ("O2", 1.0), ("N2", 3.76)):
        X[mech.species.index(name)] = x
    X /= X.sum()
    W = mech.W
    Y_fresh = X * W / (X * W).sum()
    Y_hot = mech.equilibrate_tp(hot[0], hot[1] * atm, X)
    q_hot = mech.state_from_TPY(hot[0], hot[1] * atm, Y_hot)
    q_cold = mech.state_from_TPY(cold[0], cold[1] * atm, Y_fresh)
    dx = length / nx
    xc = (np.arange(nx) + 0.5) * dx
    Q = np.where((xc < hot_x)[None, :], q_hot[:, None], q_cold[:, None])
    return Q.reshape(mech.nv, 1, nx).copy(), dx


def channel_2d(mech, nx=1024, ny=256, length=0.2, seed=3, hot_x=0.01, hot=(2500.0, 30.0), cold=(300.0, 1.0),
               block=4, amplitude=0.01):
    """The 2-D channel's base level: nx x ny cells over length x (length ny/nx)
    of stoichiometric H2/air (X H2:O2:N2 = 2:1:3.76, BlockAMR.cpp:67) at
    (T, p atm) = cold, with a hot strip x < hot_x at (T, p atm) = hot of the
    same fresh composition; at rest. Density is multiplied by 1 + amplitude r,
    r uniform in [-1, 1] per block x block cells (numpy default_rng(seed)), at
    fixed p (T scales by 1 / (1 + amplitude r)). Returns (Q (nv, ny, nx), dx)."""
    import numpy as np
    atm = 101325.0
    X = np.zeros(mech.ns)
    for name, x in (("H2", 2.0), ("O2", 1.0), ("N2", 3.76)):
        X[mech.species.index(name)] = x
    X /= X.sum()
    W = mech.W
    Y = X * W / (X * W).sum()
    dx = length / nx
    rng = np.random.default_rng(seed)
    nbx, nby = (nx + block - 1) // block, (ny + block - 1) // block
    r = rng.uniform(-1.0, 1.0, size=(nby, nbx))
    xc = (np.arange(nx) + 0.5) * dx
    hot_col = xc < hot_x
    Q = np.empty((mech.nv, ny, nx))
    # one state per (block, hot/cold) pair: blocks are uniform in r
    for by in range(nby):
        for bx in range(nbx):
            f = 1.0 + amplitude * r[by, bx]
            j0, i0 = by * block, bx * block
            cols = np.arange(i0, min(i0 + block, nx))
            for is_hot in (False, True):
                sel = cols[hot_col[cols] == is_hot]
                if sel.size == 0:
                    continue
                T, p = hot if is_hot else cold
                q = mech.state_from_TPY(T / f, p * atm, Y)
                Q[:, j0:j0 + block, sel] = q[:, None, None]
    return Q, dx


def teng_osw_base(mech, nx=144, ny=64):
    """The base level of the reference's oblique-shock wedge case
    (HyperFlow_V3/cases/teng_case1: 144 x 64 cells over 2.304 x 1.024 mm from
    origin (-0.128 mm, 0), inflow left and top, transmissive right, a wedge
    bottom whose tip is at x = 0; BlockAMR.cpp:62-67): premixed H2/air
    (X H2:O2:N2 = 2:1:3.76) at 958.1 K and 1 202 900 Pa moving at 2737.7 m/s,
    19 degrees below the x axis, fills the domain and the inflow. nx, ny scale
    the grid over the same domain. Returns a dict with Q (nv, ny, nx), inlet
    (nv), dx, dy, bc and wedge_i."""
    import math

    import numpy as np
    X = np.zeros(mech.ns)
    for name, x in (("H2", 2.0), ("O2", 1.0), ("N2", 3.76)):
        X[mech.species.index(name)] = x
    X /= X.sum()
    W = mech.W
    Y = X * W / (X * W).sum()
    speed, angle = 2737.7, 19.0 * math.acos(-1.0) / 180.0      # the reference's rounding (BlockAMR.cpp:64)
    inlet = mech.state_from_TPY(958.1, 1202900.0, Y, speed * math.cos(angle), -speed * math.sin(angle))
    lx, ly, x0 = 0.002304, 0.001024, -0.000128
    dx, dy = lx / nx, ly / ny
    from .amr import wedge_column
    Q = np.repeat(np.repeat(np.asarray(inlet)[:, None, None], ny, 1), nx, 2)
    return {"Q": Q, "inlet": np.asarray(inlet), "dx": dx, "dy": dy, "wedge_i": wedge_column(x0, dx, 0.0),
            "bc": (HF_BC_INFLOW, HF_BC_TRANSMISSIVE, HF_BC_WEDGE, HF_BC_INFLOW)}


class TwoLevelDriver:
    """Strang steps of a two-level composite: a single-box coarse level and
    one fixed interior fine box at ratio 2 (BlockAMR::run / flow / euler /
    react, BlockAMR.cpp:202-285, 400-425, without regrid).

    euler: coarse ghost fill; the fine box's two-cell ghost ring interpolated
    from the coarse level (ConservativeInterp, as FillPatchTwoLevels does for
    an interior box); both stages with face fluxes; reflux; average-down;
    roundoff repair and validation of both levels, with the first-order redo
    of the whole composite stage. react: one all-or-nothing chemistry batch
    over the active cells (the uncovered coarse cells and every fine cell),
    then average-down. stableDt: the largest CFL rate over both levels.

    Roundoff repair runs over the leaf cells of both levels with one donor
    per species, as the reference's does (the covered coarse cells are
    average-downs of the fine box).
    fine_box = (ilo, jlo, ihi, jhi) fine cells. A side of the box either lies
    on a non-periodic domain side (its ghosts then take that side's physical
    rule on the fine level, its faces are walls where the domain's are, and it
    has no flux register) or at least one coarse cell inside the level.
    """

    def __init__(self, mechanism, nx, ny, dx, dy, bc, fine_box, cfl=0.5, chem_mode=HF_CHEM_BATCHED,
                 max_attempts=18, inlet=None, wedge_i=0):
        import torch
        from .amr import ConservativeInterp
        from .physics import FluxSweep
        self.torch = torch
        self.mech = mechanism
        self.coarse = BoxDriver(mechanism, nx, ny, dx, dy, bc, cfl=cfl, chem_mode=chem_mode,
                                max_attempts=max_attempts, inlet=inlet, wedge_i=wedge_i)
        FI0, FJ0, FI1, FJ1 = (int(v) for v in fine_box)
        self._check_box((FI0, FJ0, FI1, FJ1))
        self.dxf, self.dyf = 0.5 * float(dx), 0.5 * float(dy)
        nv = mechanism.nv
        dev = dict(dtype=torch.float64, device="cuda")
        self.cfx = torch.zeros((nv, ny, nx + 1), **dev)
        self.cfy = torch.zeros((nv, ny + 1, nx), **dev)
        self.interp = ConservativeInterp(mechanism)
        self.sweep = FluxSweep(mechanism)
        self._set_box((FI0, FJ0, FI1, FJ1))
        self.regrids = 0
        self.cfl = float(cfl)
        self.max_attempts = max_attempts
        self.time, self.steps, self.retries, self.first_order_stages, self.repairs = 0.0, 0, 0, 0, 0
        self._h = mechanism.handle

    def _check_box(self, box):
        c = self.coarse
        FI0, FJ0, FI1, FJ1 = box
        if FI0 % 2 or FJ0 % 2 or FI1 % 2 == 0 or FJ1 % 2 == 0:
            raise ValueError("the fine box must be aligned to the refinement ratio 2")
        lo_hi = ((FI0 // 2, 0), (c.nx - 1 - FI1 // 2, 1), (FJ0 // 2, 2), (c.ny - 1 - FJ1 // 2, 3))
        for gap, k in lo_hi:
            if gap < 0 or (gap == 0 and c.bc[k] in (HF_BC_PERIODIC, -1)):
                raise ValueError("a fine box side must lie on a non-periodic domain side or inside the level")

    def margins(self):
        """Per side (x-lo, x-hi, y-lo, y-hi): 0 where a fine box may touch the
        domain side, 1 where it must stay a cell inside (periodic sides)."""
        return tuple(1 if self.coarse.bc[k] == HF_BC_PERIODIC else 0 for k in range(4))

    def _set_box(self, box):
        """Fine-level arrays, ghost strips, boundary handling and the active
        mask for box."""
        torch = self.torch
        FI0, FJ0, FI1, FJ1 = box
        c = self.coarse
        self.fine_box = tuple(box)
        self.nxf, self.nyf = FI1 - FI0 + 1, FJ1 - FJ0 + 1
        nv = self.mech.nv
        dev = dict(dtype=torch.float64, device="cuda")
        shape = (nv, self.nyf + 4, self.nxf + 4)
        self.Uf, self.Uf1, self.Uf2 = (torch.zeros(shape, **dev) for _ in range(3))
        self.oldf, self.backupf = torch.zeros(shape, **dev), torch.zeros(shape, **dev)
        self.ffx = torch.zeros((nv, self.nyf, self.nxf + 1), **dev)
        self.ffy = torch.zeros((nv, self.nyf + 1, self.nxf), **dev)
        self.strips = ((FI0 - 2, FJ0 - 2, FI1 + 2, FJ0 - 1), (FI0 - 2, FJ1 + 1, FI1 + 2, FJ1 + 2),
                       (FI0 - 2, FJ0, FI0 - 1, FJ1), (FI1 + 1, FJ0, FI1 + 2, FJ1))
        self.covered = (slice(FJ0 // 2, FJ1 // 2 + 1), slice(FI0 // 2, FI1 // 2 + 1))
        mask = torch.ones((c.ny, c.nx), dtype=torch.bool, device="cuda")
        mask[self.covered] = False
        self.active = mask.reshape(-1)
        # the coarse level's cells as the repair sees them: 1 a leaf, 2 covered by the box
        self.cmask = torch.where(mask, 1, 2).to(torch.uint8)
        touch = (FI0 == 0, FI1 == 2 * c.nx - 1, FJ0 == 0, FJ1 == 2 * c.ny - 1)
        self.fine_bc = (c_int32 * 4)(*[c.bc[k] if touch[k] else HF_BC_KEEP for k in range(4)])
        self.fine_walls = sum(1 << k for k in range(4)
                              if touch[k] and (c.bc[k] == HF_BC_REFLECTING or (k == 2 and c.bc[k] == HF_BC_WEDGE)))
        self.fine_wedge_i = 2 * c.wedge_i - FI0
        self.reflux_sides = sum(1 << k for k in range(4) if not touch[k])

    # ------------------------------------------------------------ state
    def set_state(self, Q):
        """Q: coarse conserved values (nv, ny, nx); the fine box starts as
        their conservative interpolation."""
        self.coarse.set_state(Q)
        self.coarse._fill(self.coarse.U)
        self.interp.interp(self.coarse.U, (-2, -2), self.Uf, (self.fine_box[0] - 2, self.fine_box[1] - 2),
                           self.fine_box)

    def fine_interior(self, A=None):
        A = self.Uf if A is None else A
        return A[:, 2:self.nyf + 2, 2:self.nxf + 2]

    def state(self):
        return self.coarse.state(), self.fine_interior().cpu().numpy().copy()

    # ------------------------------------------------------------ stages
    def _fbox(self, h=0.0):
        gx, gy = self.nxf + 4, self.nyf + 4
        plane = gx * gy
        return hf_box(self.nxf, self.nyf, self.fine_walls, self.fine_wedge_i, 0, gx, plane, 2 * gx + 2, gx, plane,
                      h / self.dxf, h / self.dyf)

    @staticmethod
    def _interior_patch(nx, ny, lo=(0, 0)):
        """hf_patch of a ghosted (nv, ny+4, nx+4) array's valid cells, cell
        (lo) at its first valid element."""
        from ._lib import hf_patch
        gx, gy = nx + 4, ny + 4
        return hf_patch(2 * gx + 2 - (int(lo[1]) * gx + int(lo[0])), gx, gx * gy)

    def _average(self, Ac, Af):
        """Average-down of the fine box into the coarse cells it covers, in place
        on the ghosted arrays."""
        FI0, FJ0, FI1, FJ1 = self.fine_box
        c = self.coarse
        fp = self._interior_patch(self.nxf, self.nyf, (FI0, FJ0))
        cp = self._interior_patch(c.nx, c.ny)
        reg = (c_int32 * 4)(FI0 // 2, FJ0 // 2, FI1 // 2, FJ1 // 2)
        check(self._h, load().hf_average_down(self._h, Af.data_ptr(), byref(fp), Ac.data_ptr(), byref(cp), reg))

    def _validate_fine(self, A):
        box = self._fbox()
        rate, ff = c_double(0.0), c_int64(-1)
        r = load().hf_validate_box(self._h, A.data_ptr(), byref(box), self.dxf, self.dyf, byref(rate), byref(ff))
        check(self._h, r, ff.value)
        return rate.value

    def _composite_once(self, src_c, src_f, dst_c, dst_f, h, first_order):
        from .amr import face_fluxes
        c = self.coarse
        FI0, FJ0, _, _ = self.fine_box
        c._fill(src_c)
        for r in self.strips:
            self.interp.interp(src_c, (-2, -2), src_f, (FI0 - 2, FJ0 - 2), r)
        if self.reflux_sides != 15:
            # ghosts beyond the domain: the physical rule on the fine level
            fb = self._fbox()
            check(self._h, load().hf_fill_ghosts_flow(self._h, src_f.data_ptr(), byref(fb), self.fine_bc, c.inlet))
        # both stages write straight into the ghosted outputs' valid cells
        for (src, dst, box, fl) in ((src_c, dst_c, c._box(h), (self.cfx, self.cfy)),
                                    (src_f, dst_f, self._fbox(h), (self.ffx, self.ffy))):
            try:
                self.sweep.stage([box], src, dst, first_order=first_order, fluxes=[fl])
            except HyperflowError as e:
                # an inadmissible raw update is repaired or rejected below, as in BoxDriver
                if e.code != HF_EDOMAIN:
                    raise
        cf, ff_ = face_fluxes(self.cfx, self.cfy), face_fluxes(self.ffx, self.ffy, (FI0, FJ0))
        cp = self._interior_patch(c.nx, c.ny)
        fb = (c_int32 * 4)(*self.fine_box)
        check(self._h, load().hf_reflux(self._h, dst_c.data_ptr(), byref(cp), byref(cf), byref(ff_), fb,
                                        self.reflux_sides, h, c.dx, c.dy))
        self._average(dst_c, dst_f)
        # repairRoundoff over the leaf cells of both levels, one donor per species
        # (BlockAMR.cpp:299-334): the covered coarse cells take the average-down
        rep, ff = c_int64(0), c_int64(-1)
        r = load().hf_repair_cells(self._h, self._leaf_blocks(dst_c, dst_f), 2, byref(rep), byref(ff))
        if r != HF_OK:
            return r, ff.value
        repaired = rep.value
        self.repairs += repaired
        if repaired:
            self._average(dst_c, dst_f)
        for A, box, dx, dy in ((dst_c, c._box(), c.dx, c.dy), (dst_f, self._fbox(), self.dxf, self.dyf)):
            r = load().hf_validate_box(self._h, A.data_ptr(), byref(box), dx, dy, None, byref(ff))
            if r != HF_OK:
                return r, ff.value
        return HF_OK, -1

    def _leaf_blocks(self, A_c, A_f):
        """hf_cells of the coarse level (mask: leaves and covered cells) and of
        the fine box (all leaves), in level order."""
        c = self.coarse
        gx, gy = c.nx + 4, c.ny + 4
        gxf, gyf = self.nxf + 4, self.nyf + 4
        return (hf_cells * 2)(hf_cells(A_c.data_ptr(), 2 * gx + 2, gx, gx * gy, c.nx, c.ny, self.cmask.data_ptr(),
                                       c.nx, c.dx, c.dy),
                              hf_cells(A_f.data_ptr(), 2 * gxf + 2, gxf, gxf * gyf, self.nxf, self.nyf, None, 0,
                                       self.dxf, self.dyf))

    def euler(self, src_c, src_f, dst_c, dst_f, h):
        r, ff = self._composite_once(src_c, src_f, dst_c, dst_f, h, False)
        if r == HF_EDOMAIN:
            self.first_order_stages += 1
            r, ff = self._composite_once(src_c, src_f, dst_c, dst_f, h, True)
        check(self._h, r, ff)

    def flow(self, h):
        c = self.coarse
        c.old.copy_(c.U)
        self.oldf.copy_(self.Uf)
        self.euler(c.U, self.Uf, c.U1, self.Uf1, h)
        self.euler(c.U1, self.Uf1, c.U2, self.Uf2, h)
        for A, old in ((c.U2, c.old), (self.Uf2, self.oldf)):
            check(self._h, load().hf_lincomb(self._h, A.data_ptr(), old.data_ptr(), 0.5, 0.5, A.numel()))
        c.U, c.U2 = c.U2, c.U
        self.Uf, self.Uf2 = self.Uf2, self.Uf
        self._average(c.U, self.Uf)
        c._validate(c.U)
        self._validate_fine(self.Uf)

    def react(self, h):
        c = self.coarse
        nv = self.mech.nv
        Cc = c.interior().reshape(nv, -1)
        Ff = self.fine_interior().reshape(nv, -1)
        batch = self.torch.cat([Cc[:, self.active], Ff], dim=1).contiguous()
        c.chem.advance(batch, h)
        na = int(self.active.sum().item())
        Cc2 = Cc.clone()
        Cc2[:, self.active] = batch[:, :na]
        c.interior().copy_(Cc2.reshape(c.interior().shape))
        self.fine_interior().copy_(batch[:, na:].reshape(self.fine_interior().shape))
        self._average(c.U, self.Uf)
        c._validate(c.U)
        self._validate_fine(self.Uf)

    def stable_dt(self):
        return self.cfl / max(self.coarse._validate(self.coarse.U), self._validate_fine(self.Uf))

    def _levels_totals(self):
        from .output import Level, totals
        c = self.coarse
        Qc = c.interior().cpu().numpy()
        Qf = self.fine_interior().cpu().numpy()
        act = self.active.reshape(c.ny, c.nx).cpu().numpy()
        z = lambda Q: np.zeros(Q.shape[1:])
        return totals([Level(0, Qc, z(Qc), z(Qc), (0, 0), c.dx, c.dy, (0, 0), act),
                       Level(1, Qf, z(Qf), z(Qf), self.fine_box[:2], self.dxf, self.dyf, (0, 0))]), \
            totals([Level(0, Qc, z(Qc), z(Qc), (0, 0), c.dx, c.dy, (0, 0), act),
                    Level(1, Qf, z(Qf), z(Qf), self.fine_box[:2], self.dxf, self.dyf, (0, 0))], absolute=True)

    def regrid(self, buffer=2, blocking=4, **tag_kw):
        """Move the refined box to the tagged region (BlockAMR::regridChecked,
        BlockAMR.cpp:367-373, with ErrorEst's criteria, :170-200): tags of the
        coarse level, one box around them (choose_fine_box), the new fine data
        by conservative interpolation with the old fine cells kept where the
        boxes overlap, average-down, and the remap defect check (relative
        change of any conserved integral above 1e-8 raises). Returns the new
        box (unchanged when nothing is tagged) and the defect."""
        from .amr import choose_fine_box, tag_cells
        c = self.coarse
        self._average(c.U, self.Uf)
        before, scale = self._levels_totals()
        c._fill(c.U)
        tags = tag_cells(self.mech, c.U, c.dx, c.dy, time=self.time, **tag_kw).cpu().numpy().astype(bool)
        box = choose_fine_box(tags, buffer, blocking, c.nx, c.ny, self.margins())
        if box is None or box == self.fine_box:
            return self.fine_box, 0.0
        old_box, old_f = self.fine_box, self.fine_interior().clone()
        self._set_box(box)
        self.interp.interp(c.U, (-2, -2), self.Uf, (box[0] - 2, box[1] - 2), box)
        # keep the old fine cells where the boxes overlap
        a0, b0 = max(box[0], old_box[0]), max(box[1], old_box[1])
        a1, b1 = min(box[2], old_box[2]), min(box[3], old_box[3])
        if a0 <= a1 and b0 <= b1:
            self.Uf[:, 2 + b0 - box[1]:3 + b1 - box[1], 2 + a0 - box[0]:3 + a1 - box[0]] = \
                old_f[:, b0 - old_box[1]:b1 - old_box[1] + 1, a0 - old_box[0]:a1 - old_box[0] + 1]
        self._average(c.U, self.Uf)
        after, _ = self._levels_totals()
        defect = float(np.max(np.abs(after - before) / np.maximum(scale, 1e-300)))
        if defect > 1e-8:
            raise RuntimeError("nonconservative remap")
        c._validate(c.U)
        self._validate_fine(self.Uf)
        self.regrids += 1
        return self.fine_box, defect

    def step(self, target=None):
        """One Strang step of both levels; returns the dt taken."""
        c = self.coarse
        dt = self.stable_dt()
        if target is not None:
            dt = min(target - self.time, dt)
        c.backup.copy_(c.U)
        self.backupf.copy_(self.Uf)
        for _ in range(self.max_attempts):
            try:
                self.flow(0.5 * dt)
                self.react(dt)
                self.flow(0.5 * dt)
                break
            except HyperflowError:
                c.U.copy_(c.backup)
                self.Uf.copy_(self.backupf)
                self.retries += 1
                dt *= 0.5
        else:
            raise RuntimeError("time advancement failed")
        self.time = target if target is not None and dt >= target - self.time else self.time + dt
        self.steps += 1
        return dt
