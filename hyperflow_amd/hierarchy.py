"""Strang-split reacting-flow steps on a block-structured AMR hierarchy, on
the GPU (BlockAMR::run / flow / euler / react / regridChecked,
HyperFlow_V3/src/BlockAMR.cpp:202-298, 299-346, 367-373, 400-425), with any
number of levels at ratio 2 and any number of boxes per fine level.

Layout. Each level is one ghosted array over its whole domain in HBM,
(nv, ny_l + 4, nx_l + 4); the level's boxes are index rectangles in it, and
a uint8 mask per level marks each cell 0 (no box), 1 (a leaf) or 2
(covered by the next finer level). Boxes of one level therefore see each
other's valid cells as ghosts without any exchange. A level-2 array of the
1024 x 256 channel is 4096 x 1024 cells, 0.4 GB per copy: the hierarchy's
five state copies and two face-flux arrays per level stay far below one
GPU's 288 GB.

One Euler stage (BlockAMR::euler) on every level, coarse to fine:
  ghosts   the ghost cells of a level's boxes that no box of the level holds
           are interpolated from the next coarser level (ConservativeInterp
           through FillPatchTwoLevels, one launch over every box's four ghost
           strips, masked to the uncovered cells), then the physical rule
           fills the level array's outer ring (FlowBoundary and the domain's
           boundary types, hf_fill_ghosts_flow);
  fluxes   one flux stage per level over all its boxes, face fluxes into the
           level's face arrays;
then, finest to coarsest, the flux-register correction of every leaf cell
beside a covered cell (hf_reflux_masked) and the average-down of the covered
cells (hf_average_down_masked); roundoff repair over the leaf cells of all
levels (hf_repair_cells) and validation of every valid cell
(hf_validate_cells), with the first-order redo of the whole composite stage.
react: one chemistry batch over the leaf cells of every level, in place
(hf_chem_advance_cells). stableDt: the validation's CFL rate over the leaves.
regrid: ErrorEst's tags per box (hf_tag_cells), new boxes by
hyperflow_amd.amr_grid.make_grids, new fine data by interpolation with the
old fine cells kept where old and new boxes overlap, average-down and the
remap defect check (relative change of a conserved integral above 1e-8
raises).

Restrictions: a fine box may not touch a periodic side (its flux register
and ghosts would have to wrap); the blocking factor is at least 4, so a fine
box either touches a domain side or stays at least two coarse cells away.
"""
from __future__ import annotations

from ctypes import byref, c_double, c_int32, c_int64

import numpy as np

from . import amr_grid
from ._lib import (HF_BC_INFLOW, HF_BC_KEEP, HF_BC_PERIODIC, HF_BC_REFLECTING, HF_BC_WEDGE, HF_CHEM_BATCHED, HF_EDOMAIN,  # noqa
                   HF_OK, HyperflowError, check, hf_box, hf_cells, hf_face_fluxes, hf_flux_stats, hf_level,
                   hf_patch, hf_tag_params, load)


class _Level:
    """Arrays and tables of one level."""

    def __init__(self, torch, nv, lev, nx, ny, dx, dy):
        dev = dict(dtype=torch.float64, device="cuda")
        self.lev, self.nx, self.ny, self.dx, self.dy = lev, nx, ny, dx, dy
        self.gx, self.gy = nx + 4, ny + 4
        shape = (nv, ny + 4, nx + 4)
        self.A = {k: torch.zeros(shape, **dev) for k in ("U", "U1", "U2", "old", "backup")}
        self.fx = torch.zeros((nv, ny, nx + 1), **dev)
        self.fy = torch.zeros((nv, ny + 1, nx), **dev)
        self.mask = torch.zeros((ny, nx), dtype=torch.uint8, device="cuda")
        self.mask_host = np.zeros((ny, nx), dtype=np.uint8)
        self.boxes = []

    def patch(self):
        """hf_patch of a level array in the level's index space."""
        return hf_patch(2 * self.gx + 2, self.gx, self.gx * self.gy)

    def cells(self, A):
        """One hf_cells block per box of the level (valid cells only, box
        order, i fastest: the reference's MFIter / BoxIterator order), with
        the level mask telling leaves from covered cells."""
        base, mask = A.data_ptr(), self.mask.data_ptr()
        return [hf_cells(base, (j0 + 2) * self.gx + i0 + 2, self.gx, self.gx * self.gy, i1 - i0 + 1, j1 - j0 + 1,
                         mask + j0 * self.nx + i0, self.nx, self.dx, self.dy) for (i0, j0, i1, j1) in self.boxes]

    def interior(self, A):
        return A[:, 2:self.ny + 2, 2:self.nx + 2]


class AmrDriver:
    """Advances a reacting-flow hierarchy on one GPU.

    nx0, ny0, dx0, dy0: level 0 (the whole domain, one box). bc: (x_lo, x_hi,
    y_lo, y_hi) boundary codes; reflecting sides (and a wedge bottom from
    column wedge_i on) are walls. max_level, blocking, max_grid_size,
    grid_eff, n_error_buf, n_proper: the regrid parameters of main.cpp:33-36
    (n_proper 2 is the reference's). tag: ErrorEst's criteria for
    hf_tag_cells (pressure_tag, temperature_tag, fuel_rate_tag; box, a tag
    rectangle (x0, x1, y0, y1) moving with velocity (u, v); mode "physics"
    (the criteria, plus the box if given) or "box" (the box alone);
    BlockAMR.cpp:170-200). chemistry False skips react and the fuel-rate
    criterion (the configuration's "chemistry", BlockAMR.cpp:271-272).
    graph True takes each step as one replayed HIP graph
    (hf_amr_strang_step, recaptured when the boxes change), bit-identical to
    the call-by-call step, which serves a step the graph hands back (a
    first-order redo or a halved dt) and graph False."""

    def __init__(self, mechanism, nx0, ny0, dx0, dy0, bc, max_level=2, cfl=0.5, chem_mode=HF_CHEM_BATCHED,
                 blocking=16, max_grid_size=64, grid_eff=0.7, n_error_buf=2, n_proper=2, regrid_interval=8,
                 max_attempts=18, inlet=None, wedge_i=0, tag=None, x0=0.0, y0=0.0, chemistry=True, graph=True):
        import torch
        self.torch = torch
        self.mech = mechanism
        self._h = mechanism.handle
        self.nv, self.ns = mechanism.nv, mechanism.ns
        self.nx0, self.ny0, self.dx0, self.dy0 = int(nx0), int(ny0), float(dx0), float(dy0)
        if blocking < 4 or blocking % 2 or self.nx0 % blocking or self.ny0 % blocking:
            raise ValueError("blocking_factor must be even, at least 4, and divide the base cells")
        self.bc = tuple(int(b) for b in bc)
        self.max_level = int(max_level)
        self.cfl, self.chem_mode, self.max_attempts = float(cfl), chem_mode, int(max_attempts)
        self.blocking, self.max_grid_size, self.grid_eff = int(blocking), int(max_grid_size), float(grid_eff)
        self.n_error_buf, self.n_proper, self.regrid_interval = int(n_error_buf), int(n_proper), int(regrid_interval)
        self.wedge_i0 = int(wedge_i)
        self.x0, self.y0 = float(x0), float(y0)
        self.chemistry = bool(chemistry)
        self.graph = bool(graph) and HF_BC_KEEP not in self.bc
        self.graph_steps = 0
        self.tag = dict(pressure_tag=0.08, temperature_tag=0.08, fuel_rate_tag=50.0, chemistry=self.chemistry,
                        box=None, velocity=(0.0, 0.0), mode="physics")
        self.tag.update(tag or {})
        if self.tag["mode"] not in ("physics", "box") or (self.tag["mode"] == "box" and self.tag["box"] is None):
            raise ValueError("invalid tagging geometry")
        if (self.tag["box"] is not None and len(self.tag["box"]) != 4) or len(self.tag["velocity"]) != 2:
            raise ValueError("invalid tagging geometry")
        self.inlet = None
        if inlet is not None:
            vals = [float(v) for v in inlet]
            if len(vals) != self.nv:
                raise ValueError("inlet must hold ns+3 conserved values")
            self.inlet = (c_double * self.nv)(*vals)
        # fine boxes keep one block away from periodic sides
        self.margins = tuple(blocking // 2 if self.bc[k] == HF_BC_PERIODIC else 0 for k in range(4))
        self.levels = [_Level(torch, self.nv, lev, self.nx0 << lev, self.ny0 << lev, self.dx0 / (1 << lev),
                              self.dy0 / (1 << lev)) for lev in range(self.max_level + 1)]
        self.finest = 0
        self.time, self.steps, self.retries, self.first_order_stages, self.repairs = 0.0, 0, 0, 0, 0
        self.regrids, self.grid_changes, self.max_remap_defect, self.chem_cells = 0, 0, 0.0, 0
        self.set_grids([[]])

    # ------------------------------------------------------------ grids
    def grids(self):
        return [[]] + [list(L.boxes) for L in self.levels[1:self.finest + 1]]

    def set_grids(self, boxes):
        """Boxes per level (boxes[0] ignored), masks and per-level tables."""
        masks = amr_grid.level_masks(list(boxes) + [[]], self.nx0, self.ny0)
        self.finest = len(masks) - 1
        for lev, L in enumerate(self.levels):
            if lev <= self.finest:
                L.boxes = [(0, 0, L.nx - 1, L.ny - 1)] if lev == 0 else [tuple(b) for b in boxes[lev]]
                L.mask_host = masks[lev]
                L.mask.copy_(self.torch.from_numpy(masks[lev]))
            else:
                L.boxes = []
        for lev in range(1, self.finest + 1):
            L = self.levels[lev]
            regions = [r for b in L.boxes for r in amr_grid.ghost_strips(b)]
            L.regions = (c_int32 * (4 * len(regions)))(*[v for r in regions for v in r])
            L.nregion = len(regions)
        for L in self.levels[:self.finest + 1]:
            L.level_bc = (c_int32 * 4)(*self.bc)
            L.wedge_i = self.wedge_i0 << L.lev
            L.box_table = (c_int32 * (4 * len(L.boxes)))(*[v for b in L.boxes for v in b])
        self._leaves = int(sum((L.mask_host == 1).sum() for L in self.levels[:self.finest + 1]))
        self._stored = int(sum((L.mask_host > 0).sum() for L in self.levels[:self.finest + 1]))

    def _box_table(self, L, h):
        """hf_box and face-flux descriptors of every box of level L."""
        boxes, fls = [], []
        for (i0, j0, i1, j1) in L.boxes:
            walls = 0
            for k, touch in enumerate((i0 == 0, i1 == L.nx - 1, j0 == 0, j1 == L.ny - 1)):
                if touch and (self.bc[k] == HF_BC_REFLECTING or (k == 2 and self.bc[k] == HF_BC_WEDGE)):
                    walls |= 1 << k
            boxes.append(hf_box(i1 - i0 + 1, j1 - j0 + 1, walls, L.wedge_i - i0, j0 * L.gx + i0, L.gx, L.gx * L.gy,
                                (j0 + 2) * L.gx + i0 + 2, L.gx, L.gx * L.gy, h / L.dx, h / L.dy))
            # box-local face (a, b) is the level's face (i0 + a, j0 + b)
            fls.append(hf_face_fluxes(L.fx.data_ptr(), j0 * (L.nx + 1) + i0, L.nx + 1, L.ny * (L.nx + 1),
                                      L.fy.data_ptr(), j0 * L.nx + i0, L.nx, (L.ny + 1) * L.nx))
        n = len(boxes)
        return (hf_box * n)(*boxes), (hf_face_fluxes * n)(*fls), n

    # ------------------------------------------------------------ state
    def set_state(self, Q):
        """Level 0's conserved values (nv, ny0, nx0); the fine levels are
        their conservative interpolation."""
        t = self.torch
        L0 = self.levels[0]
        Qt = Q if isinstance(Q, t.Tensor) else t.from_numpy(np.ascontiguousarray(Q))
        L0.interior(L0.A["U"]).copy_(Qt.reshape(self.nv, L0.ny, L0.nx))
        for lev in range(1, self.finest + 1):
            self._fill("U", lev - 1)
            self._interp_boxes(lev, "U", self.levels[lev].A["U"])
        self._average("U")

    def level_state(self, lev, A="U"):
        """(Q (nv, ny_l, nx_l) host, mask (ny_l, nx_l)) of a level."""
        L = self.levels[lev]
        return L.interior(L.A[A]).cpu().numpy().copy(), L.mask_host.copy()

    # ------------------------------------------------------------ pieces
    def _fill(self, key, lev):
        """Ghosts of level lev's boxes: the uncovered ones from level lev-1,
        then the physical rule on the level array's outer ring."""
        L = self.levels[lev]
        A = L.A[key]
        if lev > 0 and L.nregion:
            C = self.levels[lev - 1]
            cp, fp = C.patch(), L.patch()
            ext = (c_int32 * 2)(L.nx, L.ny)
            lim, ff = c_int64(0), c_int64(-1)
            r = load().hf_interp_regions(self._h, C.A[key].data_ptr(), byref(cp), A.data_ptr(), byref(fp), L.regions,
                                         L.nregion, L.mask.data_ptr(), L.nx, ext, byref(lim), byref(ff))
            check(self._h, r, ff.value)
        box = hf_box(L.nx, L.ny, 0, L.wedge_i, 0, L.gx, L.gx * L.gy, 0, L.nx, L.nx * L.ny, 0.0, 0.0)
        check(self._h, load().hf_fill_ghosts_flow(self._h, A.data_ptr(), byref(box), L.level_bc, self.inlet))

    def _interp_boxes(self, lev, key, target):
        """Every valid cell of level lev's boxes in ``target`` from level lev-1's ``key`` array."""
        L, C = self.levels[lev], self.levels[lev - 1]
        regs = (c_int32 * (4 * len(L.boxes)))(*[v for b in L.boxes for v in b])
        cp, fp = C.patch(), L.patch()
        lim, ff = c_int64(0), c_int64(-1)
        r = load().hf_interp_regions(self._h, C.A[key].data_ptr(), byref(cp), target.data_ptr(), byref(fp), regs,
                                     len(L.boxes), None, 0, None, byref(lim), byref(ff))
        check(self._h, r, ff.value)

    def _average(self, key):
        for lev in range(self.finest, 0, -1):
            F, C = self.levels[lev], self.levels[lev - 1]
            fp, cp = F.patch(), C.patch()
            reg = (c_int32 * 4)(0, 0, C.nx - 1, C.ny - 1)
            check(self._h, load().hf_average_down_masked(self._h, F.A[key].data_ptr(), byref(fp), C.A[key].data_ptr(),
                                                         byref(cp), C.mask.data_ptr(), C.nx, reg))

    def _cells(self, key):
        blocks = [b for lev in range(self.finest + 1) for b in self.levels[lev].cells(self.levels[lev].A[key])]
        n = len(blocks)
        return (hf_cells * n)(*blocks), n

    def _validate(self, key):
        blocks, n = self._cells(key)
        rate, ff = c_double(0.0), c_int64(-1)
        check(self._h, load().hf_validate_cells(self._h, blocks, n, byref(rate), byref(ff)), ff.value)
        return rate.value

    def _composite_once(self, src, dst, h, first_order):
        for lev in range(self.finest + 1):
            self._fill(src, lev)
        for lev in range(self.finest + 1):
            L = self.levels[lev]
            boxes, fls, n = self._box_table(L, h)
            st, ff = hf_flux_stats(), c_int64(-1)
            r = load().hf_flux_stage_fluxes(self._h, boxes, n, L.A[src].data_ptr(), L.A[dst].data_ptr(),
                                            int(bool(first_order)), fls, byref(st), byref(ff))
            if r not in (HF_OK, HF_EDOMAIN):      # an inadmissible raw update is repaired or rejected below
                check(self._h, r, ff.value)
        for lev in range(self.finest, 0, -1):
            F, C = self.levels[lev], self.levels[lev - 1]
            cf = hf_face_fluxes(C.fx.data_ptr(), 0, C.nx + 1, C.ny * (C.nx + 1), C.fy.data_ptr(), 0, C.nx,
                                (C.ny + 1) * C.nx)
            fl = hf_face_fluxes(F.fx.data_ptr(), 0, F.nx + 1, F.ny * (F.nx + 1), F.fy.data_ptr(), 0, F.nx,
                                (F.ny + 1) * F.nx)
            cp = C.patch()
            reg = (c_int32 * 4)(0, 0, C.nx - 1, C.ny - 1)
            check(self._h, load().hf_reflux_masked(self._h, C.A[dst].data_ptr(), byref(cp), byref(cf), byref(fl),
                                                   C.mask.data_ptr(), C.nx, reg, h, C.dx, C.dy))
        self._average(dst)
        blocks, n = self._cells(dst)
        rep, ff = c_int64(0), c_int64(-1)
        r = load().hf_repair_cells(self._h, blocks, n, byref(rep), byref(ff))
        if r != HF_OK:
            return r, ff.value
        self.repairs += rep.value
        if rep.value:
            self._average(dst)
        rate, ff = c_double(0.0), c_int64(-1)
        r = load().hf_validate_cells(self._h, blocks, n, byref(rate), byref(ff))
        return r, ff.value

    def euler(self, src, dst, h):
        r, ff = self._composite_once(src, dst, h, False)
        if r == HF_EDOMAIN:
            self.first_order_stages += 1
            r, ff = self._composite_once(src, dst, h, True)
        check(self._h, r, ff)

    def _copy(self, dst, src):
        for L in self.levels[:self.finest + 1]:
            L.A[dst].copy_(L.A[src])

    def _swap(self, a, b):
        for L in self.levels[:self.finest + 1]:
            L.A[a], L.A[b] = L.A[b], L.A[a]

    def flow(self, h):
        self._copy("old", "U")
        self.euler("U", "U1", h)
        self.euler("U1", "U2", h)
        for L in self.levels[:self.finest + 1]:
            check(self._h, load().hf_lincomb(self._h, L.A["U2"].data_ptr(), L.A["old"].data_ptr(), 0.5, 0.5,
                                             L.A["U2"].numel()))
        self._swap("U", "U2")
        self._average("U")
        self._rate = self._validate("U")

    def react(self, h):
        if not self.chemistry:
            return
        self._rate = None
        blocks, n = self._cells("U")
        from ._lib import hf_chem_stats
        st, ff = hf_chem_stats(), c_int64(-1)
        dt = (c_double * 1)(h)
        r = load().hf_chem_advance_cells(self._h, blocks, n, dt, self.chem_mode, byref(st), byref(ff))
        self.last_chem_stats = st.as_dict()
        check(self._h, r, ff.value)
        self.chem_cells += self.leaf_count()
        self._average("U")
        self._validate("U")

    def stable_dt(self):
        rate = getattr(self, "_rate", None)
        if rate is None:
            rate = self._validate("U")
        return self.cfl / rate

    def leaf_count(self):
        return self._leaves

    def stored_count(self):
        return self._stored

    def _graph_step(self, target):
        """The whole step as one HIP graph (hf_amr_strang_step). Returns dt,
        or None when the graph handed the step back with U restored."""
        levels = []
        for L in self.levels[:self.finest + 1]:
            A = L.A
            levels.append(hf_level(A["U"].data_ptr(), A["U1"].data_ptr(), A["U2"].data_ptr(), A["old"].data_ptr(),
                                   A["backup"].data_ptr(), L.fx.data_ptr(), L.fy.data_ptr(), L.mask.data_ptr(),
                                   L.nx, L.ny, L.dx, L.dy, L.box_table, len(L.boxes),
                                   L.regions if L.lev > 0 else None, L.nregion if L.lev > 0 else 0, L.wedge_i))
        table = (hf_level * len(levels))(*levels)
        dt, taken, rep = c_double(0.0), c_int32(0), c_int64(0)
        left = target - self.time if target is not None else 0.0
        r = load().hf_amr_strang_step(self._h, table, len(levels), self.levels[0].level_bc, self.inlet, self.cfl,
                                      int(self.chemistry), self.chem_mode, left, byref(dt), byref(taken), byref(rep))
        check(self._h, r)
        self._rate = None
        if not taken.value:
            return None
        self.graph_steps += 1
        self.repairs += rep.value
        if self.chemistry:
            self.chem_cells += self._leaves
        return dt.value

    def step(self, target=None):
        """One Strang step of the hierarchy, with the reference's halve-and-
        retry; a regrid follows every regrid_interval steps. Returns dt."""
        dt = self._graph_step(target) if self.graph else None
        if dt is None:
            dt = self._host_step(target)
        self.time = target if target is not None and dt >= target - self.time else self.time + dt
        self.steps += 1
        if self.max_level > 0 and self.regrid_interval > 0 and self.steps % self.regrid_interval == 0:
            self.regrid()
        return dt

    def _host_step(self, target):
        """The step call by call, with host decisions (the first-order redo
        and the halving of dt)."""
        dt = self.stable_dt()
        if target is not None:
            dt = min(target - self.time, dt)
        self._copy("backup", "U")
        for _ in range(self.max_attempts):
            try:
                self.flow(0.5 * dt)
                self.react(dt)
                self.flow(0.5 * dt)
                break
            except HyperflowError:
                self._copy("U", "backup")
                self._rate = None
                self.retries += 1
                dt *= 0.5
        else:
            raise RuntimeError("time advancement failed")
        return dt

    # ------------------------------------------------------------ regrid
    def totals(self, absolute=False):
        """Conserved integrals over the leaf cells (BlockAMR::totals, one
        device reduction over every box: hf_totals_cells)."""
        blocks, n = self._cells("U")
        out = (c_double * self.nv)()
        check(self._h, load().hf_totals_cells(self._h, blocks, n, int(bool(absolute)), out))
        return np.array(out[:])

    def tags(self, lev):
        """ErrorEst's tags on level lev's valid cells (host bool array)."""
        L = self.levels[lev]
        self._fill("U", lev)
        out = np.zeros((L.ny, L.nx), dtype=bool)
        k = self.tag
        # every box's tags go to one device buffer, read back once
        sizes = [(i1 - i0 + 1) * (j1 - j0 + 1) for (i0, j0, i1, j1) in L.boxes]
        tg_all = self.torch.zeros(sum(sizes), dtype=self.torch.uint8, device="cuda")
        at = 0
        for (i0, j0, i1, j1), size in zip(L.boxes, sizes):
            nx, ny = i1 - i0 + 1, j1 - j0 + 1
            box = hf_box(nx, ny, 0, 0, j0 * L.gx + i0, L.gx, L.gx * L.gy, 0, nx, nx * ny, 0.0, 0.0)
            box_ = k["box"]
            P = hf_tag_params(float(k["pressure_tag"]), float(k["temperature_tag"]), float(k["fuel_rate_tag"]),
                              int(k["mode"] == "physics"), int(bool(k["chemistry"])), 0, int(box_ is not None),
                              (c_double * 4)(*(box_ if box_ is not None else (0.0,) * 4)),
                              (c_double * 2)(*k["velocity"]), self.x0 + i0 * L.dx, self.y0 + j0 * L.dy, L.dx, L.dy,
                              self.time)
            n, ff = c_int64(0), c_int64(-1)
            r = load().hf_tag_cells(self._h, L.A["U"].data_ptr(), byref(box), byref(P), tg_all[at:].data_ptr(),
                                    byref(n), byref(ff))
            check(self._h, r, ff.value)
            at += size
        host = tg_all.cpu().numpy().astype(bool)
        at = 0
        for (i0, j0, i1, j1), size in zip(L.boxes, sizes):
            out[j0:j1 + 1, i0:i1 + 1] |= host[at:at + size].reshape(j1 - j0 + 1, i1 - i0 + 1)
            at += size
        return out

    def regrid(self):
        """BlockAMR::regridChecked: new boxes from the tags, new fine data,
        average-down, validation and the remap defect check. Returns the
        relative defect."""
        self._average("U")
        before, scale = self.totals(), self.totals(True)
        # tags of the levels that may be refined (make_grids reads no others)
        tags = [self.tags(lev) for lev in range(min(self.finest, self.max_level - 1) + 1)]
        valid = [L.mask_host > 0 for L in self.levels[:self.finest + 1]]
        new = amr_grid.make_grids(tags, valid, self.max_level, self.blocking, self.max_grid_size, self.grid_eff,
                                  self.n_error_buf, self.n_proper, self.margins)
        old = self.grids()
        if [list(map(tuple, b)) for b in new[1:]] == [list(map(tuple, b)) for b in (old + [[]] * 8)[1:len(new)]]:
            self.regrids += 1          # counted as the reference counts every regridChecked
            return 0.0
        self.grid_changes += 1
        old_valid = [L.mask_host > 0 for L in self.levels[:self.finest + 1]]
        old_finest = self.finest
        self.set_grids(new)
        t = self.torch
        for lev in range(1, self.finest + 1):
            L = self.levels[lev]
            self._fill("U", lev - 1)
            scratch = L.A["U2"]
            self._interp_boxes(lev, "U", scratch)
            if lev <= old_finest:
                keep = t.from_numpy(old_valid[lev] & (L.mask_host > 0)).to("cuda")
                I = L.interior(scratch)
                I.copy_(t.where(keep[None], L.interior(L.A["U"]), I))
            L.A["U"], L.A["U2"] = L.A["U2"], L.A["U"]
        self._average("U")
        self._rate = self._validate("U")
        after = self.totals()
        defect = float(np.max(np.abs(after - before) / np.maximum(scale, 1e-300)))
        self.max_remap_defect = max(self.max_remap_defect, defect)
        if defect > 1e-8:
            raise RuntimeError("nonconservative remap")
        self.regrids += 1
        return defect

    def initial_grids(self, Q):
        """Level 0 from Q, then finer levels made from the tags, one level at
        a time (InitFromScratch with every level filled from the coarser)."""
        self.set_grids([[]])
        self.set_state(Q)
        for _ in range(self.max_level):
            tags = [self.tags(lev) for lev in range(self.finest + 1)]
            valid = [L.mask_host > 0 for L in self.levels[:self.finest + 1]]
            new = amr_grid.make_grids(tags, valid, self.max_level, self.blocking, self.max_grid_size, self.grid_eff,
                                      self.n_error_buf, self.n_proper, self.margins)
            self.set_grids(new)
            self.set_state(Q)
        self._rate = None
