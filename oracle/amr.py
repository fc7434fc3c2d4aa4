"""CPU restatement of the coarse/fine transfers between block-AMR levels —
TEST INFRASTRUCTURE ONLY (the checker of hyperflow_amd.amr).

  interp        ConservativeInterp::interp (HyperFlow_V3/src/ConservativeInterp.cpp:18-51):
                minmod slopes of every coarse parent (:30-31), one theta for
                all components and all four children, halved up to 48 times
                until every child is admissible (:32-42), 0 with the parent
                itself checked otherwise (:43), children q + theta (+-sx/4
                +- sy/4) written where the fine region holds them (:44-49).
  average_down  BlockAMR::average (HyperFlow_V3/src/BlockAMR.cpp:157-159) over
                AMReX average_down (amrex @2d8c2e7e, fetched by
                HyperFlow_V3/CMakeLists.txt:22-26, not vendored): for equal
                cell volumes, coarse = (1/4) * sum of the 2 x 2 children,
                accumulated from 0 with i fastest.
  reflux        BlockAMR::euler's flux registers (HyperFlow_V3/src/BlockAMR.cpp:247-256)
                for one fine box at ratio 2 over AMReX FluxRegister (same
                pin): CrseInit stores (-h dy_c) Fc on the coarse faces of the
                box outline, FineAdd adds (h dy_f) Ff of both fine faces,
                Reflux adds -reg / V_c to the coarse cell outside a lo side
                and +reg / V_c outside a hi side.

  tag_cells     BlockAMR::ErrorEst (HyperFlow_V3/src/BlockAMR.cpp:170-200): the moving
                tag box, the relative p and T jumps to the face neighbours, the
                fuel consumption rate.

Admissibility is the oracle's properties() (physics.cpp:4-15, 46-48).

Parity status: interp is pinned by its invariants (a uniform field is
reproduced, a linear field is reproduced exactly where minmod keeps the
slope, every parent's four children average back to it), by a closed form of
the theta search on species-gradient fields (theta = the largest power of two
keeping every child's partial densities non-negative), and by the known
failure (an inadmissible parent is an error). average_down is pinned by
exact integer-valued sums and by conservation (its summation order follows
AMReX's loop; AMReX is not vendored, so against AMReX's own rounding it and
reflux are parity unpinned). reflux is pinned by
composite conservation: one Euler stage on a periodic coarse level and an
interior fine box (same h), reflux, average-down keeps every conserved total
to roundoff, which it does not without the correction, and a flipped sign or
a wrong face scale breaks it (tests/test_oracle_amr.py). tag_cells is
pinned by closed forms: a uniform gas tags nothing; a 10 % pressure step tags
exactly the two columns beside it and a 5 % one none; a temperature step at
fixed pressure likewise; the box criterion tags exactly the cells whose moved
centres lie in the box; the fuel-rate criterion at a threshold between two
known rates (the oracle's kinetics, pinned on their own) splits the cells.
"""
from __future__ import annotations

import numpy as np


class DomainError(Exception):
    pass


def minmod(a, b):
    """physics.cpp:40-44, elementwise."""
    return np.where(a * b > 0, np.copysign(np.minimum(np.abs(a), np.abs(b)), a), 0.0)


def _admissible(o, Q):
    """Q (nv, m) -> bool (m,): the reference's admissible() per column."""
    _, _, _, st = o.properties(np.ascontiguousarray(Q.T))
    return st == 0


def _child(q, theta, cx, sx, cy, sy):
    """ConservativeInterp.cpp:37,47: q + theta ((+-0.25) sx + (+-0.25) sy), no fusion."""
    return q + theta * (cx * sx + cy * sy)


_CHILDREN = ((0, 0), (1, 0), (0, 1), (1, 1))     # (a, b): fine (2i + a, 2j + b)


def interp(o, C, clo, region):
    """Interpolate coarse C onto the fine cells of region.

    C: (nv, cny, cnx); C[:, j - clo[1], i - clo[0]] is coarse cell (i, j).
    region: (ilo, jlo, ihi, jhi), fine cells, inclusive.
    Returns (F, theta, failed): F (nv, jhi-jlo+1, ihi-ilo+1) the fine values
    (NaN where the parent failed), theta (cny', cnx') per parent of
    coarsen(region), failed the row-major indices of inadmissible parents.
    """
    ilo, jlo, ihi, jhi = region
    ci0, cj0, ci1, cj1 = ilo >> 1, jlo >> 1, ihi >> 1, jhi >> 1
    nv = C.shape[0]

    def at(di, dj):
        return C[:, cj0 - clo[1] + dj:cj1 - clo[1] + 1 + dj, ci0 - clo[0] + di:ci1 - clo[0] + 1 + di]

    if ci0 - 1 < clo[0] or cj0 - 1 < clo[1] or ci1 + 1 >= clo[0] + C.shape[2] or cj1 + 1 >= clo[1] + C.shape[1]:
        raise ValueError("coarse array must hold grow(coarsen(region), 1)")
    shape = (cj1 - cj0 + 1, ci1 - ci0 + 1)
    q = at(0, 0).reshape(nv, -1)
    sx = minmod(q - at(-1, 0).reshape(nv, -1), at(1, 0).reshape(nv, -1) - q)
    sy = minmod(q - at(0, -1).reshape(nv, -1), at(0, 1).reshape(nv, -1) - q)
    m = q.shape[1]
    theta = np.ones(m)
    accepted = np.zeros(m, dtype=bool)
    pending = np.arange(m)
    for _ in range(48):
        if pending.size == 0:
            break
        good = np.ones(pending.size, dtype=bool)
        for a, b in _CHILDREN:
            ch = _child(q[:, pending], theta[pending], 0.25 if a else -0.25, sx[:, pending],
                        0.25 if b else -0.25, sy[:, pending])
            good &= _admissible(o, ch)
        accepted[pending[good]] = True
        theta[pending[~good]] *= 0.5
        pending = pending[~good]
    theta[~accepted] = 0.0
    failed = np.zeros(m, dtype=bool)
    if (~accepted).any():
        idx = np.flatnonzero(~accepted)
        failed[idx] = ~_admissible(o, q[:, idx])
    F = np.full((nv, jhi - jlo + 1, ihi - ilo + 1), np.nan)
    for a, b in _CHILDREN:
        ch = _child(q, theta, 0.25 if a else -0.25, sx, 0.25 if b else -0.25, sy).reshape(nv, *shape)
        ch[:, failed.reshape(shape)] = np.nan
        fi = 2 * np.arange(ci0, ci1 + 1) + a
        fj = 2 * np.arange(cj0, cj1 + 1) + b
        ki = (fi >= ilo) & (fi <= ihi)
        kj = (fj >= jlo) & (fj <= jhi)
        F[:, (fj[kj] - jlo)[:, None], (fi[ki] - ilo)[None, :]] = ch[:, kj][:, :, ki]
    return F, theta.reshape(shape), np.flatnonzero(failed)


def average_down(F, flo, cregion):
    """Coarse cells of cregion (ilo, jlo, ihi, jhi) from the fine array F
    (F[:, j - flo[1], i - flo[0]] is fine cell (i, j)): (nv, cny, cnx)."""
    ilo, jlo, ihi, jhi = cregion
    fi0, fj0 = 2 * ilo - flo[0], 2 * jlo - flo[1]
    fi1, fj1 = 2 * ihi + 1 - flo[0], 2 * jhi + 1 - flo[1]
    if fi0 < 0 or fj0 < 0 or fi1 >= F.shape[2] or fj1 >= F.shape[1]:
        raise ValueError("fine array must hold refine(cregion)")
    s = np.zeros((F.shape[0], jhi - jlo + 1, ihi - ilo + 1))
    for b in (0, 1):
        for a in (0, 1):
            s = s + F[:, fj0 + b:fj1 + 1:2, fi0 + a:fi1 + 1:2]
    return 0.25 * s


def reflux(Uc, clo, coarse, fine, fine_box, sides, h, dxc, dyc):
    """Apply the flux-register correction of one fine box to Uc in place.

    Uc: coarse level (nv, h, w), Uc[:, j - clo[1], i - clo[0]] is cell (i, j).
    coarse = (Fx, Fy, lo): the coarse stage's face fluxes, Fx (nv, ny, nx+1) and
    Fy (nv, ny+1, nx), whose element [:, 0, 0] is face lo of the coarse level;
    fine = (Fx, Fy, lo) likewise for the fine box's stage in the fine index
    space. fine_box = (ilo, jlo, ihi, jhi) fine cells; sides bit0 x-lo, bit1
    x-hi, bit2 y-lo, bit3 y-hi.
    """
    FI0, FJ0, FI1, FJ1 = fine_box
    ci0, cj0, ci1, cj1 = FI0 >> 1, FJ0 >> 1, FI1 >> 1, FJ1 >> 1
    cfx, cfy, cl = coarse
    ffx, ffy, fl = fine
    vol = dxc * dyc
    for side in range(4):
        if not (sides >> side) & 1:
            continue
        sgn = 1.0 if side & 1 else -1.0
        if side < 2:
            I = ci0 if side == 0 else ci1 + 1
            FI = FI0 if side == 0 else FI1 + 1
            oi = ci0 - 1 if side == 0 else ci1 + 1
            for j in range(cj0, cj1 + 1):
                reg = (-h * dyc) * cfx[:, j - cl[1], I - cl[0]]
                reg = reg + (h * (0.5 * dyc)) * ffx[:, 2 * j - fl[1], FI - fl[0]]
                reg = reg + (h * (0.5 * dyc)) * ffx[:, 2 * j + 1 - fl[1], FI - fl[0]]
                Uc[:, j - clo[1], oi - clo[0]] += (sgn * reg) / vol
        else:
            J = cj0 if side == 2 else cj1 + 1
            FJ = FJ0 if side == 2 else FJ1 + 1
            oj = cj0 - 1 if side == 2 else cj1 + 1
            for i in range(ci0, ci1 + 1):
                reg = (-h * dxc) * cfy[:, J - cl[1], i - cl[0]]
                reg = reg + (h * (0.5 * dxc)) * ffy[:, FJ - fl[1], 2 * i - fl[0]]
                reg = reg + (h * (0.5 * dxc)) * ffy[:, FJ - fl[1], 2 * i + 1 - fl[0]]
                Uc[:, oj - clo[1], i - clo[0]] += (sgn * reg) / vol


def tag_cells(o, G, dx, dy, x0=0.0, y0=0.0, time=0.0, pressure_tag=0.08, temperature_tag=0.08, fuel_rate_tag=50.0,
              physics=True, chemistry=True, fuel=0, box=None, velocity=(0.0, 0.0)):
    """Tags (ny, nx) of the valid cells of the ghosted box G (nv, ny+4, nx+4),
    first ghost layer filled. Raises DomainError on an inadmissible cell or
    neighbour (properties throws, BlockAMR.cpp:183-185)."""
    nv, gy, gx = G.shape
    nx, ny = gx - 4, gy - 4
    ns = nv - 3
    i = np.arange(nx)[None, :]
    j = np.arange(ny)[:, None]
    x = x0 + (i + 0.5) * dx - velocity[0] * time
    y = y0 + (j + 0.5) * dy - velocity[1] * time
    marked = np.zeros((ny, nx), dtype=bool)
    if box is not None:
        marked |= (x >= box[0]) & (x <= box[1]) & (y >= box[2]) & (y <= box[3])
    if not physics:
        return marked
    R = G[:, 1:ny + 3, 1:nx + 3]                      # valid cells and the first ghost ring
    p, _, T, st = o.properties(np.ascontiguousarray(R.reshape(nv, -1).T))
    p, T, st = p.reshape(ny + 2, nx + 2), T.reshape(ny + 2, nx + 2), st.reshape(ny + 2, nx + 2)
    bad = st != 0
    inner = (slice(1, ny + 1), slice(1, nx + 1))
    failed = bad[inner] | bad[0:ny, 1:nx + 1] | bad[2:ny + 2, 1:nx + 1] | bad[1:ny + 1, 0:nx] | bad[1:ny + 1, 2:nx + 2]
    if failed.any():
        raise DomainError(f"inadmissible state while tagging at cell {int(np.flatnonzero(failed.ravel())[0])}")
    pc, Tc = p[inner], T[inner]
    for (sl_l, sl_r) in (((slice(1, ny + 1), slice(0, nx)), (slice(1, ny + 1), slice(2, nx + 2))),
                         ((slice(0, ny), slice(1, nx + 1)), (slice(2, ny + 2), slice(1, nx + 1)))):
        pl, pr, Tl, Tr = p[sl_l], p[sl_r], T[sl_l], T[sl_r]
        jump = np.maximum(np.abs(pl - pc) / np.minimum(pl, pc), np.abs(pr - pc) / np.minimum(pr, pc))
        temp = np.maximum(np.abs(Tl - Tc) / np.minimum(Tl, Tc), np.abs(Tr - Tc) / np.minimum(Tr, Tc))
        marked |= (jump > pressure_tag) | (temp > temperature_tag)
    if chemistry:
        W = o.m["W"]
        Q = G[:, 2:ny + 2, 2:nx + 2]
        for jj in range(ny):
            for ii in range(nx):
                C = Q[:ns, jj, ii] / W
                w = o.wdot(Tc[jj, ii], C)
                marked[jj, ii] |= -w[fuel] * W[fuel] / Q[:ns, jj, ii].sum() > fuel_rate_tag
    return marked
