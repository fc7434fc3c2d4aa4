"""Box bookkeeping of a block-structured AMR hierarchy (host side).

The reference leaves this to AMReX (regrid with max_level, blocking_factor,
max_grid_size, grid_eff, n_error_buf and n_proper 2: HyperFlow_V3/src/
main.cpp:33-36; ErrorEst's tags, BlockAMR.cpp:170-200; the regrid every
regrid_interval steps, BlockAMR.cpp:420, checked by regridChecked,
:367-373). This module restates the published algorithm AMReX follows
(Berger & Rigoutsos, IEEE Trans. Syst. Man Cybern. 21(5), 1991: signature
cuts at holes, then at the steepest inflection of the signatures' second
difference, else bisection, until each cluster is at least grid_eff full),
with the refinement ratio fixed at 2:

  * tags of level l are grown by n_error_buf cells and coarsened by
    blocking_factor / 2 (one flag per block of the future fine boxes);
  * tags are kept only inside the proper-nesting region of level l (its
    valid cells shrunk by n_proper, rounded to whole blocks; all of the
    domain on level 0) and, when level l+2 exists, the blocks under level
    l+2's boxes grown by n_proper are tagged so that level l+1 keeps
    containing it;
  * the block flags are clustered, each cluster refined to level l+1 cells
    (so every box is a multiple of blocking_factor, aligned to it) and cut
    into pieces of at most max_grid_size cells per side.

Boxes are (ilo, jlo, ihi, jhi), inclusive cell indices of their level.
Which boxes AMReX itself would choose is not reproduced bit for bit (AMReX
is not available); what the boxes must satisfy is pinned by
tests/test_amr_grid.py: every tag covered, boxes disjoint, aligned and no
larger than max_grid_size, every box at least grid_eff full of block flags
or a single block, and proper nesting.
"""
from __future__ import annotations

import numpy as np


def grow_mask(m, n):
    """Cells within n cells (Chebyshev distance) of a set cell: the square
    neighbourhood is separable, a dilation along x then one along y."""
    if n <= 0:
        return m.copy()
    out = m.copy()
    for d in range(1, n + 1):
        out[:, d:] |= m[:, :-d]
        out[:, :-d] |= m[:, d:]
    row = out.copy()
    for d in range(1, n + 1):
        out[d:, :] |= row[:-d, :]
        out[:-d, :] |= row[d:, :]
    return out


def shrink_mask(m, n, keep_edges=True):
    """Cells whose n-neighbourhood lies inside the set; with keep_edges the
    outside of the array counts as inside (domain sides do not shrink the
    proper-nesting region)."""
    if n <= 0:
        return m.copy()
    pad = np.pad(~m, n, constant_values=not keep_edges)
    return ~grow_mask(pad, n)[n:-n, n:-n]


def coarsen_mask(m, r):
    """A flag per r x r block: set when any cell of the block is set (an OR of
    r strided slices along x, then along y)."""
    a = m[:, 0::r].copy()
    for k in range(1, r):
        a |= m[:, k::r]
    b = a[0::r, :].copy()
    for k in range(1, r):
        b |= a[k::r, :]
    return b


def boxes_mask(boxes, shape):
    m = np.zeros(shape, dtype=bool)
    for i0, j0, i1, j1 in boxes:
        m[j0:j1 + 1, i0:i1 + 1] = True
    return m


def _find_cut(sig):
    """(status, cut) for one signature: cut k splits [0, k) | [k, n).
    status 2 a hole (a zero entry; the one nearest the middle), 1 the
    steepest sign change of the second difference, 0 bisection."""
    n = len(sig)
    mid = n / 2.0
    holes = [k for k in range(1, n - 1) if sig[k] == 0]
    if holes:
        k = min(holes, key=lambda h: (abs(h - mid), h))
        return 2, k
    if n >= 4:
        lap = np.zeros(n)
        lap[1:-1] = sig[2:] - 2 * sig[1:-1] + sig[:-2]
        best, cut = 0.0, None
        for k in range(2, n - 1):
            if lap[k - 1] * lap[k] < 0:
                jump = abs(lap[k] - lap[k - 1])
                if jump > best or (jump == best and cut is not None and abs(k - mid) < abs(cut - mid)):
                    best, cut = jump, k
        if cut is not None:
            return 1, cut
    return 0, n // 2


def cluster(flags, grid_eff):
    """Berger-Rigoutsos clustering of a boolean array: boxes (i0, j0, i1, j1)
    covering every set flag, each at least grid_eff full or one cell wide in
    the direction it could not be cut."""
    out = []
    stack = []
    if flags.any():
        stack.append((0, 0, flags.shape[1] - 1, flags.shape[0] - 1))
    while stack:
        i0, j0, i1, j1 = stack.pop()
        sub = flags[j0:j1 + 1, i0:i1 + 1]
        js, is_ = np.nonzero(sub)
        if js.size == 0:
            continue
        i0, i1, j0, j1 = i0 + is_.min(), i0 + is_.max(), j0 + js.min(), j0 + js.max()
        sub = flags[j0:j1 + 1, i0:i1 + 1]
        if sub.mean() >= grid_eff or (i1 == i0 and j1 == j0):
            out.append((int(i0), int(j0), int(i1), int(j1)))
            continue
        cands = []
        for d, sig in ((0, sub.sum(axis=0)), (1, sub.sum(axis=1))):
            if len(sig) < 2:
                continue
            status, k = _find_cut(sig)
            cands.append((status, len(sig), d, k))
        status, _, d, k = max(cands)
        if d == 0:
            stack += [(i0 + k, j0, i1, j1), (i0, j0, i0 + k - 1, j1)]
        else:
            stack += [(i0, j0 + k, i1, j1), (i0, j0, i1, j0 + k - 1)]
    return sorted(out, key=lambda b: (b[1], b[0]))


def _fit_nesting(box, allowed):
    """Pieces of a block box that lie wholly in the allowed blocks (bisection)."""
    i0, j0, i1, j1 = box
    if allowed[j0:j1 + 1, i0:i1 + 1].all():
        return [box]
    if i0 == i1 and j0 == j1:
        return []
    if i1 - i0 >= j1 - j0:
        k = (i0 + i1 + 1) // 2
        halves = [(i0, j0, k - 1, j1), (k, j0, i1, j1)]
    else:
        k = (j0 + j1 + 1) // 2
        halves = [(i0, j0, i1, k - 1), (i0, k, i1, j1)]
    return [p for h in halves for p in _fit_nesting(h, allowed)]


def chop(box, max_size, blocking):
    """Cut a box into near-equal pieces of at most max_size cells per side,
    each a multiple of blocking (AMReX's maxSize with blocking alignment)."""
    i0, j0, i1, j1 = box
    pieces = [box]
    for d in (0, 1):
        nxt = []
        for b in pieces:
            lo, hi = (b[0], b[2]) if d == 0 else (b[1], b[3])
            nb = (hi - lo + 1) // blocking
            parts = -(-(hi - lo + 1) // max_size)
            base, extra = divmod(nb, parts)
            a = lo
            for p in range(parts):
                w = (base + (1 if p < extra else 0)) * blocking
                nxt.append((a, b[1], a + w - 1, b[3]) if d == 0 else (b[0], a, b[2], a + w - 1))
                a += w
        pieces = nxt
    return pieces


def make_grids(tags, valid, max_level, blocking, max_grid_size, grid_eff, n_error_buf=2, n_proper=2,
               margins=(0, 0, 0, 0)):
    """New boxes of levels 1..max_level from the tags of the existing levels.

    tags[l], valid[l]: boolean (ny_l, nx_l) arrays of level l (ny_l = ny_0 2^l)
    for the levels that exist (valid[0] all true); tags only count on valid
    cells. margins: cells per domain side (x-lo, x-hi, y-lo, y-hi, in level-0
    cells) that no fine box may enter (a periodic side needs 1). Returns a
    list per level 0..max_level: [] for level 0's place (the whole domain),
    then the boxes of each fine level ([] when a level is not needed)."""
    bf = blocking // 2
    if blocking % 2 or bf < 2:
        raise ValueError("blocking_factor must be even and at least 4")
    ny0, nx0 = valid[0].shape
    if nx0 % blocking or ny0 % blocking:
        raise ValueError("base cells must be divisible by blocking_factor")
    new = [[] for _ in range(max_level + 1)]
    top = min(len(tags) - 1, max_level - 1)
    for levc in range(top, -1, -1):
        levf = levc + 1
        shape = valid[levc].shape
        t = tags[levc] & valid[levc]
        t = grow_mask(t, n_error_buf) & np.ones(shape, dtype=bool)
        blocks = coarsen_mask(t, bf)
        # the next level up must keep containing the new level levf + 1
        if levf + 1 <= max_level and new[levf + 1]:
            up = boxes_mask(new[levf + 1], (2 * shape[0] * 2, 2 * shape[1] * 2))
            cover = coarsen_mask(up, 2)                   # in levf cells
            cover = grow_mask(cover, n_proper)
            cover = coarsen_mask(cover, 2)                # in levc cells
            blocks |= coarsen_mask(cover, bf)
        # proper nesting in level levc (whole blocks), domain margins
        nest = shrink_mask(valid[levc], n_proper if levc > 0 else 0)
        scale = 2 ** levc
        m = [k * scale for k in margins]
        if any(m):
            inner = np.zeros(shape, dtype=bool)
            inner[m[2]:shape[0] - m[3], m[0]:shape[1] - m[1]] = True
            nest &= inner
        allowed = nest.reshape(shape[0] // bf, bf, shape[1] // bf, bf).all(axis=(1, 3))
        blocks &= allowed
        boxes = []
        for b in cluster(blocks, grid_eff):
            for p in _fit_nesting(b, allowed):
                fi0, fj0 = 2 * bf * p[0], 2 * bf * p[1]
                fi1, fj1 = 2 * bf * (p[2] + 1) - 1, 2 * bf * (p[3] + 1) - 1
                boxes += chop((fi0, fj0, fi1, fj1), max_grid_size, blocking)
        new[levf] = sorted(boxes, key=lambda b: (b[1], b[0]))
        if not new[levf]:
            for k in range(levf, max_level + 1):
                new[k] = []
    # a level without boxes ends the hierarchy
    for lev in range(1, max_level + 1):
        if not new[lev]:
            for k in range(lev, max_level + 1):
                new[k] = []
            break
    return new


def level_masks(boxes, nx0, ny0):
    """Per level l: uint8 (ny_l, nx_l): 0 not valid, 1 leaf, 2 covered by
    level l+1 (the hf_cells mask). boxes[0] is ignored (level 0 is the whole
    domain); the hierarchy ends at the first level without boxes."""
    nlev = 1
    while nlev < len(boxes) and boxes[nlev]:
        nlev += 1
    masks = []
    for lev in range(nlev):
        shape = (ny0 << lev, nx0 << lev)
        valid = np.ones(shape, dtype=bool) if lev == 0 else boxes_mask(boxes[lev], shape)
        m = valid.astype(np.uint8)
        if lev + 1 < nlev:
            fine = boxes_mask(boxes[lev + 1], (shape[0] * 2, shape[1] * 2))
            cov = coarsen_mask(fine, 2)
            if (cov & ~valid).any():
                raise ValueError(f"level {lev + 1} is not nested in level {lev}")
            m[cov] = 2
        masks.append(m)
    return masks


def ghost_strips(box, ghosts=2):
    """The ghost ring of a box as four regions (bottom, top, left, right)."""
    i0, j0, i1, j1 = box
    g = ghosts
    return [(i0 - g, j0 - g, i1 + g, j0 - 1), (i0 - g, j1 + 1, i1 + g, j1 + g),
            (i0 - g, j0, i0 - 1, j1), (i1 + 1, j0, i1 + g, j1)]
