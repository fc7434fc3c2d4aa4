"""Python mirror of the level transfers around the flux sweep, on the HIP
kernels: the physical ghost fill with FlowBoundary's inflow and wedge bottom
(HyperFlow_V3/src/FlowBoundary.hpp:7-37, applied by BlockAMR::fill,
BlockAMR.cpp:137-152), coarse-to-fine conservative interpolation
(ConservativeInterp::interp, ConservativeInterp.cpp:18-51) and average-down
(BlockAMR::average, BlockAMR.cpp:157-159), and the flux-register correction
across levels (BlockAMR.cpp:247-256).

Arrays are torch float64 device tensors in the reference's SoA layout, one
(nv, h, w) tensor per patch, i fastest. A patch tensor carries the index of
its element [:, 0, 0] as ``lo = (i0, j0)`` in its level's index space. Every
computation runs in libhyperflow_hip.so; there is no CPU path.
"""
from __future__ import annotations

from ctypes import byref, c_double, c_int32, c_int64

from ._lib import (HF_BC_INFLOW, HF_BC_WEDGE, check, hf_box, hf_face_fluxes, hf_patch, hf_tag_params,  # noqa: F401
                   load)
from .chemistry import Mechanism


def _mech(mechanism):
    return mechanism if isinstance(mechanism, Mechanism) else Mechanism(mechanism)


def _device_f64(t, name):
    if not (getattr(t, "is_cuda", False) and str(t.dtype) == "torch.float64" and t.is_contiguous() and t.dim() == 3):
        raise ValueError(f"{name} must be a contiguous float64 (nv, h, w) device tensor")


def _patch(t, lo):
    """hf_patch addressing t[n, j - lo[1], i - lo[0]] as cell (i, j)."""
    h, w = t.shape[1], t.shape[2]
    return hf_patch(-(int(lo[1]) * w + int(lo[0])), w, h * w)


def wedge_column(prob_lo_x, dx, wedge_start):
    """Index of the first column whose centre lies at or past the wedge tip
    (x = lo + (i + 0.5) dx >= wedgeStart, BlockAMR.cpp:225-226; FlowBoundary.hpp:27-28).
    The reference requires the tip on a base-grid x face (BlockAMR.cpp:41-47)."""
    index = (wedge_start - prob_lo_x) / dx
    if abs(index - round(index)) > 1e-10:
        raise ValueError("wedge tip must align with a base-grid x face")
    return int(round(index))


def fill_ghosts(mechanism, U, bc, inlet=None, wedge_i=0):
    """Fill the two ghost layers of the ghosted box U (nv, ny+4, nx+4) in place.

    bc = (x_lo, x_hi, y_lo, y_hi) from HF_BC_TRANSMISSIVE, HF_BC_REFLECTING,
    HF_BC_PERIODIC, HF_BC_INFLOW (needs inlet: nv conserved values) and
    HF_BC_WEDGE (y_lo only; wall from column wedge_i on).
    """
    m = _mech(mechanism)
    _device_f64(U, "U")
    nv, gy, gx = U.shape
    if nv != m.nv:
        raise ValueError("U must hold ns+3 components")
    nx, ny = gx - 4, gy - 4
    box = hf_box(nx, ny, 0, int(wedge_i), 0, gx, gx * gy, 0, nx, nx * ny, 0.0, 0.0)
    codes = (c_int32 * 4)(*[int(b) for b in bc])
    q = None
    if inlet is not None:
        vals = [float(v) for v in inlet]
        if len(vals) != nv:
            raise ValueError("inlet must hold ns+3 conserved values")
        q = (c_double * nv)(*vals)
    check(m.handle, load().hf_fill_ghosts_flow(m.handle, U.data_ptr(), byref(box), codes, q))


class ConservativeInterp:
    """Coarse-to-fine interpolation at ratio (2, 2) (ConservativeInterp.hpp:56-68)."""

    def __init__(self, mechanism):
        self.mech = _mech(mechanism)
        self.last_limited = 0

    def interp(self, coarse, coarse_lo, fine, fine_lo, region):
        """Write the fine cells region = (ilo, jlo, ihi, jhi) (inclusive) of
        ``fine`` from ``coarse``, which must hold grow(coarsen(region), 1).
        Returns the number of parents whose slopes were limited (theta < 1);
        an inadmissible parent raises HyperflowError(HF_EDOMAIN)."""
        _device_f64(coarse, "coarse")
        _device_f64(fine, "fine")
        ilo, jlo, ihi, jhi = (int(v) for v in region)
        if ihi < ilo or jhi < jlo:
            raise ValueError("empty region")
        ci0, cj0, ci1, cj1 = ilo >> 1, jlo >> 1, ihi >> 1, jhi >> 1
        cx0, cy0 = coarse_lo
        fx0, fy0 = fine_lo
        if (ci0 - 1 < cx0 or cj0 - 1 < cy0 or ci1 + 1 >= cx0 + coarse.shape[2] or cj1 + 1 >= cy0 + coarse.shape[1]):
            raise ValueError("coarse patch must hold grow(coarsen(region), 1)")
        if ilo < fx0 or jlo < fy0 or ihi >= fx0 + fine.shape[2] or jhi >= fy0 + fine.shape[1]:
            raise ValueError("fine patch must hold region")
        if coarse.shape[0] != self.mech.nv or fine.shape[0] != self.mech.nv:
            raise ValueError("patches must hold ns+3 components")
        cp, fp = _patch(coarse, coarse_lo), _patch(fine, fine_lo)
        reg = (c_int32 * 4)(ilo, jlo, ihi, jhi)
        limited, ff = c_int64(0), c_int64(-1)
        r = load().hf_interp_coarse_to_fine(self.mech.handle, coarse.data_ptr(), byref(cp), fine.data_ptr(),
                                            byref(fp), reg, byref(limited), byref(ff))
        self.last_limited = limited.value
        check(self.mech.handle, r, ff.value)
        return limited.value


def average_down(mechanism, fine, fine_lo, coarse, coarse_lo, cregion):
    """Replace the coarse cells cregion = (ilo, jlo, ihi, jhi) (inclusive) of
    ``coarse`` by the mean of their 2 x 2 children in ``fine``."""
    m = _mech(mechanism)
    _device_f64(coarse, "coarse")
    _device_f64(fine, "fine")
    ilo, jlo, ihi, jhi = (int(v) for v in cregion)
    if ihi < ilo or jhi < jlo:
        raise ValueError("empty region")
    cx0, cy0 = coarse_lo
    fx0, fy0 = fine_lo
    if ilo < cx0 or jlo < cy0 or ihi >= cx0 + coarse.shape[2] or jhi >= cy0 + coarse.shape[1]:
        raise ValueError("coarse patch must hold cregion")
    if 2 * ilo < fx0 or 2 * jlo < fy0 or 2 * ihi + 1 >= fx0 + fine.shape[2] or 2 * jhi + 1 >= fy0 + fine.shape[1]:
        raise ValueError("fine patch must hold refine(cregion)")
    if coarse.shape[0] != m.nv or fine.shape[0] != m.nv:
        raise ValueError("patches must hold ns+3 components")
    fp, cp = _patch(fine, fine_lo), _patch(coarse, coarse_lo)
    reg = (c_int32 * 4)(ilo, jlo, ihi, jhi)
    check(m.handle, load().hf_average_down(m.handle, fine.data_ptr(), byref(fp), coarse.data_ptr(), byref(cp), reg))


def face_fluxes(fx, fy, lo=(0, 0)):
    """hf_face_fluxes for device tensors fx (nv, ny, nx+1) and fy (nv, ny+1, nx)
    whose element [:, 0, 0] is face lo of their index space."""
    for t, name in ((fx, "fx"), (fy, "fy")):
        _device_f64(t, name)
    return hf_face_fluxes(fx.data_ptr(), -(int(lo[1]) * fx.shape[2] + int(lo[0])), fx.shape[2],
                          fx.shape[1] * fx.shape[2],
                          fy.data_ptr(), -(int(lo[1]) * fy.shape[2] + int(lo[0])), fy.shape[2],
                          fy.shape[1] * fy.shape[2])


def reflux(mechanism, Uc, coarse_lo, coarse, fine, fine_box, sides, h, dxc, dyc):
    """Flux-register correction of one fine box (ratio 2) on the coarse level Uc.

    coarse = (fx, fy, lo) and fine = (fx, fy, lo): the face fluxes of the
    coarse stage and of the fine box's stage (FluxSweep.stage(..., fluxes=)),
    lo the level index of their face [:, 0, 0]. fine_box = (ilo, jlo, ihi, jhi)
    fine cells; sides: bit0 x-lo, bit1 x-hi, bit2 y-lo, bit3 y-hi (the sides
    that face coarse cells)."""
    m = _mech(mechanism)
    _device_f64(Uc, "Uc")
    ilo, jlo, ihi, jhi = (int(v) for v in fine_box)
    ci0, cj0, ci1, cj1 = ilo >> 1, jlo >> 1, ihi >> 1, jhi >> 1
    cx0, cy0 = coarse_lo
    if ci0 - 1 < cx0 or cj0 - 1 < cy0 or ci1 + 1 >= cx0 + Uc.shape[2] or cj1 + 1 >= cy0 + Uc.shape[1]:
        raise ValueError("coarse patch must hold the cells around the fine box")
    for (fx, fy, lo), (a0, b0, a1, b1) in ((coarse, (ci0, cj0, ci1 + 1, cj1 + 1)), (fine, (ilo, jlo, ihi + 1, jhi + 1))):
        if a0 < lo[0] or b0 < lo[1] or a1 > lo[0] + fx.shape[2] - 1 or b1 - 1 >= lo[1] + fx.shape[1] or \
                b1 > lo[1] + fy.shape[1] - 1 or a1 - 1 >= lo[0] + fy.shape[2]:
            raise ValueError("face fluxes must cover the fine box outline")
    cf, ff = face_fluxes(*coarse), face_fluxes(*fine)
    cp = _patch(Uc, coarse_lo)
    reg = (c_int32 * 4)(ilo, jlo, ihi, jhi)
    check(m.handle, load().hf_reflux(m.handle, Uc.data_ptr(), byref(cp), byref(cf), byref(ff), reg, int(sides),
                                     float(h), float(dxc), float(dyc)))


def tag_cells(mechanism, U, dx, dy, x0=0.0, y0=0.0, time=0.0, pressure_tag=0.08, temperature_tag=0.08,
              fuel_rate_tag=50.0, physics=True, chemistry=True, fuel=0, box=None, velocity=(0.0, 0.0)):
    """Refinement tags of BlockAMR::ErrorEst (BlockAMR.cpp:170-200) for the
    ghosted box U (nv, ny+4, nx+4) whose first ghost layer is filled: a uint8
    device tensor (ny, nx), 1 where tagged. Defaults are the reference's
    (pressure_tag and temperature_tag 0.08, fuel_rate_tag 50 /s, species 0 the
    fuel). An inadmissible cell or neighbour raises HyperflowError(HF_EDOMAIN)."""
    import torch
    m = _mech(mechanism)
    _device_f64(U, "U")
    nv, gy, gx = U.shape
    nx, ny = gx - 4, gy - 4
    b = hf_box(nx, ny, 0, 0, 0, gx, gx * gy, 0, nx, nx * ny, 0.0, 0.0)
    P = hf_tag_params(float(pressure_tag), float(temperature_tag), float(fuel_rate_tag), int(bool(physics)),
                      int(bool(chemistry)), int(fuel), int(box is not None),
                      (c_double * 4)(*(box if box is not None else (0.0, 0.0, 0.0, 0.0))),
                      (c_double * 2)(*velocity), float(x0), float(y0), float(dx), float(dy), float(time))
    tags = torch.zeros((ny, nx), dtype=torch.uint8, device=U.device)
    n, ff = c_int64(0), c_int64(-1)
    r = load().hf_tag_cells(m.handle, U.data_ptr(), byref(b), byref(P), tags.data_ptr(), byref(n), byref(ff))
    check(m.handle, r, ff.value)
    return tags


def choose_fine_box(tags, buffer, blocking, nx, ny, margin=2):
    """One refined box for the tags of a coarse level (ny, nx bool): the
    tags grown by ``buffer`` cells (AMReX's n_error_buf, HyperFlow_V3/src/
    main.cpp:35), their bounding box widened to multiples of ``blocking``
    coarse cells (blocking_factor) and kept ``margin`` cells inside the
    level (one number, or one per side: x-lo, x-hi, y-lo, y-hi). Returns fine
    cells (ilo, jlo, ihi, jhi) at ratio 2, or None when nothing is tagged.
    AMReX clusters the tags into several boxes; a single bounding box is the
    one-box driver's form of it."""
    import numpy as np
    t = np.asarray(tags, dtype=bool)
    if not t.any():
        return None
    js, is_ = np.nonzero(t)
    i0, i1 = int(is_.min()) - buffer, int(is_.max()) + buffer
    j0, j1 = int(js.min()) - buffer, int(js.max()) + buffer
    i0, j0 = (i0 // blocking) * blocking, (j0 // blocking) * blocking
    i1, j1 = (i1 // blocking + 1) * blocking - 1, (j1 // blocking + 1) * blocking - 1
    m = (margin,) * 4 if np.isscalar(margin) else tuple(margin)
    i0, j0 = max(i0, m[0]), max(j0, m[2])
    i1, j1 = min(i1, nx - 1 - m[1]), min(j1, ny - 1 - m[3])
    if i1 < i0 or j1 < j0:
        return None
    return (2 * i0, 2 * j0, 2 * i1 + 1, 2 * j1 + 1)
