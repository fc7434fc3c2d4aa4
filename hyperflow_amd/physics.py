"""Python mirror of the flux sweep of BlockAMR::euler
(HyperFlow_V3/src/BlockAMR.cpp:202-265) on the HIP kernels.

A stage takes ghosted SoA boxes (two ghost layers, BlockAMR.cpp:206) and
writes the updated valid cells: NND/minmod reconstruction, per-side
first-order fallback, HLLC with Davis wave speeds, wall-face zeroing and the
conservative update. Inadmissible results raise HyperflowError(HF_EDOMAIN),
on which the driver redoes the stage first-order (BlockAMR.cpp:257-263).
"""
from __future__ import annotations

from ctypes import byref, c_int64

from . import _lib
from ._lib import HF_EDOMAIN, check, hf_box, hf_flux_stats, load
from .chemistry import Mechanism, _ptr

WALL_XLO, WALL_XHI, WALL_YLO, WALL_YHI = 1, 2, 4, 8


def box_for_arrays(Uin, Uout, dtdx, dtdy, walls=0, wedge_i=0, in_offset=0, out_offset=0):
    """hf_box for one ghosted input array Uin (nv, ny+4, nx+4) and an output Uout (nv, ny, nx)."""
    nv, gy, gx = Uin.shape
    nx, ny = gx - 4, gy - 4
    if tuple(Uout.shape[1:]) != (ny, nx):
        raise ValueError("output must be (nv, ny, nx)")
    return hf_box(nx, ny, walls, wedge_i, in_offset, gx, gx * gy, out_offset, nx, nx * ny, dtdx, dtdy)


class FluxSweep:
    def __init__(self, mechanism):
        self.mech = mechanism if isinstance(mechanism, Mechanism) else Mechanism(mechanism)
        self.last_stats = None

    def stage(self, boxes, Uin, Uout, first_order=False, fluxes=None):
        """One Euler stage over ``boxes`` (sequence of hf_box) reading Uin, writing Uout.

        fluxes: optional, one (fx, fy) pair of device tensors per box, fx
        (nv, ny, nx+1) and fy (nv, ny+1, nx), that receive the box's face
        fluxes in box-local face indices (the flux-register input)."""
        nb = len(boxes)
        arr = (hf_box * nb)(*boxes)
        st = hf_flux_stats()
        ff = c_int64(-1)
        pin, _ = _ptr(Uin.reshape(-1) if hasattr(Uin, "reshape") else Uin)
        pout, _ = _ptr(Uout.reshape(-1) if hasattr(Uout, "reshape") else Uout)
        if fluxes is None:
            r = load().hf_flux_stage(self.mech.handle, arr, nb, pin, pout, int(bool(first_order)), byref(st),
                                     byref(ff))
        else:
            from .amr import face_fluxes
            if len(fluxes) != nb:
                raise ValueError("one (fx, fy) pair per box")
            desc = (_lib.hf_face_fluxes * nb)(*[face_fluxes(fx, fy) for fx, fy in fluxes])
            r = load().hf_flux_stage_fluxes(self.mech.handle, arr, nb, pin, pout, int(bool(first_order)), desc,
                                            byref(st), byref(ff))
        self.last_stats = st.as_dict()
        check(self.mech.handle, r, ff.value)
        return self.last_stats

    def stage_box(self, Uin, Uout, dtdx, dtdy, walls=0, first_order=False):
        return self.stage([box_for_arrays(Uin, Uout, dtdx, dtdy, walls)], Uin, Uout, first_order)
