"""ctypes binding of libhyperflow_hip.so (the C ABI in include/hyperflow.h).

The library is loaded from this package directory only. There is no fallback:
if the HIP library is missing or cannot be loaded, every call raises.
"""
from __future__ import annotations

import ctypes
from ctypes import (POINTER, Structure, byref, c_char_p, c_double, c_int, c_int32, c_int64, c_uint64,
                    c_void_p)
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libhyperflow_hip.so"

HF_OK, HF_EINVAL, HF_EDOMAIN, HF_ECONV, HF_ENOMEM, HF_EDEVICE, HF_EBUSY = range(7)
HF_CHEM_PLAIN, HF_CHEM_BATCHED = 0, 1
HF_BC_TRANSMISSIVE, HF_BC_REFLECTING, HF_BC_PERIODIC, HF_BC_INFLOW, HF_BC_WEDGE, HF_BC_KEEP = 0, 1, 2, 3, 4, 5
HF_PREC_FP64, HF_PREC_MIXED = 0, 1

ERROR_NAMES = {0: "HF_OK", 1: "HF_EINVAL", 2: "HF_EDOMAIN", 3: "HF_ECONV", 4: "HF_ENOMEM", 5: "HF_EDEVICE",
               6: "HF_EBUSY"}


class hf_options(Structure):
    _fields_ = [("rtol", c_double), ("atol", c_double), ("max_steps", c_int), ("precision", c_int),
                ("device", c_int), ("waves_per_cu", c_int), ("chunk", c_int), ("nbins", c_int), ("eager", c_int),
                ("jac_wait", c_int), ("setup_wait", c_int)]


class hf_chem_stats(Structure):
    _fields_ = [("cells", c_int64), ("steps", c_int64), ("rhs", c_int64), ("jac", c_int64), ("lu", c_int64),
                ("newton", c_int64), ("conv_failures", c_int64), ("err_failures", c_int64),
                ("max_cell_steps", c_int64), ("team", c_int), ("kernel_ms", c_double), ("chunk", c_int),
                ("graph", c_int), ("waves_per_cu", c_int), ("nbins", c_int), ("jac_wait", c_int),
                ("setup_wait", c_int), ("tuned", c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class hf_box(Structure):
    _fields_ = [("nx", c_int32), ("ny", c_int32), ("walls", c_int32), ("wedge_i", c_int32),
                ("in_offset", c_int64), ("in_pitch", c_int64), ("in_plane", c_int64),
                ("out_offset", c_int64), ("out_pitch", c_int64), ("out_plane", c_int64),
                ("dtdx", c_double), ("dtdy", c_double)]


class hf_patch(Structure):
    _fields_ = [("offset", c_int64), ("pitch", c_int64), ("plane", c_int64)]


class hf_face_fluxes(Structure):
    _fields_ = [("fx", c_void_p), ("fx_offset", c_int64), ("fx_pitch", c_int64), ("fx_plane", c_int64),
                ("fy", c_void_p), ("fy_offset", c_int64), ("fy_pitch", c_int64), ("fy_plane", c_int64)]


class hf_cells(Structure):
    _fields_ = [("U", c_void_p), ("offset", c_int64), ("pitch", c_int64), ("plane", c_int64), ("nx", c_int32),
                ("ny", c_int32), ("mask", c_void_p), ("mask_pitch", c_int64), ("dx", c_double), ("dy", c_double)]


class hf_level(Structure):
    _fields_ = [("U", c_void_p), ("U1", c_void_p), ("U2", c_void_p), ("old", c_void_p), ("backup", c_void_p),
                ("fx", c_void_p), ("fy", c_void_p), ("mask", c_void_p), ("nx", c_int32), ("ny", c_int32),
                ("dx", c_double), ("dy", c_double), ("boxes", POINTER(c_int32)), ("nbox", c_int32),
                ("ghosts", POINTER(c_int32)), ("nghost", c_int32), ("wedge_i", c_int32)]


class hf_tag_params(Structure):
    _fields_ = [("pressure_tag", c_double), ("temperature_tag", c_double), ("fuel_rate_tag", c_double),
                ("physics", c_int32), ("chemistry", c_int32), ("fuel", c_int32), ("has_box", c_int32),
                ("box", c_double * 4), ("velocity", c_double * 2), ("x0", c_double), ("y0", c_double),
                ("dx", c_double), ("dy", c_double), ("time", c_double)]


class hf_flux_stats(Structure):
    _fields_ = [("cells", c_int64), ("faces", c_int64), ("first_order_sides", c_int64), ("kernel_ms", c_double),
                ("tile_x", c_int), ("tile_y", c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_dp = POINTER(c_double)
_SIGNATURES = {
    "hf_create": (c_int, [c_char_p, POINTER(hf_options), POINTER(c_void_p)]),
    "hf_destroy": (None, [c_void_p]),
    "hf_last_error": (c_char_p, [c_void_p]),
    "hf_create_error": (c_char_p, []),
    "hf_num_species": (c_int, [c_void_p]),
    "hf_num_reactions": (c_int, [c_void_p]),
    "hf_species_name": (c_char_p, [c_void_p, c_int]),
    "hf_molecular_weights": (c_int, [c_void_p, _dp]),
    "hf_set_options": (c_int, [c_void_p, POINTER(hf_options)]),
    "hf_get_options": (c_int, [c_void_p, POINTER(hf_options)]),
    "hf_chem_advance": (c_int, [c_void_p, c_void_p, c_int64, c_int64, c_void_p, c_int, c_int,
                                POINTER(hf_chem_stats), POINTER(c_int64)]),
    "hf_flux_stage": (c_int, [c_void_p, POINTER(hf_box), c_int, c_void_p, c_void_p, c_int,
                              POINTER(hf_flux_stats), POINTER(c_int64)]),
    "hf_flux_stage_fluxes": (c_int, [c_void_p, POINTER(hf_box), c_int, c_void_p, c_void_p, c_int,
                                     POINTER(hf_face_fluxes), POINTER(hf_flux_stats), POINTER(c_int64)]),
    "hf_chem_advance_cells": (c_int, [c_void_p, POINTER(hf_cells), c_int, c_void_p, c_int, POINTER(hf_chem_stats),
                                      POINTER(c_int64)]),
    "hf_validate_cells": (c_int, [c_void_p, POINTER(hf_cells), c_int, _dp, POINTER(c_int64)]),
    "hf_totals_cells": (c_int, [c_void_p, POINTER(hf_cells), c_int, c_int, _dp]),
    "hf_repair_cells": (c_int, [c_void_p, POINTER(hf_cells), c_int, POINTER(c_int64), POINTER(c_int64)]),
    "hf_interp_regions": (c_int, [c_void_p, c_void_p, POINTER(hf_patch), c_void_p, POINTER(hf_patch),
                                  POINTER(c_int32), c_int, c_void_p, c_int64, POINTER(c_int32), POINTER(c_int64),
                                  POINTER(c_int64)]),
    "hf_average_down_masked": (c_int, [c_void_p, c_void_p, POINTER(hf_patch), c_void_p, POINTER(hf_patch), c_void_p,
                                       c_int64, POINTER(c_int32)]),
    "hf_reflux_masked": (c_int, [c_void_p, c_void_p, POINTER(hf_patch), POINTER(hf_face_fluxes),
                                 POINTER(hf_face_fluxes), c_void_p, c_int64, POINTER(c_int32), c_double, c_double,
                                 c_double]),
    "hf_flux_stage_boxes": (c_int, [c_void_p, POINTER(hf_box), c_int, POINTER(c_void_p), POINTER(c_void_p), c_int,
                                    POINTER(hf_face_fluxes), POINTER(hf_flux_stats), POINTER(c_int64)]),
    "hf_flux_stage_noalloc": (c_int, [c_void_p, POINTER(hf_box), c_int, c_void_p, c_void_p, c_int,
                                      POINTER(hf_face_fluxes), POINTER(hf_flux_stats), POINTER(c_int64)]),
    "hf_reflux": (c_int, [c_void_p, c_void_p, POINTER(hf_patch), POINTER(hf_face_fluxes), POINTER(hf_face_fluxes),
                          POINTER(c_int32), c_int32, c_double, c_double, c_double]),
    "hf_properties": (c_int, [c_void_p, c_void_p, c_int64, c_int64, c_void_p, c_void_p, c_void_p,
                              POINTER(c_int64)]),
    "hf_stable_dt": (c_int, [c_void_p, c_void_p, c_int64, c_int64, c_double, c_double, c_double, _dp]),
    "hf_chem_adjoint": (c_int, [c_void_p, c_int64, c_void_p, c_void_p, c_void_p, c_double, c_int, c_int,
                                c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "hf_chem_tangent": (c_int, [c_void_p, c_int64, c_void_p, c_void_p, c_void_p, c_double, c_int, c_void_p,
                                c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "hf_fill_ghosts": (c_int, [c_void_p, c_void_p, POINTER(hf_box), POINTER(c_int32)]),
    "hf_fill_ghosts_flow": (c_int, [c_void_p, c_void_p, POINTER(hf_box), POINTER(c_int32), _dp]),
    "hf_interp_coarse_to_fine": (c_int, [c_void_p, c_void_p, POINTER(hf_patch), c_void_p, POINTER(hf_patch),
                                         POINTER(c_int32), POINTER(c_int64), POINTER(c_int64)]),
    "hf_average_down": (c_int, [c_void_p, c_void_p, POINTER(hf_patch), c_void_p, POINTER(hf_patch),
                                POINTER(c_int32)]),
    "hf_tag_cells": (c_int, [c_void_p, c_void_p, POINTER(hf_box), POINTER(hf_tag_params), c_void_p,
                             POINTER(c_int64), POINTER(c_int64)]),
    "hf_euler_stage": (c_int, [c_void_p, c_void_p, c_void_p, POINTER(hf_box), POINTER(c_int32), _dp, c_double,
                               c_double, POINTER(c_int64), POINTER(c_int32), POINTER(c_int64)]),
    "hf_box_strang_step": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, POINTER(hf_box),
                                   POINTER(c_int32), _dp, c_double, c_double, c_double, c_int, c_double, _dp,
                                   POINTER(c_int32), POINTER(c_int64)]),
    "hf_amr_strang_step": (c_int, [c_void_p, POINTER(hf_level), c_int, POINTER(c_int32), _dp, c_double, c_int, c_int,
                                   c_double, _dp, POINTER(c_int32), POINTER(c_int64)]),
    "hf_repair_roundoff": (c_int, [c_void_p, c_void_p, POINTER(hf_box), POINTER(c_int64), POINTER(c_int64)]),
    "hf_validate_box": (c_int, [c_void_p, c_void_p, POINTER(hf_box), c_double, c_double, _dp, POINTER(c_int64)]),
    "hf_lincomb": (c_int, [c_void_p, c_void_p, c_void_p, c_double, c_double, c_int64]),
    "hf_synth_cells": (c_int, [c_void_p, c_int, c_uint64, c_int64, c_void_p, c_void_p]),
    "hf_equilibrate_tp": (c_int, [c_void_p, c_double, c_double, _dp, _dp]),
    "hf_state_from_TPY": (c_int, [c_void_p, c_double, c_double, _dp, c_double, c_double, _dp]),
}
EXPORTED = tuple(_SIGNATURES)

_lib = None


def load():
    """Load the HIP library; raises if it is absent or lacks an entry point."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is not built: run `python -m hyperflow_amd.build` "
                               "(there is no CPU fallback)")
        # torch's ROCm wheel ships its own HIP runtime under the same soname
        # (libamdhip64.so.7). Loading torch first makes this library bind to
        # that runtime, so torch tensors and these kernels share one runtime
        # (loaded the other way round, torch finds no device).
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


class HyperflowError(RuntimeError):
    def __init__(self, code, message, first_failed=-1):
        super().__init__(f"{ERROR_NAMES.get(code, code)}: {message}")
        self.code = code
        self.message = message
        self.first_failed = first_failed


def check(ctx, code, first_failed=-1):
    if code != HF_OK:
        msg = load().hf_last_error(ctx).decode() if ctx else load().hf_create_error().decode()
        raise HyperflowError(code, msg, first_failed)
