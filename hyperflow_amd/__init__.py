"""MI355X-native hot path of HyperFlow V3: stiff per-cell chemistry and
NND/HLLC flux sweeps as HIP kernels for gfx950, behind the C ABI of
include/hyperflow.h. This package is the Python mirror of the reference's
operator interface for that path; the arithmetic runs in libhyperflow_hip.so.
"""
from __future__ import annotations

from pathlib import Path

from ._lib import (HF_BC_INFLOW, HF_BC_PERIODIC, HF_BC_REFLECTING, HF_BC_TRANSMISSIVE, HF_BC_WEDGE, HF_CHEM_BATCHED,
                   HF_CHEM_PLAIN, HF_PREC_FP64, HF_PREC_MIXED, HyperflowError, LIB_PATH, load)
from .amr import ConservativeInterp, average_down, face_fluxes, fill_ghosts, reflux, tag_cells, wedge_column
from .adjoint import ChemSensitivity, fixed_step_chemistry
from .chemistry import ChemistryBatch, Mechanism
from .physics import FluxSweep

MECHANISM_DIR = Path(__file__).resolve().parent / "mechanisms"
H2_LI = MECHANISM_DIR / "H2_Li.yaml"
DME_SK39 = MECHANISM_DIR / "dme_sk39.yaml"

__all__ = ["ChemistryBatch", "Mechanism", "ChemSensitivity", "fixed_step_chemistry", "FluxSweep", "HyperflowError",
           "load", "LIB_PATH", "H2_LI", "DME_SK39", "HF_CHEM_PLAIN", "HF_CHEM_BATCHED", "HF_PREC_FP64", "HF_PREC_MIXED",
           "HF_BC_TRANSMISSIVE", "HF_BC_REFLECTING", "HF_BC_PERIODIC", "HF_BC_INFLOW", "HF_BC_WEDGE",
           "ConservativeInterp", "average_down", "face_fluxes", "fill_ghosts", "reflux", "tag_cells",
           "wedge_column"]
