"""Python mirror of the reference's chemistry operator.

``ChemistryBatch`` keeps the reference's interface
(HyperFlow_V3/src/chemistry_batch.hpp:8-17): construct from a mechanism file,
``advance(field, dt, count)``, ``threads()``, ``lastTeamSize()``; failures keep
the field unchanged and name the lowest failing cell
(chemistry_batch.cpp:40-49). The field is SoA: shape (ns+3, n), component n
of cell i at ``field[n, i]`` (the reference's State order, solver.hpp:17-18).
Device tensors stay in HBM; host arrays are staged by the library.
"""
from __future__ import annotations

import ctypes
from ctypes import byref, c_int64, c_void_p

import numpy as np

from . import _lib
from ._lib import HF_CHEM_BATCHED, HF_CHEM_PLAIN, HF_PREC_FP64, check, hf_chem_stats, hf_options, load


def _ptr(x):
    """Raw pointer and leading dimension of a float64 2-D (or 1-D) array/tensor."""
    try:
        import torch
        if isinstance(x, torch.Tensor):
            if x.dtype != torch.float64:
                raise TypeError("expected float64")
            if x.dim() == 2 and x.stride(1) != 1:
                raise ValueError("rows must be contiguous")
            return x.data_ptr(), (x.stride(0) if x.dim() == 2 else x.numel())
    except ImportError:
        pass
    if isinstance(x, np.ndarray):
        if x.dtype != np.float64:
            raise TypeError("expected float64")
        if x.ndim == 2 and x.strides[1] != 8:
            raise ValueError("rows must be contiguous")
        return x.ctypes.data, (x.strides[0] // 8 if x.ndim == 2 else x.size)
    raise TypeError(f"unsupported array type {type(x)}")


class Mechanism:
    """A loaded mechanism bound to one HIP device (owns an hf_ctx)."""

    def __init__(self, path, rtol=1e-8, atol=1e-15, max_steps=5000, precision=HF_PREC_FP64, device=0,
                 waves_per_cu=0, chunk=0, nbins=0, eager=False, jac_wait=0, setup_wait=0):
        L = load()
        opt = hf_options(rtol, atol, max_steps, precision, device, waves_per_cu, chunk, nbins, int(bool(eager)),
                         jac_wait, setup_wait)
        h = c_void_p()
        check(None, L.hf_create(str(path).encode(), byref(opt), byref(h)))
        self._h = h
        self.path = str(path)
        self.ns = L.hf_num_species(h)
        self.nr = L.hf_num_reactions(h)
        self.nv = self.ns + 3
        self.species = [L.hf_species_name(h, k).decode() for k in range(self.ns)]
        W = np.zeros(self.ns)
        L.hf_molecular_weights(h, W.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        self.W = W

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None):
            load().hf_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def options(self):
        o = hf_options()
        check(self._h, load().hf_get_options(self._h, byref(o)))
        return o

    def set_options(self, **kw):
        o = self.options()
        for k, v in kw.items():
            setattr(o, k, v)
        check(self._h, load().hf_set_options(self._h, byref(o)))

    # ---------------------------------------------------------- set-up helpers
    def equilibrate_tp(self, T, p, X):
        X = np.ascontiguousarray(X, dtype=np.float64)
        Y = np.zeros(self.ns)
        dp = ctypes.POINTER(ctypes.c_double)
        check(self._h, load().hf_equilibrate_tp(self._h, T, p, X.ctypes.data_as(dp), Y.ctypes.data_as(dp)))
        return Y

    def state_from_TPY(self, T, p, Y, u=0.0, v=0.0):
        Y = np.ascontiguousarray(Y, dtype=np.float64)
        q = np.zeros(self.nv)
        dp = ctypes.POINTER(ctypes.c_double)
        check(self._h, load().hf_state_from_TPY(self._h, T, p, Y.ctypes.data_as(dp), u, v, q.ctypes.data_as(dp)))
        return q

    def synth_cells(self, which, seed, n, with_dt=True):
        """Seeded synthetic cells: (U (nv, n), dt (n,)) host arrays."""
        U = np.zeros((self.nv, n))
        dt = np.zeros(n) if with_dt else None
        check(self._h, load().hf_synth_cells(self._h, which, seed, n, U.ctypes.data,
                                             dt.ctypes.data if with_dt else None))
        return U, dt

    def properties(self, U):
        """properties() per cell -> (p, a, T) arrays like U's kind (numpy or torch)."""
        ptr, ld = _ptr(U)
        n = U.shape[1]
        out = [np.zeros(n) for _ in range(3)] if isinstance(U, np.ndarray) else [
            __import__("torch").empty(n, dtype=U.dtype, device=U.device) for _ in range(3)]
        ff = c_int64(-1)
        r = load().hf_properties(self._h, ptr, n, ld, *[_ptr(o)[0] for o in out], byref(ff))
        check(self._h, r, ff.value)
        return tuple(out)

    def stable_dt(self, U, dx, dy, cfl):
        ptr, ld = _ptr(U)
        dt = ctypes.c_double()
        check(self._h, load().hf_stable_dt(self._h, ptr, U.shape[1], ld, dx, dy, cfl, byref(dt)))
        return dt.value


def _interval(dt, n):
    """(array, per_cell) for hf_chem_advance: a Python number, a 0-d or
    one-element array/tensor is one interval for every cell; otherwise one
    float64 value per cell is required, at least n of them."""
    if np.isscalar(dt):
        return np.array([float(dt)]), 0
    size = dt.numel() if hasattr(dt, "numel") else np.size(dt)
    if size == 1:
        if isinstance(dt, np.ndarray):
            return np.array([float(dt.reshape(-1)[0])]), 0
        return dt.reshape(1), 0
    if dt.ndim != 1 or size < n:
        raise ValueError(f"per-cell intervals: need a 1-D array of at least {n} values, got shape {tuple(dt.shape)}")
    if isinstance(dt, np.ndarray) and not dt.flags.c_contiguous:
        raise ValueError("per-cell intervals must be contiguous")
    if hasattr(dt, "is_contiguous") and not dt.is_contiguous():
        raise ValueError("per-cell intervals must be contiguous")
    return dt, 1


class ChemistryBatch:
    """ChemistryBatch of HyperFlow_V3/src/chemistry_batch.hpp:8-17 on one GPU.

    ``mode``: HF_CHEM_PLAIN (one cell per lane, FP64 — the drop-in) or
    HF_CHEM_BATCHED (persistent stiffness-binned queue)."""

    def __init__(self, mechanism, threads=0, mode=HF_CHEM_PLAIN, **options):
        if threads < 0:
            raise ValueError("chemistry_threads must be nonnegative")
        self.mech = mechanism if isinstance(mechanism, Mechanism) else Mechanism(mechanism, **options)
        self.mode = mode
        self._last_team = 1
        self.last_stats = None

    def threads(self):
        return 1

    def lastTeamSize(self):
        return self._last_team

    def advance(self, field, dt, count=None, dt_per_cell=None):
        """Advance ``count`` cells of ``field`` (SoA, (ns+3, n)) by ``dt`` in place.

        dt: scalar, or an array of per-cell intervals. Raises HyperflowError
        (field unchanged) on any failing cell; ``.first_failed`` is its index."""
        if field.shape[0] != self.mech.nv:
            raise ValueError("field must have ns+3 component rows")
        n = field.shape[1] if count is None else int(count)
        if n > field.shape[1] or n < 0:
            raise ValueError("invalid chemistry batch extent or interval")
        ptr, ld = _ptr(field)
        dta, per = _interval(dt, n)
        dptr = _ptr(dta)[0]
        st = hf_chem_stats()
        ff = c_int64(-1)
        r = load().hf_chem_advance(self.mech.handle, ptr, n, ld, dptr, per, self.mode, byref(st), byref(ff))
        self.last_stats = st.as_dict()
        check(self.mech.handle, r, ff.value)
        self._last_team = 0 if n == 0 else st.team
        return self.last_stats
