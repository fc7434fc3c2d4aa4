"""CPU oracle for the HyperFlow V3 hot path — TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
bench.py may import this package, as the checker (and the timed CPU
restatement of the reference path). The product library never loads it.

Parity status per function (how each is pinned; see the tests named):
  * NASA7 thermo, e/cv, T(e) inversion, properties: closed forms, FD
    derivative identities, numpy restatement (tests/test_oracle_thermo.py).
  * hllc / minmod: consistency, supersonic upwinding, stationary contact,
    mirror and rotation symmetry, Toro's textbook HLLC (tests/test_oracle_flux.py).
  * NND flux stage: uniform state preserved, conservation on periodic boxes
    (tests/test_oracle_flux.py).
  * kinetics: element conservation, detailed balance at equilibrium,
    Troe limits, unit-conversion known answer, numpy restatement
    (tests/test_oracle_kinetics.py).
  * BDF integrator: linear closed forms, Robertson reference values,
    scipy Radau golden trajectories (tests/test_oracle_bdf.py, tests/golden/).
  * ChemistryBatch semantics: all-or-nothing, lowest index, invalid dt,
    count 0 (tests/test_oracle_chem.py).
  * fixed-step backward-Euler objective (no reference counterpart): the
    energy identity h qdot(y_1) = rho cv (T_1 - T_0) on single steps,
    first-order convergence of T(t_end) to the converged BDF solution, and
    the ln A shift entering the rates (tests/test_oracle_adjoint.py). It is
    the forward map whose central differences pin the GPU adjoint.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

from . import mech as mech_py

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lp = ctypes.POINTER(ctypes.c_long)


def build(force=False):
    """Compile oracle/liboracle.so with g++ (OpenMP)."""
    import subprocess
    src = HERE / "oracle.cpp"
    if force or not LIB_PATH.exists() or LIB_PATH.stat().st_mtime < src.stat().st_mtime:
        subprocess.check_call(["make", "-C", str(HERE), "-s"])
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = ctypes.CDLL(str(LIB_PATH))
        L.or_mech_create.restype = ctypes.c_void_p
        L.or_mech_destroy.argtypes = [ctypes.c_void_p]
        L.or_last_error.restype = ctypes.c_char_p
        _lib = L
    return _lib


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(_dp)


def _i(a):
    return np.ascontiguousarray(a, dtype=np.int32).ctypes.data_as(_ip)


class Oracle:
    """CPU restatement bound to one mechanism file (tables read by PyYAML)."""

    def __init__(self, path):
        self.m = mech_py.load(path)
        m = self.m
        self.ns, self.nr = m["ns"], m["nr"]
        self._keep = [np.ascontiguousarray(m[k], dtype=np.float64 if m[k].dtype.kind == "f" else np.int32)
                      for k in ("W", "Tmid", "low", "high", "kind", "rev", "A", "b", "EaR", "A0", "b0", "Ea0R",
                                "troe", "troeA", "troeT3", "troeT1", "troeT2", "nu_r", "nu_p", "eff")]
        args = []
        for a in self._keep:
            args.append(a.ctypes.data_as(_dp if a.dtype == np.float64 else _ip))
        self.h = lib().or_mech_create(self.ns, self.nr, *args)
        if not self.h:
            raise RuntimeError("oracle mechanism create failed")
        self.h = ctypes.c_void_p(self.h)

    def __del__(self):
        try:
            lib().or_mech_destroy(self.h)
        except Exception:
            pass

    # ------------------------------------------------------------ thermo
    def e_cv(self, T, Y):
        e, cv = ctypes.c_double(), ctypes.c_double()
        lib().or_e_cv(self.h, ctypes.c_double(T), _d(Y), ctypes.byref(e), ctypes.byref(cv))
        return e.value, cv.value

    def temperature(self, e, Y, T0=1000.0):
        T = ctypes.c_double(T0)
        r = lib().or_temperature(self.h, ctypes.c_double(e), _d(Y), ctypes.byref(T))
        if r:
            raise ValueError("inadmissible state")
        return T.value

    def properties(self, U):
        """U (n, nv) AoS -> (p, a, T) arrays and status (0 ok, 2 inadmissible)."""
        U = np.ascontiguousarray(np.atleast_2d(U), dtype=np.float64)
        n = U.shape[0]
        out = np.zeros((n, 3))
        st = np.zeros(n, np.int32)
        lib().or_properties(self.h, _d(U), ctypes.c_long(n), out.ctypes.data_as(_dp), st.ctypes.data_as(_ip))
        return out[:, 0], out[:, 1], out[:, 2], st

    def hllc(self, L, R, n, pl, pr):
        f = np.zeros(self.ns + 3)
        lib().or_hllc(self.h, _d(L), _d(R), ctypes.c_int(n), _d(pl), _d(pr), f.ctypes.data_as(_dp))
        return f

    def wdot(self, T, C, lnA_shift=None):
        w = np.zeros(self.ns)
        sh = _d(lnA_shift) if lnA_shift is not None else None
        lib().or_wdot(self.h, ctypes.c_double(T), _d(C), w.ctypes.data_as(_dp), sh)
        return w

    # ------------------------------------------------------------ flux stage
    def flux_stage(self, Uin, dtdx, dtdy, first_order=False, walls=0, dims=2, wedge_i=0, fluxes=False):
        """Uin: (nv, ny+4, nx+4) ghosted SoA -> (status, Uout (nv, ny, nx), fallbacks, first_failed),
        plus (Fx (nv, ny, nx+1), Fy (nv, ny+1, nx)) face fluxes when fluxes is set."""
        Uin = np.ascontiguousarray(Uin, dtype=np.float64)
        nv, gy, gx = Uin.shape
        nx, ny = gx - 4, gy - 4
        out = np.zeros((nv, ny, nx))
        fb, ff = ctypes.c_long(0), ctypes.c_long(-1)
        fx = np.zeros((ny, nx + 1, nv)) if fluxes else None
        fy = np.zeros((ny + 1, nx, nv)) if fluxes else None
        r = lib().or_flux_stage(self.h, nx, ny, _d(Uin), out.ctypes.data_as(_dp), ctypes.c_double(dtdx),
                                ctypes.c_double(dtdy), int(first_order), int(walls), int(wedge_i), int(dims),
                                ctypes.byref(fb), ctypes.byref(ff),
                                fx.ctypes.data_as(_dp) if fluxes else None, fy.ctypes.data_as(_dp) if fluxes else None)
        if fluxes:
            return r, out, fb.value, ff.value, fx.transpose(2, 0, 1).copy(), fy.transpose(2, 0, 1).copy()
        return r, out, fb.value, ff.value

    # ------------------------------------------------------------ chemistry
    def chem_advance(self, U, dt, rtol=1e-10, atol=1e-14, maxsteps=5000, threads=0):
        """ChemistryBatch::advance on SoA planes U (nv, ncell), in place on success.

        Returns (status, stats dict, first_failed, message)."""
        assert U.dtype == np.float64 and U.flags.c_contiguous
        nv, n = U.shape
        dt = np.atleast_1d(np.asarray(dt, dtype=np.float64))
        per = int(dt.size > 1)
        if per:
            assert dt.size == n
        stats = np.zeros(7, np.int64)
        ff = ctypes.c_long(-1)
        r = lib().or_chem_advance(self.h, U.ctypes.data_as(_dp), ctypes.c_long(n), ctypes.c_long(n), _d(dt), per,
                                  ctypes.c_double(rtol), ctypes.c_double(atol), int(maxsteps), int(threads),
                                  stats.ctypes.data_as(_lp), ctypes.byref(ff))
        keys = ("steps", "rhs", "jac", "lu", "newton", "conv_fail", "err_fail")
        msg = lib().or_last_error().decode() if r else ""
        return r, dict(zip(keys, stats.tolist())), ff.value, msg

    def chem_fixed_be(self, rho, T0, Y0, h, nsteps, lnA_shift=None):
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        n = rho.size
        J = np.zeros(n)
        Yend = np.zeros((n, self.ns))
        Tend = np.zeros(n)
        sh = _d(lnA_shift) if lnA_shift is not None else None
        r = lib().or_chem_fixed_be(self.h, ctypes.c_long(n), _d(rho), _d(T0), _d(Y0), ctypes.c_double(h),
                                   int(nsteps), sh, J.ctypes.data_as(_dp), Yend.ctypes.data_as(_dp),
                                   Tend.ctypes.data_as(_dp))
        return r, J, Tend, Yend


def bdf_linear(A, y0, tout, rtol, atol, maxsteps=100000):
    A = np.ascontiguousarray(A, dtype=np.float64)
    y = np.array(y0, dtype=np.float64)
    st = np.zeros(7, np.int64)
    r = lib().or_bdf_linear(A.shape[0], _d(A), y.ctypes.data_as(_dp), ctypes.c_double(tout),
                            ctypes.c_double(rtol), ctypes.c_double(atol), int(maxsteps), st.ctypes.data_as(_lp))
    return r, y, st


def bdf_robertson(y0, tout, rtol, atol, maxsteps=100000):
    y = np.array(y0, dtype=np.float64)
    st = np.zeros(7, np.int64)
    r = lib().or_bdf_robertson(y.ctypes.data_as(_dp), ctypes.c_double(tout), ctypes.c_double(rtol),
                               ctypes.c_double(atol), int(maxsteps), st.ctypes.data_as(_lp))
    return r, y, st
