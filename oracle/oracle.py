"""ctypes binding of the FP64 CPU oracle (oracle/meso_oracle.h).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker. The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "meso-lts_amd"))

from meso_hip.capi import (DvsDesc, MeshDesc, StepParams, StepReport, MESO_OK, MESO_E_NUMERIC,  # noqa: E402
                           LAYOUT_VELOCITY_MAJOR, MesoError, dptr, as_f64)

LIB_PATH = os.path.join(HERE, "liboracle_meso.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MesoError(-1, f"{LIB_PATH} is not built (make -C oracle)")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        _d = C.POINTER(C.c_double)
        sig = {
            "oracle_create": ([C.POINTER(MeshDesc), C.POINTER(DvsDesc), C.POINTER(StepParams), C.POINTER(vp)], C.c_int),
            "oracle_destroy": ([vp], None),
            "oracle_last_error": ([], C.c_char_p),
            "oracle_set_threads": ([C.c_int32], None),
            "oracle_get_threads": ([], C.c_int32),
            "oracle_set_params": ([vp, C.POINTER(StepParams)], C.c_int),
            "oracle_set_mask": ([vp, C.POINTER(C.c_int32)], C.c_int),
            "oracle_init_equilibrium": ([vp, _d, _d, _d], C.c_int),
            "oracle_init_from_macros": ([vp, _d, _d, _d, _d], C.c_int),
            "oracle_set_state": ([vp, _d, C.c_int32, _d, _d, _d, _d], C.c_int),
            "oracle_get_state": ([vp, _d, C.c_int32], C.c_int),
            "oracle_step": ([vp, C.c_int32, C.POINTER(StepReport)], C.c_int),
            "oracle_get_macros": ([vp, _d, _d, _d, _d], C.c_int),
            "oracle_get_flux_moments": ([vp, _d, _d, _d], C.c_int),
            "oracle_residual": ([vp, _d], C.c_int),
            "oracle_grad": ([vp, _d, _d], C.c_int),
            "oracle_face_values": ([vp, _d], C.c_int),
            "oracle_fbar": ([vp, _d, _d], C.c_int),
        }
        for name, (args, res) in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def set_threads(n: int) -> None:
    lib().oracle_set_threads(n)


def get_threads() -> int:
    return int(lib().oracle_get_threads())


class Oracle:
    """FP64 restatement of meso's step on the descriptors of a meso_hip.Case."""

    def __init__(self, case, params: StepParams | None = None):
        L = lib()
        h = C.c_void_p()
        rc = L.oracle_create(C.byref(case.mesh), C.byref(case.dvs), C.byref(params or case.params), C.byref(h))
        if rc != MESO_OK:
            raise MesoError(rc, L.oracle_last_error().decode())
        self._h = h
        self.case = case
        self.nc, self.nv, self.nf, self.nd = case.n_cells, case.n_vel, case.n_faces, case.n_dist

    def _check(self, rc):
        if rc not in (MESO_OK,):
            raise MesoError(rc, lib().oracle_last_error().decode())

    def close(self):
        if getattr(self, "_h", None):
            lib().oracle_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, p):
        self._check(lib().oracle_set_params(self._h, C.byref(p)))

    def set_mask(self, mask=None):
        """Cells repaired instead of stopping (oracle_set_mask); None takes the
        case's `mask switch True` groups."""
        if mask is None:
            mask = self.case.cell_mask()
        m = np.ascontiguousarray(mask, dtype=np.int32)
        self._check(lib().oracle_set_mask(self._h, m.ctypes.data_as(C.POINTER(C.c_int32))))

    def init_equilibrium(self, rho=None, u=None, T=None):
        if rho is None:
            rho, u, T = self.case.initial_patch()
        self._check(lib().oracle_init_equilibrium(self._h, dptr(as_f64(rho)), dptr(as_f64(u)), dptr(as_f64(T))))

    def init_from_macros(self, rho, u, T, q):
        self._check(lib().oracle_init_from_macros(self._h, dptr(as_f64(rho)), dptr(as_f64(u)), dptr(as_f64(T)),
                                                  dptr(as_f64(q))))

    def set_state(self, f, rho, u, T=None, q=None, layout=LAYOUT_VELOCITY_MAJOR):
        self._check(lib().oracle_set_state(self._h, dptr(as_f64(f)), layout, dptr(as_f64(rho)), dptr(as_f64(u)),
                                           None if T is None else dptr(as_f64(T)),
                                           None if q is None else dptr(as_f64(q))))

    def get_state(self):
        f = np.zeros(self.nd * self.nv * self.nc)
        self._check(lib().oracle_get_state(self._h, dptr(f), LAYOUT_VELOCITY_MAJOR))
        return f.reshape(self.nd, self.nv, self.nc)

    def step(self, n=1):
        rep = StepReport()
        rc = lib().oracle_step(self._h, n, C.byref(rep))
        if rc not in (MESO_OK, MESO_E_NUMERIC):
            self._check(rc)
        return rep

    def macros(self):
        nc = self.nc
        rho, u, T, q = np.zeros(nc), np.zeros(3 * nc), np.zeros(nc), np.zeros(3 * nc)
        self._check(lib().oracle_get_macros(self._h, dptr(rho), dptr(u), dptr(T), dptr(q)))
        return dict(rho=rho, u=u.reshape(nc, 3), T=T, q=q.reshape(nc, 3))

    def flux_moments(self):
        nc = self.nc
        m0, m1, m2 = np.zeros(nc), np.zeros(3 * nc), np.zeros(nc)
        self._check(lib().oracle_get_flux_moments(self._h, dptr(m0), dptr(m1), dptr(m2)))
        return m0, m1.reshape(nc, 3), m2

    def residual(self):
        out = np.zeros(8)
        self._check(lib().oracle_residual(self._h, dptr(out)))
        return out

    def grad(self, field):
        g = np.zeros(3 * self.nc)
        self._check(lib().oracle_grad(self._h, dptr(as_f64(field)), dptr(g)))
        return g.reshape(self.nc, 3)

    def face_values(self):
        ff = np.zeros(self.nd * self.nv * self.nf)
        self._check(lib().oracle_face_values(self._h, dptr(ff)))
        return ff.reshape(self.nd, self.nv, self.nf)

    def fbar(self, cell_f):
        """f-bar+ at the faces of a velocity-major field [NV][NC] (no mixing, no boundaries)."""
        cf = as_f64(cell_f).reshape(self.nv, self.nc)
        ff = np.zeros(self.nv * self.nf)
        self._check(lib().oracle_fbar(self._h, dptr(cf), dptr(ff)))
        return ff.reshape(self.nv, self.nf)
