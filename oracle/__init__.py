"""ctypes front end of the C oracle (oracle/piks_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, where it is the checker (or the timed CPU
baseline).  The product package never imports it.

The oracle restates the reference's single-rank iteration with identical
floating-point operations; arrays come back in the reference's layouts
(f as (8, Nc8, Ntotal) = Fortran f(Ntotal, Nc8, 8), cell index
l = (i+2) + (j+1)*Nxtotal + (k+1)*Nxytotal - 1 for i, j, k in [-1, N+2]).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_ref", "libpiks_oracle.so")


class _Params(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("wall_ext_order", ctypes.c_int32), ("fluid_layer", ctypes.c_int32),
        ("ref_l", ctypes.c_double),
        ("nc_fundamental", ctypes.c_int32), ("half_range", ctypes.c_int32),
        ("chk_converge_step", ctypes.c_int32),
        ("kn", ctypes.c_double), ("press_drop", ctypes.c_double), ("accom", ctypes.c_double),
        ("exact_quadrature", ctypes.c_int32),
    ]


_lib = None


def build():
    """Compile the oracle library (gcc, -ffp-contract=off) into oracle/_ref/."""
    subprocess.run(["make", "-s", "-C", HERE, "oracle"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.po_create.restype = ctypes.c_void_p
        L.po_create.argtypes = [ctypes.POINTER(_Params), ctypes.c_void_p, ctypes.c_int64,
                                ctypes.c_char_p, ctypes.c_int]
        L.po_destroy.argtypes = [ctypes.c_void_p]
        L.po_iterate.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        L.po_chk_converge.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]
        L.po_info.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)]
        L.po_info_d.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]
        for name in ("po_f", "po_fw", "po_rho", "po_ux", "po_uy", "po_uz", "po_extcoef",
                     "po_cx", "po_cy", "po_cz", "po_w"):
            getattr(L, name).restype = ctypes.POINTER(ctypes.c_double)
            getattr(L, name).argtypes = [ctypes.c_void_p]
        for name in ("po_image", "po_which_corner", "po_vecwall", "po_mapf"):
            getattr(L, name).restype = ctypes.POINTER(ctypes.c_int32)
            getattr(L, name).argtypes = [ctypes.c_void_p]
        _lib = L
    return _lib


class Oracle:
    """One oracle solve.  `params` is a dict with para.in names (lower case):
    nx, ny, nz, wallextorder, fluidlayer, ref_l, nc_fundamental, halfrange,
    chkconvergestep, kn, pressdrop, accom (+ optional exact_quadrature)."""

    def __init__(self, params: dict, image: np.ndarray | bytes):
        L = lib()
        p = _Params(
            int(params["nx"]), int(params["ny"]), int(params["nz"]),
            int(params.get("wallextorder", 2)), int(params.get("fluidlayer", 0)),
            float(params.get("ref_l", 1.0)),
            int(params["nc_fundamental"]), int(bool(params.get("halfrange", True))),
            int(params.get("chkconvergestep", 10)),
            float(params["kn"]), float(params["pressdrop"]), float(params.get("accom", 1.0)),
            int(params.get("exact_quadrature", 0)),
        )
        raw = np.ascontiguousarray(np.frombuffer(bytes(image), dtype=np.uint8)
                                   if isinstance(image, (bytes, bytearray))
                                   else np.asarray(image, dtype=np.uint8).ravel())
        self._raw = raw
        err = ctypes.create_string_buffer(512)
        h = L.po_create(ctypes.byref(p), raw.ctypes.data, raw.size, err, 512)
        if not h:
            raise ValueError(err.value.decode())
        self._h = h
        info = (ctypes.c_int64 * 19)()
        L.po_info(h, info)
        self.info = list(info)
        (self.Nx, self.Ny, self.Nz, self.Nxt, self.Nyt, self.Nzt, self.Ntotal, self.Nc8,
         self.nwall, self.nfluid, self.ncorner) = self.info[:11]
        self.nstencil = self.info[11:19]
        d = (ctypes.c_double * 4)()
        L.po_info_d(h, d)
        self.ds, self.mu, self.diff_flux = d[0], d[1], d[2]

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().po_destroy(h)
            self._h = None

    def _dbl(self, name, n):
        ptr = getattr(lib(), name)(self._h)
        return np.ctypeslib.as_array(ptr, shape=(n,)).copy()

    def _int(self, name, n):
        ptr = getattr(lib(), name)(self._h)
        return np.ctypeslib.as_array(ptr, shape=(n,)).copy()

    def iterate(self, n=1, threads=1):
        lib().po_iterate(self._h, int(n), int(threads))

    def chk_converge(self):
        out = (ctypes.c_double * 3)()
        lib().po_chk_converge(self._h, out)
        return out[0], out[1], out[2]

    def f(self):
        return self._dbl("po_f", self.Ntotal * self.Nc8 * 8).reshape(8, self.Nc8, self.Ntotal)

    def fw(self):
        n = max(self.ncorner, 1) * self.Nc8 * 8
        return self._dbl("po_fw", n).reshape(8, self.Nc8, max(self.ncorner, 1))[:, :, :self.ncorner]

    def macro(self):
        return tuple(self._dbl(n, self.Ntotal) for n in ("po_rho", "po_ux", "po_uy", "po_uz"))

    def image(self):
        return self._int("po_image", self.Ntotal)

    def which_corner(self):
        return self._int("po_which_corner", self.Ntotal)

    def vecwall(self):
        return self._int("po_vecwall", self.nwall)

    def extcoef(self):
        return self._dbl("po_extcoef", 2 * self.nwall).reshape(2, self.nwall)

    def mapf(self):
        return self._int("po_mapf", self.nfluid)

    def velocity(self):
        return tuple(self._dbl(n, self.Nc8) for n in ("po_cx", "po_cy", "po_cz", "po_w"))

    def grid_shape(self):
        """(Nztotal, Nytotal, Nxtotal) for reshaping cell arrays."""
        return (self.Nzt, self.Nyt, self.Nxt)
