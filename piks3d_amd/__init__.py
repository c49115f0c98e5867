"""piks3d_amd -- MI355X DVM solver for rarefied gas flow through voxel images.

Python host for libpiks3d.so (C ABI in include/piks3d.h).  The solve loop
mirrors the reference program (src/main.F90:95-128): iterate, check
convergence every chkConvergeStep iterations on the mass flow through the
plane x = 2, stop when error <= eps, report the apparent permeability.

    from piks3d_amd import Solver, read_para, read_raw
    p = read_para("para.in")
    s = Solver(p, read_raw(p.image_file_name, p.nx, p.ny, p.nz))
    result = s.solve()                       # like the reference's main loop
    rho, ux, uy, uz = s.macro()              # reference cell layout

The native library is required: there is no CPU fallback in this package.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .images import read_raw, write_raw  # noqa: F401
from .params import Params, parse_para, read_para  # noqa: F401

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpiks3d.so")

PIKS_PREC_FP64 = 0
PIKS_PREC_FP32_REFINED = 1


class PiksParams(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("wall_ext_order", ctypes.c_int32), ("fluid_layer", ctypes.c_int32),
        ("ref_l", ctypes.c_double),
        ("nc_fundamental", ctypes.c_int32), ("half_range", ctypes.c_int32),
        ("chk_converge_step", ctypes.c_int32),
        ("kn", ctypes.c_double), ("press_drop", ctypes.c_double), ("accom", ctypes.c_double),
        ("precision", ctypes.c_int32), ("device", ctypes.c_int32),
        ("exact_quadrature", ctypes.c_int32), ("exact_order", ctypes.c_int32),
        ("sweep", ctypes.c_int32), ("tile", ctypes.c_int32), ("reserved", ctypes.c_int32 * 2),
    ]


class PiksInfo(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("nxt", ctypes.c_int32), ("nyt", ctypes.c_int32), ("nzt", ctypes.c_int32),
        ("ntotal", ctypes.c_int64),
        ("nc8", ctypes.c_int32), ("nwall", ctypes.c_int32), ("ncorner", ctypes.c_int32),
        ("nfluid", ctypes.c_int64), ("updates_per_iter", ctypes.c_int64),
        ("ds", ctypes.c_double), ("mu", ctypes.c_double), ("diff_flux", ctypes.c_double),
        ("mass", ctypes.c_double), ("device_bytes", ctypes.c_int64),
        ("coupled_walls", ctypes.c_int32), ("precision", ctypes.c_int32),
        ("steps_fp32", ctypes.c_int64), ("dc_block", ctypes.c_int32), ("sweep_path", ctypes.c_int32),
    ]


class ServeParams(ctypes.Structure):
    _fields_ = [
        ("device", ctypes.c_int32), ("max_slots", ctypes.c_int32),
        ("prep_threads", ctypes.c_int32), ("use_graphs", ctypes.c_int32),
        ("memory_budget", ctypes.c_int64), ("reserved", ctypes.c_int32 * 4),
    ]


class ServeResult(ctypes.Structure):
    _fields_ = [
        ("status", ctypes.c_int32), ("error_code", ctypes.c_int32),
        ("steps", ctypes.c_int32), ("steps32", ctypes.c_int32),
        ("converged", ctypes.c_int32), ("reserved0", ctypes.c_int32),
        ("permeability", ctypes.c_double), ("mass", ctypes.c_double), ("error", ctypes.c_double),
        ("queue_ms", ctypes.c_double), ("run_ms", ctypes.c_double), ("total_ms", ctypes.c_double),
        ("device_bytes", ctypes.c_int64),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if not k.startswith("reserved")}


REQ_QUEUED, REQ_RUNNING, REQ_DONE, REQ_FAILED = 0, 1, 2, 3

# every entry point include/piks3d.h declares, with its ctypes signature
_API = {
    "piks_create": (ctypes.c_int, [ctypes.POINTER(PiksParams), ctypes.c_void_p, ctypes.c_int64,
                                   ctypes.POINTER(ctypes.c_void_p)]),
    "piks_iterate": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "piks_set_precision": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "piks_set_correction_block": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "piks_mass_flow_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "piks_chk_converge": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]),
    "piks_mass_flow": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]),
    "piks_get_macro": (ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_void_p] * 4),
    "piks_get_f": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "piks_get_fw": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "piks_get_labels": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "piks_get_info": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(PiksInfo)]),
    "piks_preprocess": (ctypes.c_int, [ctypes.POINTER(PiksParams), ctypes.c_void_p, ctypes.c_int64,
                                       ctypes.c_void_p, ctypes.POINTER(PiksInfo)]),
    "piks_last_timing": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]),
    "piks_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "piks_last_error": (ctypes.c_char_p, []),
    "piks_create_multi": (ctypes.c_int, [ctypes.POINTER(PiksParams), ctypes.c_void_p, ctypes.c_int64,
                                         ctypes.c_int32, ctypes.POINTER(ctypes.c_double),
                                         ctypes.POINTER(ctypes.c_void_p)]),
    "piks_chk_converge_multi": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]),
    "piks_get_macro_rhs": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32] + [ctypes.c_void_p] * 4),
    "piks_serve_create": (ctypes.c_int, [ctypes.POINTER(ServeParams), ctypes.POINTER(ctypes.c_void_p)]),
    "piks_serve_submit": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(PiksParams), ctypes.c_void_p,
                                         ctypes.c_int64, ctypes.c_int32, ctypes.c_double, ctypes.c_double,
                                         ctypes.POINTER(ctypes.c_int64)]),
    "piks_serve_poll": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.POINTER(ServeResult)]),
    "piks_serve_wait": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.POINTER(ServeResult)]),
    "piks_serve_drain": (ctypes.c_int, [ctypes.c_void_p]),
    "piks_serve_destroy": (ctypes.c_int, [ctypes.c_void_p]),
}

_lib = None


class PiksError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("piks error %d: %s" % (code, msg))
        self.code = code


def lib():
    """Load libpiks3d.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("%s is missing: build it with `python -m piks3d_amd.build`" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _API.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise PiksError(rc, lib().piks_last_error().decode())


def to_cparams(p: Params, precision=None, exact_order=None, device=None) -> PiksParams:
    prec = p.precision if precision is None else precision
    prec_code = {"fp64": PIKS_PREC_FP64, "fp32": PIKS_PREC_FP32_REFINED}.get(prec, prec)
    cp = PiksParams()
    cp.nx, cp.ny, cp.nz = p.nx, p.ny, p.nz
    cp.wall_ext_order = p.wall_ext_order
    cp.fluid_layer = p.fluid_layer
    cp.ref_l = p.ref_l
    cp.nc_fundamental = p.nc_fundamental
    cp.half_range = int(bool(p.half_range))
    cp.chk_converge_step = p.chk_converge_step
    cp.kn, cp.press_drop, cp.accom = p.kn, p.press_drop, p.accom
    cp.precision = int(prec_code)
    cp.device = p.device if device is None else device
    cp.exact_quadrature = int(bool(p.exact_quadrature))
    cp.exact_order = int(bool(p.exact_order if exact_order is None else exact_order))
    cp.sweep = int(p.sweep)
    cp.tile = int(p.tile)
    return cp


def preprocess(params: Params, image):
    """Labels (reference cell order) and counts of an image, on the host only."""
    raw = np.ascontiguousarray(np.asarray(image, dtype=np.uint8).ravel())
    cp = to_cparams(params)
    info = PiksInfo()
    _check(lib().piks_preprocess(ctypes.byref(cp), raw.ctypes.data, raw.size, None, ctypes.byref(info)))
    labels = np.empty(info.ntotal, dtype=np.int32)
    _check(lib().piks_preprocess(ctypes.byref(cp), raw.ctypes.data, raw.size, labels.ctypes.data,
                                 ctypes.byref(info)))
    return labels, info


class Solver:
    """One DVM solve on one GPU (see module docstring)."""

    def __init__(self, params: Params, image, precision=None, exact_order=None, device=None):
        self.params = params
        raw = np.ascontiguousarray(np.asarray(image, dtype=np.uint8).ravel())
        self._cp = to_cparams(params, precision, exact_order, device)
        h = ctypes.c_void_p()
        _check(lib().piks_create(ctypes.byref(self._cp), raw.ctypes.data, raw.size, ctypes.byref(h)))
        self._h = h
        self.info = self.get_info()
        self.mass = 1.0
        self.history = []

    def close(self):
        if getattr(self, "_h", None):
            lib().piks_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def get_info(self) -> PiksInfo:
        info = PiksInfo()
        _check(lib().piks_get_info(self._h, ctypes.byref(info)))
        return info

    def iterate(self, n=1):
        _check(lib().piks_iterate(self._h, int(n)))

    def set_precision(self, precision):
        code = {"fp64": PIKS_PREC_FP64, "fp32": PIKS_PREC_FP32_REFINED}.get(precision, precision)
        _check(lib().piks_set_precision(self._h, int(code)))
        self.info = self.get_info()

    def set_correction_block(self, m):
        """Fix the defect-correction block length (m >= 1) of a refined
        context, or return to the adaptive one (m = 0)."""
        _check(lib().piks_set_correction_block(self._h, int(m)))
        self.info = self.get_info()

    def chk_converge(self):
        out = (ctypes.c_double * 3)()
        _check(lib().piks_chk_converge(self._h, out))
        self.mass = out[0]
        return out[0], out[1], out[2]

    def mass_flow(self):
        m = ctypes.c_double()
        _check(lib().piks_mass_flow(self._h, ctypes.byref(m)))
        return m.value

    def macro(self):
        n = self.info.ntotal
        arrs = [np.empty(n) for _ in range(4)]
        _check(lib().piks_get_macro(self._h, *[a.ctypes.data for a in arrs]))
        return tuple(arrs)

    def f(self):
        i = self.info
        out = np.empty(8 * i.nc8 * i.ntotal)
        _check(lib().piks_get_f(self._h, out.ctypes.data))
        return out.reshape(8, i.nc8, i.ntotal)

    def fw(self):
        i = self.info
        out = np.empty(max(i.ncorner, 1) * 8 * i.nc8)
        _check(lib().piks_get_fw(self._h, out.ctypes.data))
        return out[: i.ncorner * 8 * i.nc8].reshape(8, i.nc8, i.ncorner)

    def labels(self):
        out = np.empty(self.info.ntotal, dtype=np.int32)
        _check(lib().piks_get_labels(self._h, out.ctypes.data))
        return out

    def timing(self):
        ms = (ctypes.c_double * 4)()
        _check(lib().piks_last_timing(self._h, ms))
        return list(ms)

    def solve(self, max_step=None, chk_step=None, eps=None, callback=None, refine_switch=None):
        """The reference's main loop (src/main.F90:95-121): returns a dict with
        steps, permeability, mass, error and the check history.

        A context created with precision 'fp32' runs the library's FP64
        defect correction (FP32 correction iterations between FP64 residual
        iterations, include/piks3d.h piks_set_precision); `steps32` in the
        result counts its FP32 iterations.  refine_switch is accepted and
        ignored."""
        p = self.params
        max_step = p.max_step if max_step is None else max_step
        chk = p.chk_converge_step if chk_step is None else chk_step
        eps = p.eps if eps is None else eps
        start32 = self.get_info().steps_fp32
        error = 1.0
        perm = 0.0
        step = 0
        while step < max_step:
            # iterate to the next check (or the end) in one call, like the
            # reference's loop does one iteration at a time
            n = min(chk - step % chk, max_step - step)
            self.iterate(n)
            step += n
            if step % chk == 0:
                m, perm, error = self.chk_converge()
                self.history.append((step, m, perm, error))
                if callback:
                    callback(step, m, perm, error)
            if error <= eps:
                if step < chk:
                    m, perm, error = self.chk_converge()
                    self.history.append((step, m, perm, error))
                break
        self.info = self.get_info()
        return dict(steps=step, steps32=int(self.info.steps_fp32 - start32), permeability=perm, mass=self.mass,
                    error=error, history=list(self.history))


class Server:
    """Continuous-batching request server on one GPU (piks_serve_* in
    include/piks3d.h): submit returns at once; requests are preprocessed on
    host threads, admitted as slots and device memory free up, and run
    concurrently.  Each result is the same as a Solver running the same
    request alone."""

    def __init__(self, device=0, max_slots=8, prep_threads=4, use_graphs=True, memory_budget=0, tuning="auto"):
        """tuning: the sweep-schedule table of piks3d_amd.autotune applied to
        requests that choose no schedule ("auto": the package's tuning.json
        when present, None: none, or a path)."""
        from . import autotune
        self.tuning = autotune.load(autotune.TABLE if tuning == "auto" else tuning) if tuning else None
        sp = ServeParams()
        sp.device, sp.max_slots, sp.prep_threads = device, max_slots, prep_threads
        sp.use_graphs = int(bool(use_graphs))
        sp.memory_budget = int(memory_budget)
        h = ctypes.c_void_p()
        _check(lib().piks_serve_create(ctypes.byref(sp), ctypes.byref(h)))
        self._h = h

    def submit(self, params: Params, image, max_step=None, eps=None, precision=None, refine_switch=None):
        """Queue one solve (the reference's loop with params' chkConvergeStep);
        returns its ticket."""
        raw = np.ascontiguousarray(np.asarray(image, dtype=np.uint8).ravel())
        cp = to_cparams(params, precision)
        if cp.sweep == 0 and self.tuning:
            from . import autotune
            cp.sweep, cp.tile = autotune.choose(self.tuning, max(params.nx, params.ny, params.nz),
                                                params.nc_fundamental)
        t = ctypes.c_int64()
        _check(lib().piks_serve_submit(
            self._h, ctypes.byref(cp), raw.ctypes.data, raw.size,
            int(params.max_step if max_step is None else max_step),
            float(params.eps if eps is None else eps),
            float(params.refine_switch if refine_switch is None else refine_switch), ctypes.byref(t)))
        return t.value

    def poll(self, ticket):
        r = ServeResult()
        _check(lib().piks_serve_poll(self._h, ticket, ctypes.byref(r)))
        return r.as_dict()

    def wait(self, ticket):
        r = ServeResult()
        _check(lib().piks_serve_wait(self._h, ticket, ctypes.byref(r)))
        return r.as_dict()

    def drain(self):
        _check(lib().piks_serve_drain(self._h))

    def close(self):
        if getattr(self, "_h", None):
            lib().piks_serve_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class MultiSolver(Solver):
    """Several Knudsen numbers on one geometry (piks_create_multi): one
    context, one sweep per hyperplane for all right-hand sides."""

    def __init__(self, params: Params, image, kn_values, precision=None, exact_order=None, device=None):
        self.params = params
        self.kn_values = [float(k) for k in kn_values]
        raw = np.ascontiguousarray(np.asarray(image, dtype=np.uint8).ravel())
        self._cp = to_cparams(params, precision, exact_order, device)
        kn = (ctypes.c_double * len(self.kn_values))(*self.kn_values)
        h = ctypes.c_void_p()
        _check(lib().piks_create_multi(ctypes.byref(self._cp), raw.ctypes.data, raw.size,
                                       len(self.kn_values), kn, ctypes.byref(h)))
        self._h = h
        self.info = self.get_info()
        self.mass = 1.0
        self.history = []

    @property
    def nrhs(self):
        return len(self.kn_values)

    def chk_converge_all(self):
        """(mass, permeability, error) arrays over the right-hand sides."""
        out = (ctypes.c_double * (3 * self.nrhs))()
        _check(lib().piks_chk_converge_multi(self._h, out))
        a = np.array(out[:]).reshape(self.nrhs, 3)
        return a[:, 0], a[:, 1], a[:, 2]

    def macro_rhs(self, r):
        n = self.info.ntotal
        arrs = [np.empty(n) for _ in range(4)]
        _check(lib().piks_get_macro_rhs(self._h, int(r), *[a.ctypes.data for a in arrs]))
        return tuple(arrs)

    def solve(self, max_step=None, chk_step=None, eps=None, callback=None, refine_switch=None):
        """The reference's loop for every right-hand side at once: stops when
        each has error <= eps (or at max_step).  An FP32-refined context runs
        the defect correction for all right-hand sides together, its block
        length following the largest error."""
        p = self.params
        max_step = p.max_step if max_step is None else max_step
        chk = p.chk_converge_step if chk_step is None else chk_step
        eps = p.eps if eps is None else eps
        start32 = self.get_info().steps_fp32
        error = np.ones(self.nrhs)
        perm = np.zeros(self.nrhs)
        step = 0
        while step < max_step:
            n = min(chk - step % chk, max_step - step)
            self.iterate(n)
            step += n
            if step % chk:
                continue
            m, perm, error = self.chk_converge_all()
            self.history.append((step, m.copy(), perm.copy(), error.copy()))
            if callback:
                callback(step, m, perm, error)
            if np.all(error <= eps):
                break
        self.info = self.get_info()
        return dict(steps=step, steps32=int(self.info.steps_fp32 - start32), kn=list(self.kn_values),
                    permeability=perm, error=error, history=list(self.history))
