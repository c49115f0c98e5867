"""piks_mass_flow_noalloc: the allocation-free variant of piks_mass_flow (its
plane staged in the library's fixed per-device array) gives the same mass
flow bit for bit, call after call, the second call with a smaller plane; a
plane beyond the array's size is refused."""
import ctypes

import numpy as np
import pytest

from piks3d_amd import Solver, images, lib
from piks3d_amd.params import Params

pytestmark = pytest.mark.gpu


def _solver(n, seed):
    img = images.sphere_pack(n, 0.40, 3.0, 6.0, seed=seed)
    p = Params(nx=n, ny=n, nz=n, fluid_layer=2, wall_ext_order=2, nc_fundamental=2, half_range=True,
               chk_converge_step=10, max_step=100, eps=0.0, kn=1.0, press_drop=1.0, accom=1.0)
    return Solver(p, img, precision="fp64")


class _DeviceDoubles:
    """n doubles of device memory for the entry point's output, allocated by
    the test through the HIP runtime the library itself uses."""

    def __init__(self, n):
        lib()
        self.hip = ctypes.CDLL("libamdhip64.so.7")
        self.n = n
        self.ptr = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(8 * n)) == 0

    def at(self, q):
        return ctypes.c_void_p(self.ptr.value + 8 * q)

    def read(self):
        h = np.zeros(self.n)
        # hipMemcpyDeviceToHost; the caller has synchronised the writer's stream
        assert self.hip.hipMemcpy(ctypes.c_void_p(h.ctypes.data), self.ptr, ctypes.c_size_t(8 * self.n), 2) == 0
        return h

    def free(self):
        self.hip.hipFree(self.ptr)


def _noalloc(s, dptr):
    f = lib().piks_mass_flow_noalloc
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    return f(s._h, dptr)


def test_noalloc_mass_flow_equals_mass_flow():
    out = _DeviceDoubles(2)
    try:
        for q, (n, steps) in enumerate([(48, 7), (20, 3)]):   # the second plane is smaller
            print("noalloc call %d: %d^3 image, %d iterations" % (q, n, steps), flush=True)
            with _solver(n, 30 + q) as s:
                s.iterate(steps)
                assert _noalloc(s, out.at(q)) == 0
                ref = s.mass_flow()               # synchronises the context's stream
                got = out.read()[q]
                assert got == ref, (n, got, ref)
                assert ref != 0.0
    finally:
        out.free()


def test_noalloc_refuses_a_plane_beyond_its_array():
    out = _DeviceDoubles(1)
    img = np.zeros((730, 730, 4), dtype=np.uint8)      # Ny * Nz = 532900 > 524288
    p = Params(nx=4, ny=730, nz=730, fluid_layer=0, wall_ext_order=2, nc_fundamental=2, half_range=True,
               chk_converge_step=10, max_step=1, eps=0.0, kn=1.0, press_drop=1.0, accom=1.0)
    try:
        with Solver(p, img, precision="fp64") as s:
            rc = _noalloc(s, out.at(0))
            assert rc == -1
            msg = lib().piks_last_error().decode()
            assert "allocation-free limit" in msg, msg
    finally:
        out.free()
