"""The request server (piks_serve_*): continuous batching of independent
solves.  Each request's result must be bit-for-bit the result of the same
request solved alone through piks_create/piks_iterate/piks_chk_converge: the
server runs the same kernels in the same order per solve (one iteration per
HIP graph launch), so the stated tolerance is zero."""
import numpy as np
import pytest

from piks3d_amd import REQ_DONE, REQ_FAILED, PiksError, Server, Solver, images
from piks3d_amd.params import Params

pytestmark = pytest.mark.gpu


def _request(n, phi, kn, nc, seed, max_step=200, eps=1e-7):
    img = images.sphere_pack(n, phi, 2.5, 5.0, seed=seed)
    p = Params(nx=n, ny=n, nz=n, fluid_layer=2, wall_ext_order=2, nc_fundamental=nc, half_range=True,
               chk_converge_step=10, max_step=max_step, eps=eps, kn=kn, press_drop=1.0, accom=1.0,
               exact_quadrature=(nc == 3))
    return p, img


MIX = [  # side, porosity, Kn, Nc_fundamental, seed, precision
    (24, 0.45, 0.5, 2, 1, "fp64"),
    (32, 0.40, 2.0, 3, 2, "fp64"),
    (20, 0.35, 0.05, 4, 3, "fp64"),
    (28, 0.50, 8.0, 2, 4, "fp32"),
    (36, 0.30, 1.0, 4, 5, "fp64"),
    (16, 0.45, 0.2, 3, 6, "fp32"),
    (24, 0.40, 3.0, 2, 7, "fp64"),
]


def test_results_in_the_mix_equal_results_alone():
    reqs = [(_request(n, phi, kn, nc, seed), prec) for n, phi, kn, nc, seed, prec in MIX]
    alone = []
    for (p, img), prec in reqs:
        with Solver(p, img, precision=prec) as s:
            alone.append(s.solve())
    # three slots for seven requests: later ones are admitted as earlier finish
    with Server(max_slots=3, prep_threads=2) as srv:
        tickets = [srv.submit(p, img, precision=prec) for (p, img), prec in reqs]
        got = [srv.wait(t) for t in tickets]
    for a, b, spec in zip(alone, got, MIX):
        assert b["status"] == REQ_DONE, (spec, b)
        assert b["steps"] == a["steps"] and b["steps32"] == a["steps32"], (spec, a, b)
        assert b["permeability"] == a["permeability"], (spec, a["permeability"], b["permeability"])
        assert b["error"] == a["error"] and b["mass"] == a["mass"], spec
        assert b["queue_ms"] >= 0 and b["run_ms"] > 0 and b["total_ms"] >= b["run_ms"]


def test_admission_by_memory_and_failures():
    small = _request(16, 0.45, 1.0, 2, 11, max_step=30)
    big = _request(40, 0.40, 1.0, 4, 12, max_step=30)
    bad_img = small[1].copy()
    bad_img[3, 3, 3] = 9
    with Server(max_slots=4, prep_threads=2, memory_budget=64 << 20) as srv:
        t_small = [srv.submit(*small) for _ in range(3)]
        t_big = srv.submit(*big)
        t_bad = srv.submit(small[0], bad_img)
        srv.drain()
        res = [srv.poll(t) for t in t_small]
        rb, rbad = srv.poll(t_big), srv.poll(t_bad)
        with pytest.raises(PiksError):
            srv.poll(10 ** 9)
    assert all(r["status"] == REQ_DONE for r in res)
    assert len({r["permeability"] for r in res}) == 1            # same request, same answer
    assert all(r["device_bytes"] <= 64 << 20 for r in res)
    assert rb["status"] == REQ_FAILED and rb["error_code"] == -6  # over the memory budget
    assert rbad["status"] == REQ_FAILED and rbad["error_code"] == -2


def test_many_small_requests_overlap():
    reqs = [_request(16 + 4 * (q % 4), 0.4, 0.5 + q, 2, 100 + q, max_step=60) for q in range(12)]
    with Server(max_slots=6, prep_threads=4) as srv:
        tickets = [srv.submit(p, img) for p, img in reqs]
        srv.drain()
        res = [srv.poll(t) for t in tickets]
    assert all(r["status"] == REQ_DONE for r in res)
    # the last request waited for a slot, then ran like the others
    assert res[-1]["queue_ms"] > 0
    with Solver(*reqs[5]) as s:
        alone = s.solve()
    assert res[5]["permeability"] == alone["permeability"]
