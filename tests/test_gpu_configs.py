"""The benchmark configurations at their full sizes on the HIP path.

* 200^3 sphere pack (seed 2024, porosity 0.40, radii 4-10, Kn 1, 8^3
  velocities, wallExtOrder 2): 30 iterations against the reference program's
  own run of the same input (tests/golden/c2_spheres200.npz, written by
  tests/golden/make_golden.py from the reference built from its sources, one
  thread): the three check lines, the final Ux plane x = 2 and the fluid
  moment sums, bit for bit with exact_order = 1 and within 1e-10 in the
  default reduction order.  The same pack with wallExtOrder 1 (the iteration
  the reference keeps bounded on sphere packs): the refined precision against
  plain FP64 at equal iteration counts, |dK| <= 1e-8 K.
* 300^3 Gaussian random field (correlation length 6, porosity 0.20,
  percolating in x) at Kn 0.01, 1 and 10: refined against FP64 at equal
  iteration counts (1e-8 on K), and the permeability's independence of the
  pressure drop (linearity: f scales with pressDrop, K = m / pressDrop ...).
* The serving stream's request mix (sides 32-128, 6^3 / 8^3 / 12^3 velocity
  grids): every result equal, bit for bit, to the same request solved alone.
* The 160^3 random field (correlation length 5, porosity 0.25) with all 32
  Knudsen numbers of the sweep (log-spaced over 1e-3..10) as right-hand sides
  of multi-RHS contexts: every right-hand side equal, bit for bit, to its own
  single solve.
"""
import hashlib
import os
import time

import numpy as np
import pytest

from conftest import GOLDEN
from piks3d_amd import REQ_DONE, MultiSolver, Server, Solver, images
from piks3d_amd.params import Params

pytestmark = pytest.mark.gpu

RTOL_RUN = 1e-10


def _c2_params(**kw):
    p = Params(nx=200, ny=200, nz=200, fluid_layer=4, wall_ext_order=2, nc_fundamental=4, half_range=True,
               chk_converge_step=10, max_step=30, eps=0.0, kn=1.0, press_drop=1.0, accom=1.0)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


_C2_IMG = {}


def _c2_image():
    if "img" not in _C2_IMG:
        _C2_IMG["img"] = images.sphere_pack(200, 0.40, 4.0, 10.0, seed=2024)
    return _C2_IMG["img"]


@pytest.mark.parametrize("exact", [True, False])
def test_c2_spheres200_against_reference_run(exact):
    print("200^3 sphere pack, reference run, exact_order=%d" % exact, flush=True)
    z = np.load(os.path.join(GOLDEN, "c2_spheres200.npz"))
    img = _c2_image()
    assert hashlib.sha256(np.ascontiguousarray(img).tobytes()).hexdigest() == str(z["image_sha256"])
    p = _c2_params()
    steps = [int(s) for s in z["conv_step"]]
    conv = z["conv"]
    with Solver(p, img, precision="fp64", exact_order=exact) as s:
        lab = s.labels().astype(np.int8)
        assert hashlib.sha256(lab.tobytes()).hexdigest() == str(z["label_sha256"])
        done = 0
        for q, st in enumerate(steps):
            s.iterate(st - done)
            done = st
            got = np.array(s.chk_converge())
            print("step %d: %r (reference %r)" % (st, got.tolist(), conv[q].tolist()), flush=True)
            if exact:
                assert np.array_equal(got, conv[q]), (st, got, conv[q])
            else:
                assert abs(got[0] - conv[q][0]) <= RTOL_RUN * abs(conv[q][0]), (st, got, conv[q])
                assert abs(got[1] - conv[q][1]) <= RTOL_RUN * abs(conv[q][1]), (st, got, conv[q])
        mac = np.stack(s.macro())
    fluid = lab == 0
    sums = mac[:, fluid].sum(axis=1)
    sq = (mac[:, fluid] ** 2).sum(axis=1)
    ux = mac[1].reshape(204, 204, 212)[2:-2, 2:-2, 3]
    if exact:
        assert np.array_equal(ux, z["ux_plane2_final"])
        assert np.array_equal(sums, z["macro_final_fluidsum"])
        assert np.array_equal(sq, z["macro_final_fluidsq"])
    else:
        assert np.abs(ux - z["ux_plane2_final"]).max() <= RTOL_RUN * np.abs(z["ux_plane2_final"]).max()
        ref = z["macro_final_fluidsum"]
        for q in range(4):
            assert abs(sums[q] - ref[q]) <= RTOL_RUN * abs(ref[1]), (q, sums[q], ref[q])
        assert np.abs(sq - z["macro_final_fluidsq"]).max() <= RTOL_RUN * np.abs(z["macro_final_fluidsq"]).max()


def _refined_vs_fp64(p, img, what):
    t = time.perf_counter()
    with Solver(p, img, precision="fp64") as a:
        r64 = a.solve()
    t64 = time.perf_counter() - t
    t = time.perf_counter()
    with Solver(p, img, precision="fp32") as b:
        r32 = b.solve()
    t32 = time.perf_counter() - t
    k64, k32 = r64["permeability"], r32["permeability"]
    print("%s: %d steps (%d FP32), K %.15e vs FP64 %.15e (rel %.2e); %.1f s vs %.1f s" % (
        what, r32["steps"], r32["steps32"], k32, k64, abs(k32 - k64) / abs(k64), t32, t64), flush=True)
    assert r64["steps"] == r32["steps"] == p.max_step
    assert r32["steps32"] > 0
    assert abs(k32 - k64) <= 1e-8 * abs(k64), (what, k32, k64)


def test_c2_spheres200_refined_matches_fp64():
    p = _c2_params(wall_ext_order=1, max_step=300)
    _refined_vs_fp64(p, _c2_image(), "200^3 sphere pack, Kn 1, wallExtOrder 1")


_C3_IMG = {}


def _c3_image():
    if "img" not in _C3_IMG:
        _C3_IMG["img"] = images.gaussian_field(300, 0.20, 6.0, seed=2024)
    return _C3_IMG["img"]


def _c3_params(kn, **kw):
    p = Params(nx=300, ny=300, nz=300, fluid_layer=4, wall_ext_order=1, nc_fundamental=4, half_range=True,
               chk_converge_step=10, max_step=200, eps=0.0, kn=kn, press_drop=1.0, accom=1.0)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


@pytest.mark.parametrize("kn", [0.01, 1.0, 10.0])
def test_c3_grf300_refined_matches_fp64(kn):
    _refined_vs_fp64(_c3_params(kn), _c3_image(), "300^3 random field, Kn %g" % kn)


def test_c3_grf300_permeability_independent_of_pressure_drop():
    img = _c3_image()
    ks = []
    for pd in (1.0, 0.25):
        with Solver(_c3_params(1.0, press_drop=pd, wall_ext_order=2), img, precision="fp64") as s:
            s.iterate(20)
            ks.append(s.chk_converge()[1])
    print("K at pressDrop 1: %.17e, at 0.25: %.17e" % tuple(ks), flush=True)
    assert abs(ks[0] - ks[1]) <= 1e-12 * abs(ks[0])


def _serve_mix():
    """Requests of the serving stream's distribution (bench.serve_specs),
    chosen to cover every cube side and every velocity grid, 12 in all."""
    import bench
    specs = bench.serve_specs(64, 7, 30, 0.0)
    chosen, seen = [], set()
    for q, (p, phi, s) in enumerate(specs):
        key = (p.nx, p.nc_fundamental)
        if key not in seen:
            seen.add(key)
            chosen.append(q)
    chosen = sorted(chosen[:12])
    return [(specs[q][0], bench.serve_image(*specs[q])) for q in chosen]


def test_c4_serving_mix_equals_alone():
    mix = _serve_mix()
    sides = {p.nx for p, _ in mix}
    grids = {p.nc_fundamental for p, _ in mix}
    print("mix:", [(p.nx, p.nc_fundamental, round(p.kn, 4)) for p, _ in mix], flush=True)
    assert sides == {32, 48, 64, 96, 128} and grids == {3, 4, 6}
    alone = []
    for p, img in mix:
        with Solver(p, img, precision="fp64") as s:
            alone.append(s.solve())
    with Server(max_slots=4, prep_threads=4) as srv:
        tickets = [srv.submit(p, img, precision="fp64") for p, img in mix]
        got = [srv.wait(t) for t in tickets]
    for (p, _), a, b in zip(mix, alone, got):
        assert b["status"] == REQ_DONE, (p.nx, p.nc_fundamental, b)
        assert b["steps"] == a["steps"] == p.max_step
        assert b["permeability"] == a["permeability"], (p.nx, p.nc_fundamental, a, b)
        assert b["mass"] == a["mass"] and b["error"] == a["error"]


def test_c5_grf160_32kn_each_rhs_equals_single_solve():
    img = images.gaussian_field(160, 0.25, 5.0, seed=2024)
    p = Params(nx=160, ny=160, nz=160, fluid_layer=4, wall_ext_order=2, nc_fundamental=4, half_range=True,
               chk_converge_step=10, max_step=3, eps=0.0, kn=1.0, press_drop=1.0, accom=1.0)
    kns = list(np.logspace(-3, 1, 32))
    steps = 3
    group = 8
    for g0 in range(0, len(kns), group):
        ks = kns[g0:g0 + group]
        with MultiSolver(p, img, ks, precision="fp64") as ms:
            ms.iterate(steps)
            multi = [np.stack(ms.macro_rhs(r)) for r in range(len(ks))]
            m_all, k_all, e_all = ms.chk_converge_all()
        for r, kn in enumerate(ks):
            p.kn = kn
            with Solver(p, img, precision="fp64") as s:
                s.iterate(steps)
                single = np.stack(s.macro())
                m, k, e = s.chk_converge()
            assert np.array_equal(multi[r], single), kn
            assert (m_all[r], k_all[r], e_all[r]) == (m, k, e), kn
        print("Kn %.3g..%.3g: %d right-hand sides equal their single solves" % (ks[0], ks[-1], len(ks)),
              flush=True)
