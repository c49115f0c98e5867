"""FP64 defect correction with FP32 correction iterations (precision
'fp32', PIKS_PREC_FP32_REFINED) against the plain FP64 solve.

The reference's iteration is an affine map f <- T(f) (src/solver.F90:111-1940
with the pressure-drop source of src/solver.F90:1832-1862).  A refined
context runs blocks: one FP64 iteration of the state S that also yields the
residual r = T(S) - S, then FP32 iterations of the correction
e <- A(e + r) with the linear part A; S + e + r is then, in exact arithmetic,
the plain iteration's state, so FP32 rounds only the correction.  The bounds
stated here: at equal iteration counts the permeability is within 1e-8 of
the plain FP64 solve's, and the state within single-precision distance of
the correction (far below single-precision distance of the state).  With a
block length of 1 the refined context is the FP64 iteration bit for bit.
"""
import numpy as np
import pytest

from piks3d_amd import PiksError, Solver, images
from piks3d_amd.params import Params

pytestmark = pytest.mark.gpu


def _pack(n=40, kn=1.0, seed=5):
    img = images.sphere_pack(n, 0.40, 3.0, 6.0, seed=seed)
    p = Params(nx=n, ny=n, nz=n, fluid_layer=2, wall_ext_order=2, nc_fundamental=4,
               half_range=True, chk_converge_step=10, max_step=20000, eps=1e-10,
               kn=kn, press_drop=1.0, accom=1.0)
    return p, img


@pytest.mark.parametrize("block", [2, 5, 12])
def test_correction_blocks_track_fp64(block, monkeypatch):
    # fixed block lengths from the first iteration on: the correction is then
    # as large as the state itself (the worst case for FP32 rounding)
    p, img = _pack()
    monkeypatch.setenv("PIKS_DC_BLOCK", str(block))
    with Solver(p, img, precision="fp64") as a, Solver(p, img, precision="fp32") as b:
        assert b.get_info().dc_block == block
        for n in (block, 3, 2 * block + 1):
            a.iterate(n)
            b.iterate(n)
            ra, rb = np.stack(a.macro()), np.stack(b.macro())
            scale = np.abs(ra).max()
            assert scale > 0 and np.abs(ra - rb).max() <= 1e-6 * scale
        fa, fb = a.f(), b.f()
        assert np.abs(fa - fb).max() <= 1e-6 * np.abs(fa).max()
        # FP32 rounding really happened
        assert not np.array_equal(fa, fb)
        assert b.get_info().steps_fp32 > 0


def test_block_length_one_is_the_fp64_iteration(monkeypatch):
    p, img = _pack(n=24)
    monkeypatch.setenv("PIKS_DC_BLOCK", "1")
    with Solver(p, img, precision="fp64") as a, Solver(p, img, precision="fp32") as b:
        for _ in range(3):
            a.iterate(4)
            b.iterate(4)
            assert a.chk_converge() == b.chk_converge()
        assert np.array_equal(a.f(), b.f())
        assert b.get_info().steps_fp32 == 0


@pytest.mark.parametrize("kn", [1.0, 10.0])
def test_refined_solve_reaches_the_fp64_permeability(kn):
    # The reference's iteration has a slowly drifting mode on this image (its
    # check error stops falling near 1e-7 and K keeps creeping in the 6th
    # digit), so both solves run the same number of iterations (eps = 0) and
    # are compared there, at every check.
    p, img = _pack(kn=kn)
    p.eps = 0.0
    p.max_step = 600
    with Solver(p, img, precision="fp64") as a:
        r64 = a.solve()
    with Solver(p, img, precision="fp32") as b:
        r32 = b.solve()
        assert b.info.precision == 1
    print("kn %g: %d steps (%d FP32 corrections), K %.12e vs %.12e" % (
        kn, r32["steps"], r32["steps32"], r32["permeability"], r64["permeability"]))
    assert r64["steps"] == r32["steps"] == p.max_step
    assert r32["steps32"] >= 0.5 * r32["steps"]
    for (s64, m64, k64, e64), (s32, m32, k32, e32) in zip(r64["history"], r32["history"]):
        assert s64 == s32
        assert abs(k32 - k64) <= 1e-8 * abs(k64), (s64, k32, k64)


def test_precision_switch_rules(monkeypatch):
    p, img = _pack(n=24)
    with Solver(p, img, precision="fp64") as a:
        with pytest.raises(PiksError) as e:
            a.set_precision("fp32")
        assert e.value.code == -7
        with pytest.raises(PiksError) as e:
            a.set_precision(5)
        assert e.value.code == -1
        a.set_precision("fp64")                  # no-op
    monkeypatch.setenv("PIKS_DC_BLOCK", "8")
    with Solver(p, img, precision="fp32") as b, Solver(p, img, precision="fp64") as a:
        assert b.info.precision == 1
        a.iterate(22)
        b.iterate(22)                            # blocks of 8: corrections pending after 22
        b.set_precision("fp64")                  # folds them into the FP64 state
        assert b.info.precision == 0
        fa, fb = a.f(), b.f()
        assert np.abs(fa - fb).max() <= 1e-6 * np.abs(fa).max()
        b.set_precision("fp32")
        assert np.array_equal(b.f(), fb)         # switching back changes nothing stored


def test_correction_iterations_replayed_from_graphs(monkeypatch):
    # every iteration kind of a block (FP64 with and without folding, first
    # and later correction) replayed from its captured graph gives the
    # launched iterations' bits; one iteration per call with the check (its
    # own kernel and a synchronisation) between replays
    p, img = _pack(n=24)
    monkeypatch.setenv("PIKS_DC_BLOCK", "4")
    runs = []
    for graph in ("0", "1"):
        monkeypatch.setenv("PIKS_GRAPH", graph)
        with Solver(p, img, precision="fp32") as s:
            lines = []
            for _ in range(11):
                s.iterate(1)
                lines.append(s.mass_flow())
            s.set_precision("fp64")
            runs.append((lines, s.f()))
    assert runs[0][0] == runs[1][0]
    assert np.array_equal(runs[0][1], runs[1][1])


@pytest.mark.parametrize("env", [("PIKS_SWEEP_NT", "2"), ("PIKS_SWEEP_NT", "4"), ("PIKS_OCT_GROUPS", "1"),
                                 ("PIKS_OCT_GROUPS", "8"), ("PIKS_DC_ACTIVE", "0")])
def test_sweep_launch_shapes_keep_the_correction_bits(env, monkeypatch):
    # the FP64 block iterations and the FP32 corrections with several sweep
    # tasks per wave, the octants swept in groups, or the correction's
    # non-swept slots walked in full instead of only those with a nonzero
    # residual, give the default bits
    p, img = _pack(n=24)
    monkeypatch.setenv("PIKS_DC_BLOCK", "4")
    runs = []
    for v in (None, env[1]):
        if v is None:
            monkeypatch.delenv(env[0], raising=False)
        else:
            monkeypatch.setenv(env[0], v)
        with Solver(p, img, precision="fp32") as s:
            s.iterate(9)
            lines = s.chk_converge()
            s.set_precision("fp64")
            runs.append((lines, s.f()))
    assert runs[0][0] == runs[1][0]
    assert np.array_equal(runs[0][1], runs[1][1])
