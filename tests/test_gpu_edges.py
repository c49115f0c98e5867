"""HIP path vs the oracle on cases the golden geometries leave out.

* velocity grids whose per-octant count is not one wave: 27 (the exact 6^3
  rule), 216 (12^3: three full passes of 64 lanes and one of 24) and 512
  (16^3: eight passes), in exact mode (bit for bit) and fast mode (1e-12);
* images at the edges of the domain's behaviour: an x-blocked image (a solid
  plane across the flow: no mass flow at steady state), a fully fluid image, the
  smallest image the reference runs (3 x 3 x 3, with fluid layers), and a
  ragged image whose sides are all different;
* the FP32-refined precision on a multi-pass velocity grid.
The oracle is the C restatement pinned bit for bit to the reference's dumps
(test_oracle.py); here it runs the same seeded inputs.
"""
import numpy as np
import pytest

from conftest import interior_mask

pytestmark = pytest.mark.gpu

RTOL_FAST = 1e-12


def _params(nx, ny, nz, nc, exact_quadrature=False, fluid_layer=2, wall_ext_order=2, steps=3,
            kn=0.5, accom=1.0):
    from piks3d_amd.params import Params
    return Params(image_file_name="img.raw", nx=nx, ny=ny, nz=nz, wall_ext_order=wall_ext_order,
                  fluid_layer=fluid_layer, nc_fundamental=nc, half_range=True,
                  exact_quadrature=exact_quadrature, chk_converge_step=1, max_step=steps, kn=kn,
                  press_drop=1.0, accom=accom, eps=0.0)


def _boxes(nx, ny, nz, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    img = np.zeros((nz, ny, nx), dtype=np.uint8)
    for _ in range(4):
        x0, y0, z0 = rng.integers(0, nx - 2), rng.integers(0, ny - 2), rng.integers(0, nz - 2)
        img[z0:z0 + rng.integers(1, 4), y0:y0 + rng.integers(1, 4), x0:x0 + rng.integers(1, 4)] = 1
    return img


def _compare(p, img, exact, steps):
    from oracle import Oracle
    from piks3d_amd import Solver
    s = Solver(p, img, precision="fp64", exact_order=exact)
    o = Oracle(p.as_oracle_dict(), img)
    inner = interior_mask(s.info.nx, s.info.ny, s.info.nz)
    try:
        assert np.array_equal(s.labels(), o.image())
        for it in range(1, steps + 1):
            s.iterate(1)
            o.iterate(1)
            fg, fo = s.f()[:, :, inner], o.f()[:, :, inner]
            mg, mo = s.macro(), o.macro()
            if exact:
                assert np.array_equal(fg, fo), "iter %d: f differs" % it
                assert np.array_equal(s.fw(), o.fw()), "iter %d: fw differs" % it
                for q in range(4):
                    assert np.array_equal(mg[q][inner], mo[q][inner]), "iter %d: moment %d" % (it, q)
                assert s.chk_converge() == o.chk_converge()
            else:
                scale = max(np.abs(fo).max(), 1e-300)
                assert np.abs(fg - fo).max() / scale <= RTOL_FAST, "iter %d: f" % it
                # a velocity moment that vanishes by symmetry is compared on
                # the scale of the density moment (round-off of tree sums)
                rho_scale = np.abs(mo[0][inner]).max()
                for q in range(4):
                    sc = max(np.abs(mo[q][inner]).max(), rho_scale, 1e-300)
                    assert np.abs(mg[q][inner] - mo[q][inner]).max() / sc <= RTOL_FAST, "iter %d: moment %d" % (it, q)
                a, b = s.chk_converge(), o.chk_converge()
                assert abs(a[1] - b[1]) <= RTOL_FAST * max(abs(b[1]), 1e-300)
        return s.chk_converge()
    finally:
        s.close()


@pytest.mark.parametrize("nc,exact_quadrature", [(3, True), (6, False), (8, False)])
@pytest.mark.parametrize("exact", [True, False])
def test_multi_pass_velocity_grids(nc, exact_quadrature, exact):
    print("case nc=%d exact_order=%d" % (nc, exact), flush=True)
    p = _params(9, 10, 11, nc, exact_quadrature=exact_quadrature, steps=2 if nc == 8 else 3)
    _compare(p, _boxes(9, 10, 11, seed=nc), exact, p.max_step)


@pytest.mark.parametrize("exact", [True, False])
def test_blocked_image(exact):
    """A two-voxel solid plane across x (a one-voxel plane is thinner than
    the reference keeps as a wall): parity over the first iterations, then no
    net mass flow once the transient has decayed."""
    print("case blocked exact_order=%d" % exact, flush=True)
    from piks3d_amd import Solver
    img = _boxes(10, 9, 8, seed=5)
    img[:, :, 4:6] = 1
    p = _params(10, 9, 8, 2, steps=4)
    _compare(p, img, exact, p.max_step)
    if not exact:
        with Solver(p, img, precision="fp64") as s:
            s.iterate(500)
            perm = s.chk_converge()[1]
        assert abs(perm) < 1e-12, perm


@pytest.mark.parametrize("exact", [True, False])
def test_fully_fluid_image(exact):
    print("case fluid exact_order=%d" % exact, flush=True)
    p = _params(8, 7, 6, 2, steps=3)
    _compare(p, np.zeros((6, 7, 8), dtype=np.uint8), exact, p.max_step)


@pytest.mark.parametrize("exact", [True, False])
def test_smallest_image(exact):
    print("case 3x3x3 exact_order=%d" % exact, flush=True)
    img = np.zeros((3, 3, 3), dtype=np.uint8)
    img[1, 1, 1] = 1
    p = _params(3, 3, 3, 2, steps=3)
    _compare(p, img, exact, p.max_step)


def test_ragged_sides_wall_order_3():
    print("case ragged order 3", flush=True)
    p = _params(13, 7, 5, 2, wall_ext_order=3, accom=0.7, steps=3)
    _compare(p, _boxes(13, 7, 5, seed=11), True, p.max_step)


def test_refined_precision_on_a_multi_pass_grid(monkeypatch):
    """Defect correction on a 216-velocity-per-octant grid (blocks of 3: one
    FP64 iteration, two FP32 corrections): after the same number of
    iterations the moments stay within single-precision distance of the FP64
    path, and switching to FP64 mid-block folds the correction and leaves the
    iteration running."""
    print("case refined nc=6", flush=True)
    from piks3d_amd import PIKS_PREC_FP64, Solver
    p = _params(9, 10, 11, 6, steps=6)
    img = _boxes(9, 10, 11, seed=6)
    monkeypatch.setenv("PIKS_DC_BLOCK", "3")
    with Solver(p, img, precision="fp64") as a, Solver(p, img, precision="fp32") as b:
        a.iterate(4)
        b.iterate(4)
        ra, rb = np.stack(a.macro()), np.stack(b.macro())
        assert np.abs(ra - rb).max() <= 1e-5 * max(np.abs(ra).max(), 1e-300)
        b.set_precision(PIKS_PREC_FP64)
        a.iterate(2)
        b.iterate(2)
        ka, kb = a.chk_converge()[1], b.chk_converge()[1]
        assert abs(ka - kb) <= 1e-5 * abs(ka)


def _cube_lattice(n, period=5, size=3, off=2):
    img = np.zeros((n, n, n), dtype=np.uint8)
    for a in range(off, n - size, period):
        for b in range(off, n - size, period):
            for c in range(off, n - size, period):
                img[a:a + size, b:b + size, c:c + size] = 1
    return img


def test_more_walls_than_one_launch_grid():
    """316 342 walls (faces, edges and 3-fold corners of a lattice of solid
    cubes): more than the 65 536 blocks x 4 waves one capped wall launch
    covers, so every wall kernel must stride over its list.  Bit for bit
    against the oracle after two iterations."""
    print("case cube lattice 120^3", flush=True)
    p = _params(120, 120, 120, 2, steps=2)
    img = _cube_lattice(120)
    from piks3d_amd import Solver
    with Solver(p, img, precision="fp64") as s:
        assert s.info.nwall > 65536 * 4
    _compare(p, img, True, 2)


def test_exact_order_moments_beyond_one_launch_grid():
    """A fully fluid 168^3 image has 4.9 M fluid cells, more than the 65 536
    blocks x 64 threads of one capped launch of the serial-order moments
    kernel: exact and fast modes must agree on every cell's moments."""
    print("case fluid 168^3 exact vs fast", flush=True)
    from piks3d_amd import Solver
    p = _params(168, 168, 168, 2, steps=2)
    img = np.zeros((168, 168, 168), dtype=np.uint8)
    out = []
    for exact in (True, False):
        with Solver(p, img, precision="fp64", exact_order=exact) as s:
            assert s.info.nfluid > 65536 * 64
            s.iterate(2)
            out.append(np.stack(s.macro()))
    scale = np.abs(out[1]).max()
    assert np.abs(out[0] - out[1]).max() <= 1e-12 * scale
