"""Write the golden fixtures in tests/golden/ from the reference itself.

The reference (iPACT-Platform/PIKS3D, Fortran) is compiled from its own
sources by oracle/Makefile into oracle/_ref/ together with a dump driver
(oracle/ref_dump_driver.F90) that runs the reference's modules unchanged and
writes their state in binary.  This script builds that, runs one case per
fixture on one MPI rank with OMP_NUM_THREADS=1 (the reference's wall loop is
order dependent when two walls face each other across a one-voxel pore, so a
single thread is its well-defined order), and saves inputs and outputs as
compressed .npz files.  Nothing from the reference's sources is stored: only
the case inputs and the reference's outputs.

Run once in the build container (needs /root/reference):
    python tests/golden/make_golden.py [case ...]
"""
from __future__ import annotations

import hashlib
import os
import shutil
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)

from piks3d_amd import images  # noqa: E402
from piks3d_amd.params import Params, format_para  # noqa: E402

ORACLE_DIR = os.path.join(REPO, "oracle")
DUMP = os.path.join(ORACLE_DIR, "_ref", "piks_ref_dump")
RUNS = os.path.join(ORACLE_DIR, "_ref", "runs")
TUTORIAL_RAW = "/root/reference/tutorials/cubicSpherePacking/singleSphere100^3.raw"


def box_case(shape, nboxes, smin, smax, seed):
    return images.box_blocks(shape, nboxes, smin, smax, seed)


# name -> (image factory, Params overrides, dump args, velocity subset or None)
CASES = {
    # every one of the 26 wall labels, 27 three-fold corners
    "boxes_n2_a1_e2": (lambda: box_case((10, 11, 12), 8, 2, 5, 1),
                       dict(fluid_layer=2, nc_fundamental=2, kn=0.5, press_drop=0.1, accom=1.0,
                            wall_ext_order=2, chk_converge_step=2, max_step=6),
                       ["f1", "f2", "m3", "m4", "m5", "f6"], None),
    "boxes_n2_a05_e1": (lambda: box_case((10, 11, 12), 8, 2, 5, 1),
                        dict(fluid_layer=2, nc_fundamental=2, kn=2.0, press_drop=0.05, accom=0.5,
                             wall_ext_order=1, chk_converge_step=3, max_step=6),
                        ["f2", "m4", "f6"], None),
    # fluidLayer = 0: two walls whose extrapolation reads another wall
    "coupled_fl0_a06": (lambda: box_case((9, 10, 11), 6, 2, 4, 7),
                        dict(fluid_layer=0, nc_fundamental=2, kn=0.3, press_drop=0.1, accom=0.6,
                             wall_ext_order=2, chk_converge_step=2, max_step=5),
                        ["f1", "m2", "m3", "m4", "f5"], None),
    # wallExtOrder = 3 with walls on the first-order side conditions
    "order3_fl0_a07": (lambda: box_case((9, 10, 11), 6, 2, 4, 172),
                       dict(fluid_layer=0, nc_fundamental=2, kn=0.1, press_drop=0.2, accom=0.7,
                            wall_ext_order=3, chk_converge_step=2, max_step=5),
                       ["f1", "m3", "f5"], None),
    # 8^3 velocities on walls of every kind, velocity subset stored
    "boxes_n4_a08": (lambda: box_case((8, 9, 10), 6, 2, 4, 1),
                     dict(fluid_layer=2, nc_fundamental=4, kn=1.0, press_drop=0.1, accom=0.8,
                          wall_ext_order=2, chk_converge_step=1, max_step=3),
                     ["m1", "m2", "f3"], [0, 21, 42, 63]),
    # full-range Gauss-Hermite grid
    "fullrange_n2": (lambda: box_case((8, 9, 10), 6, 2, 4, 1),
                     dict(fluid_layer=2, nc_fundamental=2, half_range=False, kn=0.7,
                          press_drop=0.1, accom=1.0, wall_ext_order=2, chk_converge_step=1,
                          max_step=3),
                     ["m1", "f3"], None),
    # the sphere-pack and Gaussian-random-field image families at small size
    "spheres40_n2": (lambda: images.sphere_pack(40, 0.40, 4.0, 10.0, seed=11),
                     dict(fluid_layer=4, nc_fundamental=2, kn=1.0, press_drop=1.0, accom=1.0,
                          wall_ext_order=2, chk_converge_step=5, max_step=30),
                     ["m30"], None),
    "grf40_n2": (lambda: images.gaussian_field(40, 0.25, 5.0, seed=5),
                 dict(fluid_layer=4, nc_fundamental=2, kn=0.05, press_drop=1.0, accom=1.0,
                      wall_ext_order=2, chk_converge_step=5, max_step=30),
                 ["m30"], None),
}

# velocity-grid tables for every n the reference supports
for _n in (2, 4, 6, 8, 12, 16):
    for _half in (True, False):
        CASES["vgrid_%s%d" % ("h" if _half else "f", _n)] = (
            lambda: np.zeros((3, 3, 3), dtype=np.uint8),
            dict(fluid_layer=0, nc_fundamental=_n, half_range=_half, kn=1.0, press_drop=0.1,
                 accom=1.0, wall_ext_order=2, chk_converge_step=1, max_step=1),
            [], "velocity")

# long runs (history of the convergence check only)
LONG_CASES = {
    # the CPU-runnable configuration: 64^3 centred sphere r = 28, 8^3 velocities, Kn 0.1
    "c1_sphere64": (lambda: images.centered_sphere(64, 28.0),
                    dict(fluid_layer=4, nc_fundamental=4, kn=0.1, press_drop=1.0, accom=1.0,
                         wall_ext_order=2, chk_converge_step=10, max_step=100000, eps=1e-6),
                    ["mfinal"], "history"),
    # the 200^3 sphere pack of the headline throughput workload (seed 2024),
    # 8^3 velocities, Kn 1, wallExtOrder 2: 30 iterations, checked every 10.
    # The image is regenerated from its seed (its SHA-256 is in the fixture)
    # and only the check lines, the final Ux plane x = 2 and the moment sums
    # are kept.
    "c2_spheres200": (lambda: images.sphere_pack(200, 0.40, 4.0, 10.0, seed=2024),
                      dict(fluid_layer=4, nc_fundamental=4, kn=1.0, press_drop=1.0, accom=1.0,
                           wall_ext_order=2, chk_converge_step=10, max_step=30, eps=0.0),
                      ["mfinal"], "history_compact"),
    # the reference's own tutorial (conf/createSphere.F90 image), 100 steps
    "tutorial100": (lambda: images.quarter_sphere(100, 100, 100, 38.0),
                    dict(fluid_layer=0, nc_fundamental=4, kn=0.1, press_drop=0.1, accom=1.0,
                         wall_ext_order=1, chk_converge_step=10, max_step=100, eps=1e-7),
                    ["mfinal"], "history"),
}


def run_case(name, factory, over, dump_args, subset, threads=1):
    img = factory()
    nz, ny, nx = img.shape
    fl = over.get("fluid_layer", 0)
    p = Params(image_file_name="img.raw", nx=nx, ny=ny, nz=nz, eps=over.pop("eps", 1e-30))
    for k, v in over.items():
        setattr(p, k, v)
    p.save_step = 1 << 30
    p.save_last = False
    d = os.path.join(RUNS, name)
    shutil.rmtree(d, ignore_errors=True)
    os.makedirs(os.path.join(d, "out"))
    images.write_raw(os.path.join(d, "img.raw"), img)
    with open(os.path.join(d, "para.in"), "w") as fh:
        fh.write(format_para(p))
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    r = subprocess.run([DUMP, "out"] + dump_args, cwd=d, env=env, capture_output=True, text=True)
    if r.returncode != 0 or "DUMP_DONE" not in r.stdout:
        raise RuntimeError("reference run %s failed:\n%s\n%s" % (name, r.stdout[-2000:], r.stderr[-2000:]))
    out = os.path.join(d, "out")

    def rd(fn, dt=np.float64):
        return np.fromfile(os.path.join(out, fn), dtype=dt)

    sc = {}
    for line in open(os.path.join(out, "setup_scalars.txt")):
        parts = line.split()
        sc[parts[0]] = parts[1:]
    Ntotal = int(sc["grid"][6])
    Nc8 = int(sc["nc"][1])
    ncorner = int(sc["counts"][2])
    compact = subset == "history_compact"
    if compact:
        subset = "history"
    fx = dict(
        nx=nx, ny=ny, nz=nz,
        params=np.array([p.wall_ext_order, p.fluid_layer, p.nc_fundamental, int(p.half_range),
                         p.chk_converge_step, p.max_step], dtype=np.int64),
        fparams=np.array([p.ref_l, p.kn, p.press_drop, p.accom, p.eps]),
        cx=rd("setup_cx.bin"), cy=rd("setup_cy.bin"), cz=rd("setup_cz.bin"), w=rd("setup_w.bin"),
        diff_flux=float(sc["DiffFlux"][0].replace("E", "e")), ds=float(sc["ds"][0]),
        mu=float(sc["mu"][0]),
        counts=np.array([int(v) for v in sc["counts"]] + [int(v) for v in sc["nstencil"]],
                        dtype=np.int64),
        stdout=r.stdout,
    )
    if compact:
        fx["image_sha256"] = hashlib.sha256(np.ascontiguousarray(img).tobytes()).hexdigest()
    else:
        fx["image"] = img
    label = rd("setup_image.bin", np.int32).astype(np.int8) if subset != "velocity" else None
    if compact:
        fx["label_sha256"] = hashlib.sha256(label.tobytes()).hexdigest()
    elif subset != "velocity":
        fx.update(label=label,
                  which_corner=rd("setup_which_corner.bin", np.int32),
                  vecwall=rd("setup_vecWall.bin", np.int32),
                  extcoef=rd("setup_extCoef.bin"),
                  mapf=rd("setup_mapF.bin", np.int32))
    if subset != "velocity":
        conv = np.fromfile(os.path.join(out, "conv.bin"),
                           dtype=np.dtype([("step", "<i4"), ("mass", "<f8"), ("perm", "<f8"),
                                           ("error", "<f8")]))
        fx["conv_step"] = conv["step"].astype(np.int64)
        fx["conv"] = np.stack([conv["mass"], conv["perm"], conv["error"]], axis=1) if len(conv) else np.zeros((0, 3))
        for a in dump_args:
            tag = a[1:] if a[0] in "fm" and a not in ("ffinal", "mfinal") else "final"
            mac = np.stack([rd("%s_%s.bin" % (q, tag)) for q in ("rho", "ux", "uy", "uz")])
            if subset == "history":
                # interior fluid cells only, as float64
                fluid = label == 0
                fx["macro_%s_fluidsum" % tag] = mac[:, fluid].sum(axis=1)
                fx["macro_%s_fluidsq" % tag] = (mac[:, fluid] ** 2).sum(axis=1)
                Nxt = nx + 2 * fl + 4
                Nyt = ny + 4
                ux = mac[1].reshape(nz + 4, Nyt, Nxt)
                fx["ux_plane2_%s" % tag] = ux[2:-2, 2:-2, 3]    # x = 2 (column = 2)
                continue
            fx["macro_%s" % tag] = mac
            if a[0] == "f" or a == "ffinal":
                f = rd("f_%s.bin" % tag).reshape(8, Nc8, Ntotal)
                if ncorner > 0:
                    fw = rd("fw_%s.bin" % tag).reshape(8, Nc8, ncorner)
                else:
                    fw = np.zeros((8, Nc8, 0))
                if subset is not None:
                    f = f[:, subset, :]
                    fw = fw[:, subset, :]
                    fx["vsubset"] = np.array(subset, dtype=np.int64)
                fx["f_%s" % tag] = f
                fx["fw_%s" % tag] = fw
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **fx)
    shutil.rmtree(d, ignore_errors=True)
    return fx


def main(argv):
    subprocess.run(["make", "-s", "-C", ORACLE_DIR, "ref"], check=True)
    names = argv or list(CASES)
    for name in names:
        if name in CASES:
            factory, over, dump_args, subset = CASES[name]
            run_case(name, factory, dict(over), list(dump_args), subset)
        elif name in LONG_CASES:
            factory, over, dump_args, subset = LONG_CASES[name]
            # the sphere packs have walls that read other walls: one thread
            # gives the reference's serial wall order
            threads = 1 if name == "c2_spheres200" else int(os.environ.get("GOLDEN_THREADS", "8"))
            run_case(name, factory, dict(over), list(dump_args), subset, threads=threads)
        else:
            raise SystemExit("unknown case " + name)
        print("wrote", name, flush=True)
    if os.path.exists(TUTORIAL_RAW):
        h = hashlib.sha256(open(TUTORIAL_RAW, "rb").read()).hexdigest()
        with open(os.path.join(HERE, "tutorial_image.sha256"), "w") as fh:
            fh.write(h + "  singleSphere100^3.raw\n")


if __name__ == "__main__":
    main(sys.argv[1:])
