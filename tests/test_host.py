"""Host-side pieces that need no GPU: the C ABI library and its exports, the
product's image preprocessing against the oracle, para.in, image generators."""
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import REPO, golden_names, load_golden, params_from_golden


def test_library_exports_every_declared_entry_point():
    import piks3d_amd
    header = open(os.path.join(REPO, "include", "piks3d.h")).read()
    declared = set(re.findall(r"^\s*(?:int|const char\*)\s+(piks_\w+)\s*\(", header, re.M))
    assert len(declared) >= 12
    lib = piks3d_amd.lib()
    for name in declared:
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", piks3d_amd.LIB_PATH], capture_output=True,
                         text=True).stdout
    exported = set(re.findall(r"\bT (piks_\w+)", out))
    assert declared <= exported
    assert set(piks3d_amd._API) == declared


@pytest.mark.parametrize("name", [n for n in golden_names() if not n.startswith(("vgrid", "tutorial"))])
def test_product_preprocessing_matches_oracle(name):
    from oracle import Oracle
    from piks3d_amd import preprocess
    z = load_golden(name)
    p = params_from_golden(z)
    labels, info = preprocess(p, z["image"])
    assert np.array_equal(labels, z["label"].astype(np.int32))
    o = Oracle(p.as_oracle_dict(), z["image"])
    assert (info.nwall, info.ncorner, info.nfluid) == (o.nwall, o.ncorner, o.nfluid)


def test_preprocess_rejects_bad_bytes_and_x_boundary_corners():
    from piks3d_amd import PiksError, preprocess
    from piks3d_amd.params import Params
    p = Params(nx=6, ny=5, nz=5, nc_fundamental=2, kn=1.0, press_drop=0.1, fluid_layer=2)
    img = np.zeros((5, 5, 6), np.uint8)
    img[2, 2, 3] = 7
    with pytest.raises(PiksError) as e:
        preprocess(p, img)
    assert e.value.code == -2
    with pytest.raises(PiksError) as e:
        preprocess(p, np.zeros(10, np.uint8))
    assert e.value.code == -2
    # a block touching x = 1 with fluidLayer 0 makes 3-fold corners in the
    # first two x layers: the reference aborts in MPI, the library refuses it
    p0 = Params(nx=8, ny=8, nz=8, nc_fundamental=2, kn=1.0, press_drop=0.1, fluid_layer=0)
    img = np.zeros((8, 8, 8), np.uint8)
    img[3:6, 3:6, 1:4] = 1
    with pytest.raises(PiksError) as e:
        preprocess(p0, img)
    assert e.value.code == -4


def test_para_in_round_trip_and_reference_format():
    from piks3d_amd.params import format_para, parse_para
    text = """&physicalNml
imageFileName = 'singleSphere100^3.raw'
Nx = 100
Ny = 100
Nz = 100
wallExtOrder = 1
fluidLayer = 0
Ref_L = 1.d0
/

&velocityNml
Nc_fundamental = 4
halfRange = T
/

&mpiNml
mpi_xdim = 2
mpi_ydim = 2
mpi_zdim = 1
/

&solverNml
maxStep = 100
chkConvergeStep = 10
saveStep = 50
eps = 1.0d-7
saveLast = T
saveFormat = 1
/

&flowNml
Kn = 1.0d-1
pressDrop = 1.0d-1
accom = 1.d0
/
"""
    p = parse_para(text)
    assert (p.image_file_name, p.nx, p.nz, p.wall_ext_order, p.fluid_layer) == (
        "singleSphere100^3.raw", 100, 100, 1, 0)
    assert (p.nc_fundamental, p.half_range, p.max_step, p.chk_converge_step) == (4, True, 100, 10)
    assert (p.eps, p.kn, p.press_drop, p.accom) == (1e-7, 0.1, 0.1, 1.0)
    q = parse_para(format_para(p, gpu=True))
    for k in ("nx", "ny", "nz", "wall_ext_order", "fluid_layer", "ref_l", "nc_fundamental",
              "half_range", "max_step", "chk_converge_step", "eps", "kn", "press_drop", "accom"):
        assert getattr(q, k) == getattr(p, k), k
    # defaults of src/parameters.F90:91-98
    d = parse_para("&physicalNml\nNx=4\n/\n&velocityNml\n/\n&solverNml\n/\n&flowNml\n/\n")
    assert (d.wall_ext_order, d.fluid_layer, d.ref_l, d.save_last, d.save_format) == (2, 0, 1.0, True, 1)
    with pytest.raises(ValueError):
        parse_para(text.replace("mpi_zdim = 1", "mpi_zdim = 1\nblock_repx = 2"))


def test_image_generators_are_seeded_and_hit_their_porosity():
    from piks3d_amd import images
    a = images.sphere_pack(40, 0.4, 4, 10, seed=3)
    b = images.sphere_pack(40, 0.4, 4, 10, seed=3)
    assert np.array_equal(a, b) and 0.39 <= images.porosity(a) <= 0.40
    g = images.gaussian_field(32, 0.25, 4.0, seed=1, percolate_x=False)
    assert abs(images.porosity(g) - 0.25) < 1e-3
    s = images.centered_sphere(64, 28.0)
    assert int(s.sum()) == 92096  # solid voxels of the 64^3 configuration


def test_workload_images_match_their_committed_hashes():
    """Every benchmark image regenerates to the SHA-256 committed with the
    fixtures (tests/golden/make_workload_hashes.py wrote them)."""
    import hashlib
    import sys
    sys.path.insert(0, os.path.join(REPO, "tests", "golden"))
    from make_workload_hashes import workload_images
    want = {}
    with open(os.path.join(REPO, "tests", "golden", "workload_images.sha256")) as fh:
        for line in fh:
            h, name = line.split()
            want[name] = h
    got = {name: hashlib.sha256(img.tobytes()).hexdigest() for name, img in workload_images()}
    assert got == want
