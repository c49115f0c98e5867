"""Host checks of the single-launch sweep's records (build_flow_records in
piks3d_amd/csrc/geometry.cpp), compiled from the library's own geometry
sources with g++ and run on three images: a 40^3 sphere pack, the 64^3
centred sphere of the CPU-runnable configuration and a 64^3 pack of the
headline workload's kind.  tests/flow_records_check.cpp lists the checks:
waits on the nearest swept upwind neighbour per axis, every swept source
updated by an earlier task of the octant, every switched read addressed as
the hyperplane kernels' switched_source addresses it (src/solver.F90:184-204
and analogues decide those switches), and the static split over 8, 64 and
768 waves per octant drains, also when the records go strip by strip
(bands of y, one after another).  The same program with mistakes planted in the
records must fail."""
import os
import shutil
import subprocess

import pytest

from piks3d_amd import images

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(REPO, "piks3d_amd", "csrc")

CASES = [
    ("pack40", lambda: images.sphere_pack(40, 0.40, 3.0, 6.0, seed=5), 2),
    ("sphere64", lambda: images.centered_sphere(64, 28.0), 4),
    ("pack64", lambda: images.sphere_pack(64, 0.40, 4.0, 10.0, seed=2024), 4),
]


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no g++")
    out = tmp_path_factory.mktemp("flow_records") / "flow_records_check"
    subprocess.run([gxx, "-O2", "-std=c++17", "-I" + CSRC, "-I" + os.path.join(REPO, "include"), "-o", str(out),
                    os.path.join(HERE, "flow_records_check.cpp"), os.path.join(CSRC, "geometry.cpp"),
                    "-lpthread"], check=True)
    return str(out)


def _run(checker, tmp_path, name, make, layers, *extra):
    img = make()
    path = str(tmp_path / (name + ".raw"))
    images.write_raw(path, img)
    nx, ny, nz = img.shape[2], img.shape[1], img.shape[0]
    return subprocess.run([checker, path, str(nx), str(ny), str(nz), str(layers), "2", *extra],
                          capture_output=True, text=True, timeout=300)


@pytest.mark.parametrize("strips", [1, 3, 64])
@pytest.mark.parametrize("name,make,layers", CASES, ids=[c[0] for c in CASES])
def test_flow_records(checker, tmp_path, name, make, layers, strips):
    # strips > 1: each octant's tasks band of y by band of y (64: more bands
    # than the 40^3 image has rows, one row per band)
    r = _run(checker, tmp_path, name, make, layers, "strips=%d" % strips)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "bad 0" in r.stdout
    assert "switched reads 0 " not in r.stdout   # the images exercise switched reads


def test_planted_mistakes_are_found(checker, tmp_path):
    name, make, layers = CASES[0]
    r = _run(checker, tmp_path, name, make, layers, "plant")
    print(r.stdout)
    assert "planted: wait mark 1, switched read 1" in r.stdout
    assert r.returncode == 1
    assert "FAIL wait mark" in r.stdout and "FAIL switched f source" in r.stdout
