"""The drop-in Fortran host (piks3d_gpu) against the reference program's own output.

The reference's stdout for every golden case is stored in its fixture; the
host, run on the same para.in and image, must print the same convergence-check
lines (src/output.F90:156: iStep, mass, permeability, error in ES15.6), write
Results.dat in the reference's format, and write a VTI field whose flags are
the reference's labels.  With exact_order = T its field files in all three
formats equal the reference program's own files byte for byte.
"""
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import REPO, load_golden, params_from_golden

pytestmark = pytest.mark.gpu

HOST = os.path.join(REPO, "piks3d_amd", "piks3d_gpu")
CHECK_LINE = re.compile(r"^\s+(\d+)\s+(\S+E[+-]\d+)\s+(\S+E[+-]\d+)\s+(\S+E[+-]\d+)\s*$")


def _check_lines(text):
    return [m.groups() for m in map(CHECK_LINE.match, text.splitlines()) if m]


@pytest.mark.parametrize("name", ["spheres40_n2", "boxes_n2_a1_e2", "grf40_n2"])
def test_fortran_host_prints_the_reference_lines(name, tmp_path):
    if not os.path.exists(HOST):
        pytest.fail("piks3d_gpu is not built (python -m piks3d_amd.build)")
    from piks3d_amd import images
    from piks3d_amd.params import format_para
    z = load_golden(name)
    p = params_from_golden(z)
    p.image_file_name = "img.raw"
    p.save_step = 1 << 30
    p.save_last = True
    p.save_format = 1
    images.write_raw(str(tmp_path / "img.raw"), z["image"])
    (tmp_path / "para.in").write_text(format_para(p))
    r = subprocess.run([HOST], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    got = _check_lines(r.stdout)
    want = _check_lines(str(z["stdout"]))
    assert len(want) > 0 and got == want
    res = (tmp_path / "Results.dat").read_text().splitlines()
    assert len(res) == len(want)
    for line, (step, mass, perm, err) in zip(res, want):
        f = line.split()
        assert f[1:4] == [mass, perm, err] and f[4] == step
    # eps is never reached in these cases: like the reference's DO loop, the
    # step counter ends one past maxStep
    assert re.search(r"Walltime=\s+\S+, Steps=\s+%d, K=\s+%s" % (p.max_step + 1, re.escape(want[-1][2])),
                     r.stdout)
    # VTI flags are the labels of the reference's image over the interior
    vti = (tmp_path / "Field_000.vti").read_text().splitlines()
    i0 = vti.index('<DataArray type="Int32" Name="flag" format="ascii">') + 1
    nx, ny, nz = int(z["nx"]) + 2 * p.fluid_layer, int(z["ny"]), int(z["nz"])
    flags = np.array([int(t) for t in vti[i0:i0 + nx * ny * nz]])
    lab = z["label"].reshape(nz + 4, ny + 4, nx + 4)[2:-2, 2:-2, 2:-2].ravel()
    assert np.array_equal(flags, lab)
    assert os.path.exists(tmp_path / "Field.pvti")


OUTPUTS = "outputs_boxes_n2_a1_e2"


@pytest.mark.parametrize("fmt", [1, 2, 3])
def test_fortran_host_writes_the_reference_files(fmt, tmp_path):
    """With exact_order = T every value equals the reference's bit for bit, so
    the field files (VTI + PVTI, Tecplot, legacy VTK) and Results.dat equal
    the reference program's own files byte for byte (tests/golden/
    make_outputs.py wrote those from the reference)."""
    print("case output format %d" % fmt, flush=True)
    if not os.path.exists(HOST):
        pytest.fail("piks3d_gpu is not built (python -m piks3d_amd.build)")
    from piks3d_amd import images
    from piks3d_amd.params import format_para
    want = load_golden(OUTPUTS)
    z = load_golden("boxes_n2_a1_e2")
    p = params_from_golden(z)
    p.image_file_name = "img.raw"
    p.save_step = 1 << 30
    p.save_last = True
    p.save_format = fmt
    p.exact_order = True
    images.write_raw(str(tmp_path / "img.raw"), z["image"])
    (tmp_path / "para.in").write_text(format_para(p, gpu=True))
    r = subprocess.run([HOST], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    names = [k.split("/", 1)[1] for k in want.files if k.startswith("%d/" % fmt)]
    assert names
    for n in names:
        got = (tmp_path / n).read_bytes()
        ref = bytes(want["%d/%s" % (fmt, n)])
        if got != ref:
            gl, rl = got.decode().splitlines(), ref.decode().splitlines()
            bad = next((q for q, (a, b) in enumerate(zip(gl, rl)) if a != b), min(len(gl), len(rl)))
            pytest.fail("%s differs from the reference's at line %d: %r vs %r" % (
                n, bad + 1, gl[bad] if bad < len(gl) else None, rl[bad] if bad < len(rl) else None))
