import glob
import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)
GOLDEN = os.path.join(REPO, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


def golden_names(prefix=""):
    """Golden cases (inputs plus the reference's state).  The outputs_* files
    hold the reference's field files instead, and the compact fixtures of
    large runs (an image hash instead of the image) only check lines and
    summaries: both are loaded by name."""
    names = []
    for p in sorted(glob.glob(os.path.join(GOLDEN, prefix + "*.npz"))):
        if os.path.basename(p).startswith("outputs_"):
            continue
        with np.load(p) as z:
            if "image_sha256" in z.files:
                continue
        names.append(os.path.basename(p)[:-4])
    return names


def load_golden(name):
    return np.load(os.path.join(GOLDEN, name + ".npz"))


def params_from_golden(z):
    """Params of a golden case (the fixture stores para.in's values)."""
    from piks3d_amd.params import Params
    ip, fp = z["params"], z["fparams"]
    return Params(image_file_name="img.raw", nx=int(z["nx"]), ny=int(z["ny"]), nz=int(z["nz"]),
                  wall_ext_order=int(ip[0]), fluid_layer=int(ip[1]), nc_fundamental=int(ip[2]),
                  half_range=bool(ip[3]), chk_converge_step=int(ip[4]), max_step=int(ip[5]),
                  ref_l=float(fp[0]), kn=float(fp[1]), press_drop=float(fp[2]),
                  accom=float(fp[3]), eps=float(fp[4]))


def interior_mask(Nx, Ny, Nz):
    """Cells with 1 <= i,j,k <= N in the ghosted grid (reference cell order)."""
    m = np.zeros((Nz + 4, Ny + 4, Nx + 4), dtype=bool)
    m[2:-2, 2:-2, 2:-2] = True
    return m.ravel()
