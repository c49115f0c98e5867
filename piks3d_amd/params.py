"""para.in: the solver's configuration surface.

Mirrors src/parameters.F90:78-136 of the reference: five Fortran namelists
read in order from ``para.in``

    &physicalNml  imageFileName, Nx, Ny, Nz, wallExtOrder, fluidLayer, Ref_L
    &velocityNml  Nc_fundamental, halfRange
    &mpiNml       mpi_xdim, mpi_ydim, mpi_zdim, block_repx, block_repy, block_repz
    &solverNml    maxStep, chkConvergeStep, saveStep, eps, saveLast, saveFormat
    &flowNml      Kn, pressDrop, accom

with the reference's defaults (wallExtOrder = 2, fluidLayer = 0, Ref_L = 1,
saveLast = T, saveFormat = 1, block_rep* = 1; src/parameters.F90:91-98).  This
build adds one optional namelist the reference ignores:

    &gpuNml       precision ('fp64' | 'fp32'), device, exact_quadrature, exact_order,
                  refineSwitch

The mpi* values are accepted and ignored (one GPU solves the whole domain);
block_rep* other than 1 (the reference's weak-scaling replication) is refused.
"""
from __future__ import annotations

import dataclasses
import re


@dataclasses.dataclass
class Params:
    image_file_name: str = ""
    nx: int = 0
    ny: int = 0
    nz: int = 0
    wall_ext_order: int = 2
    fluid_layer: int = 0
    ref_l: float = 1.0
    nc_fundamental: int = 0
    half_range: bool = True
    mpi_xdim: int = 1
    mpi_ydim: int = 1
    mpi_zdim: int = 1
    max_step: int = 0
    chk_converge_step: int = 1
    save_step: int = 1 << 30
    eps: float = 0.0
    save_last: bool = True
    save_format: int = 1
    kn: float = 0.0
    press_drop: float = 0.0
    accom: float = 1.0
    precision: str = "fp64"
    device: int = 0
    exact_quadrature: bool = False
    exact_order: bool = False
    refine_switch: float = 1.0e-6
    # sweep schedule (piks_params.sweep / tile): 0 the library's default,
    # 1 hyperplanes, 2 tiles of tile = tx | ty << 8 | tz << 16
    sweep: int = 0
    tile: int = 0

    # sizes after the fluid layers are added (src/parameters.F90:121)
    @property
    def nx_total(self) -> int:
        return self.nx + 2 * self.fluid_layer

    def image_bytes(self) -> int:
        return self.nx * self.ny * self.nz

    def as_oracle_dict(self) -> dict:
        return dict(nx=self.nx, ny=self.ny, nz=self.nz, wallextorder=self.wall_ext_order,
                    fluidlayer=self.fluid_layer, ref_l=self.ref_l,
                    nc_fundamental=self.nc_fundamental, halfrange=self.half_range,
                    chkconvergestep=self.chk_converge_step, kn=self.kn,
                    pressdrop=self.press_drop, accom=self.accom,
                    exact_quadrature=int(self.exact_quadrature))


_KEYS = {
    "physicalnml": {"imagefilename": "image_file_name", "nx": "nx", "ny": "ny", "nz": "nz",
                    "wallextorder": "wall_ext_order", "fluidlayer": "fluid_layer",
                    "ref_l": "ref_l"},
    "velocitynml": {"nc_fundamental": "nc_fundamental", "halfrange": "half_range"},
    "mpinml": {"mpi_xdim": "mpi_xdim", "mpi_ydim": "mpi_ydim", "mpi_zdim": "mpi_zdim",
               "block_repx": None, "block_repy": None, "block_repz": None},
    "solvernml": {"maxstep": "max_step", "chkconvergestep": "chk_converge_step",
                  "savestep": "save_step", "eps": "eps", "savelast": "save_last",
                  "saveformat": "save_format"},
    "flownml": {"kn": "kn", "pressdrop": "press_drop", "accom": "accom"},
    "gpunml": {"precision": "precision", "device": "device",
               "exact_quadrature": "exact_quadrature", "exact_order": "exact_order",
               "refineswitch": "refine_switch"},
}


def _value(tok: str):
    t = tok.strip().rstrip(",").strip()
    if not t:
        raise ValueError("empty namelist value")
    if t[0] in "'\"":
        return t[1:-1] if t[-1] == t[0] else t[1:]
    low = t.lower()
    if low in ("t", ".true.", "true", ".t."):
        return True
    if low in ("f", ".false.", "false", ".f."):
        return False
    if re.fullmatch(r"[-+]?\d+", t):
        return int(t)
    return float(low.replace("d", "e"))


def parse_namelists(text: str) -> dict:
    """Return {group: {key: value}} for every &group ... / block."""
    out = {}
    group = None
    for raw in text.splitlines():
        line = raw
        # strip comments outside quotes
        q = None
        for i, ch in enumerate(line):
            if ch in "'\"":
                q = None if q == ch else (ch if q is None else q)
            elif ch == "!" and q is None:
                line = line[:i]
                break
        line = line.strip()
        if not line:
            continue
        if line.startswith("&"):
            group = line[1:].split()[0].lower()
            out.setdefault(group, {})
            rest = line[1 + len(group):].strip()
            line = rest
            if not line:
                continue
        if group is None:
            continue
        end = line.endswith("/")
        if end:
            line = line[:-1]
        for m in re.finditer(r"([A-Za-z_][A-Za-z0-9_]*)\s*=\s*('[^']*'|\"[^\"]*\"|[^,\s]+)", line):
            out[group][m.group(1).lower()] = _value(m.group(2))
        if end:
            group = None
    return out


def read_para(path: str = "para.in") -> Params:
    with open(path) as fh:
        return parse_para(fh.read())


def parse_para(text: str) -> Params:
    groups = parse_namelists(text)
    p = Params()
    for g, keys in _KEYS.items():
        for k, v in groups.get(g, {}).items():
            if k not in keys:
                raise ValueError("unknown variable %s in namelist &%s" % (k, g))
            attr = keys[k]
            if attr is None:
                if int(v) != 1:
                    raise ValueError("%s = %s: block replication is not supported" % (k, v))
                continue
            cur = getattr(p, attr)
            if isinstance(cur, bool):
                v = bool(v)
            elif isinstance(cur, int):
                v = int(v)
            elif isinstance(cur, float):
                v = float(v)
            setattr(p, attr, v)
    for g in ("physicalnml", "velocitynml", "solvernml", "flownml"):
        if g not in groups:
            raise ValueError("could not read namelist &%s" % g)
    return p


def _fmt(v):
    if isinstance(v, bool):
        return "T" if v else "F"
    if isinstance(v, int):
        return str(v)
    if isinstance(v, float):
        return repr(v).replace("e", "d") if "e" in repr(v) else repr(v) + "d0"
    return "'%s'" % v


def format_para(p: Params, gpu: bool = False) -> str:
    lines = [
        "&physicalNml",
        "imageFileName = %s" % _fmt(p.image_file_name),
        "Nx = %d" % p.nx, "Ny = %d" % p.ny, "Nz = %d" % p.nz,
        "wallExtOrder = %d" % p.wall_ext_order,
        "fluidLayer = %d" % p.fluid_layer,
        "Ref_L = %s" % _fmt(float(p.ref_l)),
        "/", "",
        "&velocityNml",
        "Nc_fundamental = %d" % p.nc_fundamental,
        "halfRange = %s" % _fmt(bool(p.half_range)),
        "/", "",
        "&mpiNml",
        "mpi_xdim = 1", "mpi_ydim = 1", "mpi_zdim = 1",
        "/", "",
        "&solverNml",
        "maxStep = %d" % p.max_step,
        "chkConvergeStep = %d" % p.chk_converge_step,
        "saveStep = %d" % p.save_step,
        "eps = %s" % _fmt(float(p.eps)),
        "saveLast = %s" % _fmt(bool(p.save_last)),
        "saveFormat = %d" % p.save_format,
        "/", "",
        "&flowNml",
        "Kn = %s" % _fmt(float(p.kn)),
        "pressDrop = %s" % _fmt(float(p.press_drop)),
        "accom = %s" % _fmt(float(p.accom)),
        "/",
    ]
    if gpu:
        lines += ["", "&gpuNml", "precision = '%s'" % p.precision, "device = %d" % p.device,
                  "exact_quadrature = %s" % _fmt(bool(p.exact_quadrature)),
                  "exact_order = %s" % _fmt(bool(p.exact_order)),
                  "refineSwitch = %s" % _fmt(float(p.refine_switch)), "/"]
    return "\n".join(lines) + "\n"
