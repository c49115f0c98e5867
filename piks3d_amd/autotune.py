"""Autotuner of the sweep schedule per request class.

One source iteration's sweep is a chain of dependent launches: one per
oriented hyperplane (i + j + k = const, all octants), or, with the tiled
schedule, one per tile diagonal, each wave walking a tile of tx * ty * tz
cells of one octant plane by plane.  Which is faster depends on the request:
large images keep every launch full and favour the hyperplane sweep's short
waves; small ones are bound by the launch chain's length, which tiles cut by
their side.  This module times every candidate schedule on a representative
request of each class of the serving stream (cube side, velocity grid:
seeded sphere pack at porosity 0.4, Kn 1, 4 fluid layers, wallExtOrder 2)
and keeps the fastest per class in a table (JSON).  Server.submit applies the
table to requests that do not choose a schedule themselves.

    python -m piks3d_amd.autotune [--out piks3d_amd/tuning.json] [--sides 32,48,...]

Timing: device time of whole iterations (HIP events, piks_last_timing), the
best of two runs of --steps iterations after two warm-up iterations.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "tuning.json")

SIDES = (32, 48, 64, 96, 128)
GRIDS = (3, 4, 6)          # Nc_fundamental of the 6^3, 8^3, 12^3 velocity grids


def tile_code(tx, ty, tz):
    return tx | ty << 8 | tz << 16


def tile_shape(code):
    return code & 255, (code >> 8) & 255, code >> 16


# (sweep, tile) candidates: the hyperplane sweep and tiled sweeps (the tiled
# kernel serves grids of at most 64 velocities per octant)
TILES = [tile_code(*t) for t in ((2, 2, 2), (2, 2, 4), (4, 2, 4), (4, 4, 4), (4, 2, 8), (4, 4, 8))]


def candidates(nc):
    c = [(1, 0)]
    if nc ** 3 <= 64:
        c += [(2, t) for t in TILES]
    return c


def label(sweep, tile):
    if sweep == 1:
        return "hyperplanes"
    return "tiles %dx%dx%d" % tile_shape(tile)


def representative(side, nc, seed=11):
    from . import images
    from .params import Params
    img = images.sphere_pack(side, 0.40, 4.0, 10.0, seed=seed)
    p = Params(nx=side, ny=side, nz=side, fluid_layer=4, wall_ext_order=2, nc_fundamental=nc, half_range=True,
               exact_quadrature=(nc == 3), chk_converge_step=10, max_step=10, eps=0.0, kn=1.0,
               press_drop=1.0, accom=1.0)
    return p, img


def time_schedule(p, img, sweep, tile, steps):
    from . import Solver
    p.sweep, p.tile = sweep, tile
    best = math.inf
    with Solver(p, img, precision="fp64") as s:
        s.iterate(2)
        for _ in range(2):
            s.iterate(steps)
            best = min(best, s.timing()[3] / steps)
    return best


def tune(sides=SIDES, grids=GRIDS, steps=5, log=print):
    rows = []
    for nc in grids:
        for side in sides:
            p, img = representative(side, nc)
            ms = {}
            for sweep, tile in candidates(nc):
                t = time_schedule(p, img, sweep, tile, steps)
                ms[label(sweep, tile)] = round(t, 4)
                log("side %d nc %d %-14s %.3f ms/iteration" % (side, nc, label(sweep, tile), t))
            sweep, tile = min(candidates(nc), key=lambda c: ms[label(*c)])
            rows.append({"side": side, "nc": nc, "sweep": sweep, "tile": tile, "choice": label(sweep, tile),
                         "ms_per_iteration": ms})
    return rows


def load(path=TABLE):
    if not path or not os.path.exists(path):
        return None
    with open(path) as fh:
        return json.load(fh)


def choose(table, side, nc):
    """(sweep, tile) of the table's class nearest in cube side with the same
    velocity grid; (0, 0), the library's default, without one."""
    if not table:
        return 0, 0
    rows = [r for r in table.get("classes", []) if r["nc"] == nc]
    if not rows:
        return 0, 0
    r = min(rows, key=lambda r: abs(math.log(r["side"] / max(side, 1))))
    return int(r["sweep"]), int(r["tile"])


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=TABLE)
    ap.add_argument("--sides", default=",".join(map(str, SIDES)))
    ap.add_argument("--grids", default=",".join(map(str, GRIDS)))
    ap.add_argument("--steps", type=int, default=5)
    a = ap.parse_args(argv)
    t0 = time.time()
    rows = tune([int(x) for x in a.sides.split(",")], [int(x) for x in a.grids.split(",")], a.steps,
                log=lambda *x: print(*x, file=sys.stderr, flush=True))
    table = {"what": "sweep schedule per request class (piks3d_amd/autotune.py)",
             "steps_timed": a.steps, "tuning_s": round(time.time() - t0, 1), "classes": rows}
    with open(a.out, "w") as fh:
        json.dump(table, fh, indent=1)
    print(json.dumps(table))


if __name__ == "__main__":
    main()
