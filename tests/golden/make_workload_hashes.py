"""Write tests/golden/workload_images.sha256: the SHA-256 of every benchmark
image, regenerated from its seeded generator (piks3d_amd.images), so that a
change to a generator, a seed or a size shows up as a failing test rather than
as a silently different workload.  Lines: "<sha256>  <name>".

    python tests/golden/make_workload_hashes.py
"""
from __future__ import annotations

import hashlib
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)


def workload_images():
    """(name, uint8 image) for the benchmark workloads: the 64^3 centred
    sphere, the 200^3 sphere pack, the 300^3 and 160^3 random fields, and the
    first eight requests of the serving stream (seed 7)."""
    import bench
    from piks3d_amd import images
    yield "sphere64_r28", images.centered_sphere(64, 28.0)
    yield "spheres200_phi040_seed%d" % bench.SEED, bench.workload()[1]
    yield "grf300_phi020_l6_seed7", images.gaussian_field(300, 0.20, 6.0, seed=7)
    yield "grf160_phi025_l5_seed7", images.gaussian_field(160, 0.25, 5.0, seed=7)
    for q, (p, img, _) in enumerate(bench.serve_workload(8, 7, 400, 1e-6)):
        yield "serve_seed7_req%d_side%d" % (q, p.nx), img


def main():
    lines = ["%s  %s" % (hashlib.sha256(img.tobytes()).hexdigest(), name) for name, img in workload_images()]
    with open(os.path.join(HERE, "workload_images.sha256"), "w") as fh:
        fh.write("\n".join(lines) + "\n")
    print("\n".join(lines))


if __name__ == "__main__":
    main()
