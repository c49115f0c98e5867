"""The sweep's launch variants give the default path's results.

A context reads twelve environment settings at creation (piks3d.hip,
ctx_create_host; geometry.cpp): PIKS_HOST_THREADS (host threads of the image
preprocessing, whose loops write disjoint data), PIKS_XCD_REMAP
(XCD-contiguous task ranges), PIKS_SWEEP_BLOCK (64 or 256 threads per sweep
workgroup), PIKS_GRAPH (one HIP graph per
iteration), PIKS_SWEEP_GENERIC (the multi-right-hand-side kernel for a single
solve), PIKS_SWEEP_TILES and PIKS_TILE (the tiled sweep, one wave per tile of
one octant walking its planes, and its tile shape), PIKS_PLANES_FUSED (0: the
inlet/outlet and the two symmetry directions as three launches instead of
one), PIKS_SWEEP_NT (tasks per wave of the single-solve sweep, 1-4, whose
loads are all issued before the first update), PIKS_OCT_GROUPS (the octants
swept in 1, 2, 4 or 8 groups, as concurrent launch chains on streams of
their own; 4 by default) and PIKS_OCT_SERIAL (the groups one after the other on the
context's stream) and PIKS_SWEEP_RECIP (fast
mode: multiply by a tabulated reciprocal of the denominator).  All but the
last compute every update with the same
operations in the same order, so f, fw and the moments after several
iterations equal the default path's bit for bit, in both reduction modes; the
reciprocal variant rounds once more per update and is held to the fast mode's
1e-12 against the oracle.
"""
import numpy as np
import pytest

from conftest import interior_mask, load_golden, params_from_golden

pytestmark = pytest.mark.gpu

VARIANTS = [
    {"PIKS_HOST_THREADS": "1"},
    {"PIKS_HOST_THREADS": "7"},
    {"PIKS_XCD_REMAP": "0"},
    {"PIKS_SWEEP_BLOCK": "256"},
    {"PIKS_GRAPH": "1"},
    {"PIKS_SWEEP_GENERIC": "1"},
    {"PIKS_SWEEP_BLOCK": "256", "PIKS_SWEEP_GENERIC": "1", "PIKS_GRAPH": "1"},
    {"PIKS_SWEEP_TILES": "1"},
    {"PIKS_SWEEP_TILES": "1", "PIKS_TILE": "4x4x16"},
    {"PIKS_SWEEP_TILES": "1", "PIKS_TILE": "8x2x3", "PIKS_GRAPH": "1"},
    {"PIKS_PLANES_FUSED": "0"},
    {"PIKS_SWEEP_NT": "1"},
    {"PIKS_SWEEP_NT": "2"},
    {"PIKS_SWEEP_NT": "3", "PIKS_XCD_REMAP": "0"},
    {"PIKS_SWEEP_NT": "4", "PIKS_GRAPH": "1"},
    {"PIKS_OCT_GROUPS": "1"},
    {"PIKS_OCT_GROUPS": "2"},
    {"PIKS_OCT_GROUPS": "4", "PIKS_SWEEP_GENERIC": "1"},
    {"PIKS_OCT_GROUPS": "8", "PIKS_GRAPH": "1", "PIKS_SWEEP_NT": "2"},
    {"PIKS_OCT_GROUPS": "4", "PIKS_OCT_SERIAL": "1"},
    {"PIKS_OCT_GROUPS": "2", "PIKS_GRAPH": "1"},
]
CASES = ["boxes_n2_a1_e2", "spheres40_n2"]
STEPS = 4


def _run(p, img, exact, env, monkeypatch):
    from piks3d_amd import Solver
    for k in ("PIKS_XCD_REMAP", "PIKS_SWEEP_BLOCK", "PIKS_GRAPH", "PIKS_SWEEP_GENERIC", "PIKS_SWEEP_RECIP",
              "PIKS_HOST_THREADS", "PIKS_SWEEP_TILES", "PIKS_TILE", "PIKS_PLANES_FUSED", "PIKS_SWEEP_NT",
              "PIKS_OCT_GROUPS", "PIKS_OCT_SERIAL"):
        monkeypatch.delenv(k, raising=False)
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    with Solver(p, img, precision="fp64", exact_order=exact) as s:
        # one iteration per call, with the check's own kernel and a
        # synchronisation between calls: under PIKS_GRAPH every call replays
        # the captured iteration, and each replay is compared
        per_step = []
        for _ in range(STEPS):
            s.iterate(1)
            per_step.append((np.stack(s.macro()), s.chk_converge()))
        out = (s.f(), s.fw(), per_step, s.chk_converge())
    for k in env:
        monkeypatch.delenv(k, raising=False)
    return out


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("exact", [True, False])
@pytest.mark.parametrize("env", VARIANTS, ids=lambda e: ",".join("%s=%s" % kv for kv in e.items()))
def test_variant_equals_default(name, exact, env, monkeypatch):
    print("case %s exact_order=%d %s" % (name, exact, env), flush=True)
    z = load_golden(name)
    p = params_from_golden(z)
    ref = _run(p, z["image"], exact, {}, monkeypatch)
    got = _run(p, z["image"], exact, env, monkeypatch)
    for a, b, what in zip(got[:2], ref[:2], ("f", "fw")):
        assert np.array_equal(a, b), "%s differs under %s" % (what, env)
    for it, ((ma, ca), (mb, cb)) in enumerate(zip(got[2], ref[2])):
        assert np.array_equal(ma, mb), "moments differ after iteration %d under %s" % (it + 1, env)
        assert ca == cb
    assert got[3] == ref[3]


@pytest.mark.parametrize("name", CASES)
def test_reciprocal_variant_within_fast_tolerance(name, monkeypatch):
    print("case %s reciprocal" % name, flush=True)
    from oracle import Oracle
    z = load_golden(name)
    p = params_from_golden(z)
    got = _run(p, z["image"], False, {"PIKS_SWEEP_RECIP": "1"}, monkeypatch)
    o = Oracle(p.as_oracle_dict(), z["image"])
    o.iterate(STEPS)
    inner = interior_mask(int(z["nx"]) + 2 * p.fluid_layer, int(z["ny"]), int(z["nz"]))
    fg, fo = got[0][:, :, inner], o.f()[:, :, inner]
    assert np.abs(fg - fo).max() <= 1e-12 * np.abs(fo).max()
    assert abs(got[3][1] - o.chk_converge()[1]) <= 1e-12 * abs(o.chk_converge()[1])
