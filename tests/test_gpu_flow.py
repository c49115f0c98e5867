"""The single-launch sweep (PIKS_SWEEP_FLOW=1: the whole sweep of an
iteration in one launch, octant o on XCD o, tasks waiting for their upwind
neighbours' flags) computes every update with the same operations in the same
order as the hyperplane launches, so f, fw, the moments and the check after
every iteration equal the default path's bit for bit, in both reduction
modes and with graph replays; it is refused where it does not apply."""
import numpy as np
import pytest

from piks3d_amd import PiksError, Solver, images
from piks3d_amd.params import Params

pytestmark = pytest.mark.gpu


def _pack(n, seed, nc=4):
    img = images.sphere_pack(n, 0.40, 3.0, 6.0, seed=seed)
    p = Params(nx=n, ny=n, nz=n, fluid_layer=2, wall_ext_order=2, nc_fundamental=nc, half_range=True,
               chk_converge_step=10, max_step=100, eps=0.0, kn=1.0, press_drop=1.0, accom=1.0)
    return p, img


def _run(p, img, exact, env, monkeypatch, steps=4):
    for k in ("PIKS_SWEEP_FLOW", "PIKS_GRAPH", "PIKS_FLOW_WAVES"):
        monkeypatch.delenv(k, raising=False)
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    with Solver(p, img, precision="fp64", exact_order=exact) as s:
        per = []
        for _ in range(steps):
            s.iterate(1)
            per.append((np.stack(s.macro()), s.chk_converge()))
        return s.f(), s.fw(), per


@pytest.mark.parametrize("env", [{"PIKS_SWEEP_FLOW": "1"}, {"PIKS_SWEEP_FLOW": "1", "PIKS_GRAPH": "1"},
                                 {"PIKS_SWEEP_FLOW": "1", "PIKS_FLOW_WAVES": "64"}],
                         ids=lambda e: ",".join("%s=%s" % kv for kv in e.items()))
@pytest.mark.parametrize("exact", [False, True])
def test_flow_sweep_equals_hyperplane_launches(env, exact, monkeypatch):
    p, img = _pack(40, 5)
    print("40^3 pack, exact_order=%d, %s" % (exact, env), flush=True)
    ref = _run(p, img, exact, {}, monkeypatch)
    got = _run(p, img, exact, env, monkeypatch)
    assert np.array_equal(got[0], ref[0])
    assert np.array_equal(got[1], ref[1])
    for (ma, ca), (mb, cb) in zip(got[2], ref[2]):
        assert np.array_equal(ma, mb)
        assert ca == cb


def test_flow_sweep_refused_where_it_does_not_apply(monkeypatch):
    monkeypatch.setenv("PIKS_SWEEP_FLOW", "1")
    p, img = _pack(24, 3, nc=3)           # 27 velocities per octant
    p.exact_quadrature = True
    with pytest.raises(PiksError) as e:
        Solver(p, img, precision="fp64").close()
    assert e.value.code == -1
    p, img = _pack(24, 3)
    with pytest.raises(PiksError):
        Solver(p, img, precision="fp32").close()
