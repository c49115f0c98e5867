"""The single-launch sweep (the whole sweep of an iteration in one launch,
octant o on XCD o, tasks waiting for their upwind neighbours' flags; the
default for FP64 single solves with 64 velocities per octant on a context
with its own stream) computes every update with the same operations in the
same order as the hyperplane launches (PIKS_SWEEP_FLOW=0).  So f, fw, the
moments and the check after every iteration equal that path's bit for bit:
in both reduction modes, with graph replays, with few waves per octant, and
in the claims mode a launch falls back to when an octant's waves are not all
resident (forced here from the host with a zero poll limit on the kernel's
own bounded wait for them), and with the next task's upwind loads issued
early (with the current task's stores; by default only for images whose
hyperplanes hold at least two tasks per wave).  The sweep is refused where it does not apply,
and the automatic choice leaves those contexts and the settings that shape
the hyperplane sweep alone."""
import numpy as np
import pytest

from piks3d_amd import PiksError, Solver, images, lib
from piks3d_amd.params import Params

pytestmark = pytest.mark.gpu

_ENV = ("PIKS_SWEEP_FLOW", "PIKS_GRAPH", "PIKS_FLOW_WAVES", "PIKS_FLOW_BARRIER", "PIKS_FLOW_EARLY", "PIKS_FLOW_STRIPS",
        "PIKS_OCT_GROUPS")


def _pack(n, seed, nc=4):
    img = images.sphere_pack(n, 0.40, 3.0, 6.0, seed=seed)
    p = Params(nx=n, ny=n, nz=n, fluid_layer=2, wall_ext_order=2, nc_fundamental=nc, half_range=True,
               chk_converge_step=10, max_step=100, eps=0.0, kn=1.0, press_drop=1.0, accom=1.0)
    return p, img


def _setenv(env, monkeypatch):
    for k in _ENV:
        monkeypatch.delenv(k, raising=False)
    for k, v in env.items():
        monkeypatch.setenv(k, v)


def _run(p, img, exact, env, monkeypatch, steps=4):
    _setenv(env, monkeypatch)
    with Solver(p, img, precision="fp64", exact_order=exact) as s:
        path = s.info.sweep_path
        per = []
        for _ in range(steps):
            s.iterate(1)
            per.append((np.stack(s.macro()), s.chk_converge()))
        return s.f(), s.fw(), per, path


@pytest.mark.parametrize("env", [{}, {"PIKS_SWEEP_FLOW": "1"}, {"PIKS_GRAPH": "1"}, {"PIKS_FLOW_WAVES": "64"},
                                 {"PIKS_FLOW_BARRIER": "0"}, {"PIKS_FLOW_EARLY": "1"},
                                 {"PIKS_FLOW_EARLY": "1", "PIKS_FLOW_WAVES": "64"},
                                 {"PIKS_FLOW_EARLY": "1", "PIKS_FLOW_BARRIER": "0"}, {"PIKS_FLOW_STRIPS": "3"},
                                 {"PIKS_FLOW_STRIPS": "5", "PIKS_FLOW_EARLY": "1", "PIKS_FLOW_WAVES": "64"}],
                         ids=lambda e: ",".join("%s=%s" % kv for kv in e.items()) or "default")
@pytest.mark.parametrize("exact", [False, True])
def test_flow_sweep_equals_hyperplane_launches(env, exact, monkeypatch):
    p, img = _pack(40, 5)
    print("40^3 pack, exact_order=%d, %s" % (exact, env), flush=True)
    ref = _run(p, img, exact, {"PIKS_SWEEP_FLOW": "0"}, monkeypatch)
    got = _run(p, img, exact, env, monkeypatch)
    assert ref[3] == 0 and got[3] == 2
    assert np.array_equal(got[0], ref[0])
    assert np.array_equal(got[1], ref[1])
    for (ma, ca), (mb, cb) in zip(got[2], ref[2]):
        assert np.array_equal(ma, mb)
        assert ca == cb


@pytest.mark.parametrize("early,strips", [("0", "1"), ("1", "1"), ("1", "4")])
def test_flow_sweep_on_a_larger_pack(early, strips, monkeypatch):
    # many tasks per wave (each of the waves takes tens of tasks per octant);
    # with strips, each octant's tasks go band of y by band of y
    p, img = _pack(72, 11)
    print("72^3 pack, early loads %s, strips %s" % (early, strips), flush=True)
    ref = _run(p, img, False, {"PIKS_SWEEP_FLOW": "0"}, monkeypatch, steps=2)
    got = _run(p, img, False, {"PIKS_FLOW_EARLY": early, "PIKS_FLOW_STRIPS": strips}, monkeypatch, steps=2)
    assert got[3] == 2
    assert np.array_equal(got[0], ref[0])
    assert got[2][-1][1] == ref[2][-1][1]


def test_flow_sweep_refused_where_it_does_not_apply(monkeypatch):
    _setenv({"PIKS_SWEEP_FLOW": "1"}, monkeypatch)
    p, img = _pack(24, 3, nc=3)           # 27 velocities per octant
    p.exact_quadrature = True
    with pytest.raises(PiksError) as e:
        Solver(p, img, precision="fp64").close()
    assert e.value.code == -1
    p, img = _pack(24, 3)
    with pytest.raises(PiksError):
        Solver(p, img, precision="fp32").close()


def test_automatic_choice(monkeypatch):
    # where the single launch does not apply, or a setting asks for the
    # hyperplane sweep's shape, the automatic choice keeps the launches
    p, img = _pack(24, 3, nc=3)
    p.exact_quadrature = True
    _setenv({}, monkeypatch)
    with Solver(p, img, precision="fp64") as s:
        assert s.info.sweep_path == 0
    p, img = _pack(24, 3)
    with Solver(p, img, precision="fp32") as s:
        assert s.info.sweep_path == 0
        # the automatic choice's refusal is not an error and leaves no message
        assert not lib().piks_last_error()
    with Solver(p, img, precision="fp64") as s:
        assert s.info.sweep_path == 2
    _setenv({"PIKS_OCT_GROUPS": "8"}, monkeypatch)
    with Solver(p, img, precision="fp64") as s:
        assert s.info.sweep_path == 0
    _setenv({"PIKS_SWEEP_FLOW": "0"}, monkeypatch)
    with Solver(p, img, precision="fp64") as s:
        assert s.info.sweep_path == 0


def test_concurrent_single_launch_sweeps(monkeypatch):
    # two contexts, each with its own stream, iterating at the same time from
    # two host threads: their single launches compete for the CUs, so an
    # octant whose waves are not all resident claims its tasks instead; each
    # result equals the same context's run alone, bit for bit
    import threading
    _setenv({}, monkeypatch)
    cases = [_pack(56, 21), _pack(48, 22)]
    alone = []
    for p, img in cases:
        with Solver(p, img, precision="fp64") as s:
            assert s.info.sweep_path == 2
            s.iterate(6)
            alone.append((s.f(), s.chk_converge()))
    solvers = [Solver(p, img, precision="fp64") for p, img in cases]
    errors = []

    def work(s):
        try:
            for _ in range(6):
                s.iterate(1)
        except Exception as e:  # reported below, in the test's thread
            errors.append(e)

    try:
        threads = [threading.Thread(target=work, args=(s,)) for s in solvers]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        assert not errors, errors
        for s, (f, chk) in zip(solvers, alone):
            assert s.info.sweep_path == 2
            assert np.array_equal(s.f(), f)
            assert s.chk_converge() == chk
    finally:
        for s in solvers:
            s.close()
