"""Several Knudsen numbers on one geometry (piks_create_multi).  Each
right-hand side runs the same kernels with its own mu and state, so its
iterates must equal those of a single-Kn context bit for bit; the oracle
pins the single-Kn context."""
import numpy as np
import pytest

from conftest import load_golden, params_from_golden
from piks3d_amd import MultiSolver, Solver, images
from piks3d_amd.params import Params

pytestmark = pytest.mark.gpu

KNS = [0.3, 1.0, 5.0, 0.07, 2.5]


@pytest.mark.parametrize("exact", [True, False])
@pytest.mark.parametrize("name", ["boxes_n2_a05_e1", "coupled_fl0_a06"])
def test_each_rhs_equals_its_single_solve(name, exact):
    z = load_golden(name)
    p = params_from_golden(z)
    img = z["image"]
    with MultiSolver(p, img, KNS, exact_order=exact) as ms:
        ms.iterate(5)
        multi = [ms.macro_rhs(r) for r in range(len(KNS))]
        m_all, k_all, e_all = ms.chk_converge_all()
    for r, kn in enumerate(KNS):
        p.kn = kn
        with Solver(p, img, exact_order=exact) as s:
            s.iterate(5)
            single = s.macro()
            m, k, e = s.chk_converge()
        for a, b in zip(multi[r], single):
            assert np.array_equal(a, b), (name, kn)
        assert (m_all[r], k_all[r], e_all[r]) == (m, k, e)


def test_refined_multi_matches_fp64_multi():
    n = 32
    img = images.sphere_pack(n, 0.40, 3.0, 6.0, seed=9)
    p = Params(nx=n, ny=n, nz=n, fluid_layer=2, wall_ext_order=2, nc_fundamental=4, half_range=True,
               chk_converge_step=10, max_step=500, eps=0.0, kn=1.0, press_drop=1.0, accom=1.0)
    kns = [0.5, 2.0, 8.0]
    with MultiSolver(p, img, kns, precision="fp64") as a:
        r64 = a.solve()
    with MultiSolver(p, img, kns, precision="fp32") as b:
        r32 = b.solve()
    assert 0 < r32["steps32"] < r32["steps"] == r64["steps"] == 500
    rel = np.abs(np.asarray(r32["permeability"]) - r64["permeability"]) / np.abs(r64["permeability"])
    assert np.all(rel <= 1e-8), rel


def test_single_rhs_calls_on_a_multi_context():
    # piks_chk_converge and piks_mass_flow on a context with several
    # right-hand sides report the first one (and must not write past it)
    z = load_golden("boxes_n2_a05_e1")
    p = params_from_golden(z)
    with MultiSolver(p, z["image"], KNS) as ms:
        ms.iterate(3)
        m0 = ms.mass_flow()
        m, k, e = ms.chk_converge()
        m_all, k_all, e_all = ms.chk_converge_all()
    assert m0 == m == m_all[0]
    assert k == k_all[0]
    assert e_all[0] == 0.0                    # the stored mass was updated by chk_converge


@pytest.mark.parametrize("env", [{"PIKS_MULTI_RB": "2"}, {"PIKS_MULTI_RB": "4"},
                                 {"PIKS_SWEEP_BLOCK": "256", "PIKS_MULTI_RB": "2"}, {"PIKS_SWEEP_GENERIC": "1"}],
                         ids=lambda e: ",".join("%s=%s" % kv for kv in e.items()))
def test_sweep_layouts_keep_the_bits(env, monkeypatch):
    # the several-Kn sweep with 2 or 4 right-hand sides per wave sharing one
    # task record, with 256-thread workgroups, or through the generic kernel
    # (instead of the single-solve task kernel, one right-hand side per wave)
    # gives the same iterates, in both precisions
    z = load_golden("boxes_n2_a05_e1")
    p = params_from_golden(z)
    for prec in ("fp64", "fp32"):
        runs = []
        for v in (None, env):
            for k in ("PIKS_MULTI_RB", "PIKS_SWEEP_BLOCK", "PIKS_SWEEP_GENERIC"):
                monkeypatch.delenv(k, raising=False)
            for k, val in (v or {}).items():
                monkeypatch.setenv(k, val)
            monkeypatch.setenv("PIKS_DC_BLOCK", "3")
            with MultiSolver(p, z["image"], KNS, precision=prec) as ms:
                ms.iterate(7)
                runs.append([np.stack(ms.macro_rhs(r)) for r in range(len(KNS))])
        for a, b in zip(*runs):
            assert np.array_equal(a, b), (prec, env)
