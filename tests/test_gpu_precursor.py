"""GPU parity of the precursor-inflow pieces against the oracle, through the
C ABI: planes pushed with cuda_dom_turb_planes_push, the PRECURSOR ghost
fill of cuda_dom_BC and cuda_dom_BC_star, and cuda_yank_turb_planes with
the planes pulled back by cuda_dom_turb_planes_pull.  Ghost values and
planes are bit-exact (the same separately rounded operations)."""
import ctypes as C

import numpy as np
import pytest

from bluebottle_amd import abi, sim
from oracle.pyoracle import OracleState

from test_oracle_precursor import precursor_case, configs18

pytestmark = pytest.mark.gpu

CASES = {
    # every velocity face PRECURSOR, on a grid with dx, dy and dz distinct
    "all": lambda: precursor_case(n=(12, 10, 8), hi=(1.2, 1.0, 0.8)),
    # an x inflow/outflow pair, periodic in y and z
    "inflow_x": lambda: precursor_case(n=(16, 12, 10),
                                       faces=(1, 1, 0, 0, 0, 0),
                                       hi=(1.0, 0.75, 0.625)),
}

FIELDS = ("u", "v", "w", "u_star", "v_star", "w_star")


def _setup(which, seed):
    case = CASES[which]()
    s = sim.Simulation(case)
    rng = np.random.default_rng(seed)
    for nm in FIELDS + ("p",):
        a = s.array(nm)
        a[:] = rng.standard_normal(a.size)
    for nm in s.PLANES:
        a = s.array(nm)
        a[:] = rng.standard_normal(a.size)
    s.to_device()
    return case, s


@pytest.mark.parametrize("which", sorted(CASES))
def test_precursor_ghost_fill(which):
    case, s = _setup(which, 5)
    o = OracleState.from_sim(s)
    cfg = configs18(case)
    s.hip.cuda_dom_turb_planes_push(cfg)
    s.hip.cuda_dom_BC()
    s.hip.cuda_dom_BC_star()
    s.pull()
    o.call("dom_BC")
    o.call("dom_BC_star")
    for nm in FIELDS + ("p",):
        assert np.array_equal(s.array(nm), o.array(nm)), nm
    s.close()


@pytest.mark.parametrize("which", sorted(CASES))
def test_yank_turb_planes(which):
    case, s = _setup(which, 6)
    cfg = configs18(case)
    # the planes on the device start from the host's values
    s.hip.cuda_dom_turb_planes_push(cfg)
    o = OracleState.from_sim(s)
    d = s.dom
    pos = (C.c_double * 9)(d.xs + 0.37 * d.xl, 0, 0, 0, d.ys + 0.61 * d.yl,
                           0, 0, 0, d.zs + 0.43 * d.zl)
    vel = (C.c_double * 18)(*[0.01 * (q + 1) for q in range(18)])
    s.hip.cuda_yank_turb_planes(cfg, pos, vel)
    s.hip.cuda_dom_turb_planes_pull(cfg)
    s.pull()
    lib = o.lib
    lib.orc_yank_turb_planes.argtypes = [C.c_void_p, C.POINTER(C.c_int),
                                         C.POINTER(C.c_double),
                                         C.POINTER(C.c_double)]
    lib.orc_yank_turb_planes(o.ptr, cfg, pos, vel)
    for nm in s.PLANES:
        assert np.array_equal(s.array(nm), o.array(nm)), nm
    for nm in ("u", "v", "w"):
        assert np.array_equal(s.array(nm), o.array(nm)), nm
    s.close()


# ---------------------------------------------------------------------------
# the precursor domain: energy, linear forcing, colocation, whole steps
from test_oracle_turb import turb_case  # noqa: E402


def _turb_state(seed, turbA=0.6, n=(12, 10, 8)):
    case = turb_case(n=n, hi=(1.0, 0.8, 0.7), turbA=turbA, seed=seed)
    s = sim.Simulation(case)
    rng = np.random.default_rng(seed)
    for nm in ("f_x", "f_y", "f_z"):
        a = s.array(nm)
        a[:] = rng.standard_normal(a.size)
    s.to_device()
    return s


def test_compute_energy():
    s = _turb_state(21)
    o = OracleState.from_sim(s)
    k = s.hip.cuda_compute_energy()
    want = o.lib.orc_compute_energy(o.ptr)
    assert k == pytest.approx(want, rel=1e-13)
    s.close()


def test_compute_turb_forcing():
    s = _turb_state(22)
    o = OracleState.from_sim(s)
    s.hip.cuda_compute_turb_forcing()
    s.pull()
    o.lib.orc_compute_turb_forcing(o.ptr)
    for nm in ("f_x", "f_y", "f_z"):
        a, b = s.array(nm), o.array(nm)
        assert np.allclose(a, b, rtol=1e-12, atol=1e-12 * np.abs(b).max()), nm
    # the velocity is read only
    for nm in ("u", "v", "w"):
        assert np.array_equal(s.array(nm), o.array(nm)), nm
    s.close()


@pytest.mark.parametrize("comp,name", [(0, "cuda_colocate_Gfx"),
                                       (1, "cuda_colocate_Gfy"),
                                       (2, "cuda_colocate_Gfz")])
def test_colocate(comp, name):
    """Into the device phi (only its first Gcc.s3 values may change)."""
    s = _turb_state(23 + comp)
    d = s.dom
    phi = s.array("phi")
    phi[:] = -7.0
    s.hip.cuda_dom_push()
    o = OracleState.from_sim(s)
    src = C.POINTER(C.POINTER(C.c_double)).in_dll(s.host, "_" + "uvw"[comp])[0]
    dst = C.POINTER(C.POINTER(C.c_double)).in_dll(s.host, "_phi")[0]
    ddom = C.POINTER(C.c_void_p).in_dll(s.host, "_dom")[0]
    getattr(s.hip, name)(C.cast(src, C.c_void_p), C.cast(dst, C.c_void_p),
                         ddom)
    s.pull()
    want = np.zeros(d.Gcc.s3)
    o.lib.orc_colocate(o.ptr, comp, want.ctypes.data_as(C.POINTER(C.c_double)))
    got = s.array("phi")
    assert np.array_equal(got[:d.Gcc.s3], want)
    assert (got[d.Gcc.s3:] == -7.0).all()
    s.close()


def test_precursor_steps():
    """Set-up and three precursor steps (forcing, predictor, solve,
    projection) against the oracle, both solves at a 1e-13 tolerance."""
    case = turb_case(n=(16, 12, 10), hi=(1.0, 0.75, 0.625), turbA=0.4,
                     seed=31)
    s = sim.Simulation(case)
    s.to_device()
    o = OracleState.from_sim(s)
    s.prepare()
    o.lib.orc_prepare_turb(o.ptr)
    s.pull()
    assert runtime_dt() == pytest.approx(o.s.dt, rel=1e-14)
    for _ in range(3):
        s.step()
        o.lib.orc_step_turb(o.ptr)
        s.pull()
        for nm in ("u", "v", "w", "p", "f_x"):
            a, b = s.array(nm), o.array(nm)
            err = np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)
            assert err < 1e-10, (nm, err)
        assert runtime_dt() == pytest.approx(o.s.dt, rel=1e-10)
    s.close()


def runtime_dt():
    from bluebottle_amd import runtime
    return runtime.get_scalar("dt")
