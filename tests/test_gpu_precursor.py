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
