"""cuda_lamb_err (src/cuda_quadrature.cu:316-398, src/quadrature_kernel.cu:
578-654) against the oracle's literal restatement of the reference's
repeated selection, on coefficient sets chosen to reach every branch of the
selection: equal squares (the lowest index picked first), squares at or
below FLT_MIN (never picked), no qualifying coefficient at all (index 0
picked first, then zeroes), NaN and infinite values, a negative lamb_cut
(so the picks past the qualifying ones score), and both Lamb orders whose
6 * stride coefficients fit one (90) and two (126) per lane.  One particle,
so the returned maximum is that particle's error; results are compared as
doubles, bit for bit."""
import math

import numpy as np
import pytest

from bluebottle_amd import runtime, sim
from oracle.pyoracle import OracleState

pytestmark = pytest.mark.gpu

COEFFS = ("pnm_re", "pnm_im", "phinm_re", "phinm_im", "chinm_re", "chinm_im")


def _case(order):
    p = sim.Particle(x=0.5, y=0.5, z=0.5, r=0.15, order=order)
    return sim.Case(n=(12, 12, 12), nu=0.01, init="quiescent", particles=[p])


def _sets(n6, rng):
    tiny = math.sqrt(np.finfo(np.float32).tiny) * 0.5
    out = {}
    out["random"] = (rng.standard_normal(n6), rng.standard_normal(n6), 1e-4)
    a = rng.standard_normal(n6)
    a[[3, 17, 40, n6 - 1]] = [2.5, -2.5, 2.5, -2.5]
    out["ties_at_max"] = (a, a * (1 + 1e-3 * rng.standard_normal(n6)), 1e-4)
    b = np.round(rng.standard_normal(n6), 1)
    out["many_ties"] = (b, b + 0.01, 0.05)
    out["all_zero"] = (np.zeros(n6), rng.standard_normal(n6), 1e-4)
    c = np.zeros(n6)
    c[0] = tiny
    c[5] = -tiny
    out["all_tiny"] = (c, rng.standard_normal(n6), 1e-4)
    d = rng.standard_normal(n6) * 1e-3
    d[::7] = tiny
    out["some_tiny"] = (d, rng.standard_normal(n6) * 1e-3, 0.0)
    e = rng.standard_normal(n6)
    e[[2, 9, 70 % n6]] = np.nan
    out["nan"] = (e, rng.standard_normal(n6), 1e-4)
    f = rng.standard_normal(n6)
    f[[11, 33]] = [np.inf, -np.inf]
    out["inf"] = (f, rng.standard_normal(n6), 1e-4)
    g = rng.standard_normal(n6)
    g[rng.integers(0, n6, n6 // 3)] = 0.0
    out["negative_cut"] = (g, rng.standard_normal(n6), -0.5)
    out["negative_cut_all_zero"] = (np.zeros(n6), rng.standard_normal(n6), -1.0)
    h = rng.standard_normal(n6)
    out["unchanged"] = (h, h.copy(), 0.0)
    return out


@pytest.mark.parametrize("order", [4, 5])
def test_lamb_err_selection_cases(order):
    s = sim.Simulation(_case(order))
    s.to_device()
    try:
        stride = runtime.get_scalar("coeff_stride", "i")
        n6 = 6 * stride
        assert n6 == (90 if order == 4 else 126)
        rng = np.random.default_rng(order)
        for name, (cur, prev, cut) in _sets(n6, rng).items():
            for c, nm in enumerate(COEFFS):
                s.array(nm)[:] = cur[c * stride:(c + 1) * stride]
                s.array(nm + "0")[:] = prev[c * stride:(c + 1) * stride]
            runtime.set_scalar("lamb_cut", cut)
            s.hip.cuda_part_push()
            o = OracleState.from_sim(s)
            a = s.hip.cuda_lamb_err()
            b = o.lib.orc_lamb_err(o.ptr)
            assert a == b or (math.isnan(a) and math.isnan(b)), (name, a, b)
    finally:
        s.close()
