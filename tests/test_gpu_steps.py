"""GPU parity of whole time steps (src/bluebottle.c:598-711 order) against
the FP64 oracle, with no resynchronisation between steps: the errors
printed are what accumulates over three steps, Lamb sub-iterations
included.  Integer cell flags must stay bit-exact throughout."""
import pytest

from bluebottle_amd import sim
from oracle.pyoracle import OracleState

from helpers import compare_ints, rel_err
from test_gpu_parity import CASES, TOL_SOLVE

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case_name", ["settling", "channel", "tg_free",
                                       "tg_particles", "channel_free"])
def test_full_steps(case_name):
    print("case", case_name, flush=True)
    case = CASES[case_name]()
    s = sim.Simulation(case)
    s.to_device()
    s.prepare()
    s.pull()
    o = OracleState.from_sim(s)
    worst = {}
    for step in range(3):
        ia = s.step()
        ib = o.call("step")
        s.pull()
        assert ia == ib, "Lamb sub-iterations differ at step %d" % step
        for f in ("u", "v", "w", "p", "phi"):
            worst[f] = max(worst.get(f, 0.0), rel_err(s.array(f), o.array(f)))
        compare_ints(s, o)
    print("full steps %s: %s" % (case_name, {k: "%.2e" % v for k, v in
                                              worst.items()}))
    for f, e in worst.items():
        assert e <= TOL_SOLVE * 10, "%s: %.3e" % (f, e)
    s.close()
