"""Golden fixtures (tests/golden/*.npz, written by
tests/golden/make_fixtures.py): one set-up plus one time step of two small
seeded cases as the FP64 oracle computes them.

* The oracle must reproduce them bit for bit (a regression pin on the
  oracle and on the seeded case set-up).
* The HIP path, through the C ABI on the GPU, must match them: integer
  cell flags bit-exact, the Lamb sub-iteration count exactly, fields,
  coefficients and particle records to 1e-8 relative (one Poisson solve at
  rel. tolerance 1e-10/1e-11 along a different dot-product summation
  order; the full-step GPU tests measure 1e-10 to 4e-10).
"""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_fixtures as mf  # noqa: E402

from helpers import rel_err  # noqa: E402

TOL_STEP = 1e-8


def _load(name):
    with np.load(os.path.join(HERE, "golden", name + ".npz")) as f:
        return {k: f[k] for k in f.files}


@pytest.mark.parametrize("name", sorted(mf.cases()))
def test_oracle_reproduces_fixture(name):
    want = _load(name)
    s, o, it = mf.oracle_step(mf.cases()[name])
    got = mf.snapshot(o, it)
    s.close()
    assert sorted(got) == sorted(want)
    for k in want:
        assert np.array_equal(got[k], want[k]), k


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(mf.cases()))
def test_hip_step_matches_fixture(name):
    from bluebottle_amd import sim
    want = _load(name)
    s = sim.Simulation(mf.cases()[name])
    s.to_device()
    s.prepare()
    it = s.step()
    s.pull()
    assert it == int(want["lamb_iters"])
    for f in mf.INTS:
        assert np.array_equal(s.array(f), want[f]), f
    worst = {}
    for f in mf.REAL:
        worst[f] = rel_err(s.array(f), want[f])
    if s.case.particles:
        for c in mf.COEFFS:
            worst[c] = rel_err(s.array(c), want[c])
        parts = s.parts()
        for k in mf.PART:
            worst["part_" + k] = rel_err(
                np.array([getattr(p, k) for p in parts]), want["part_" + k])
    s.close()
    print("golden %s: %s" % (name, {k: "%.1e" % v for k, v in worst.items()}))
    bad = {k: v for k, v in worst.items() if not v <= TOL_STEP}
    assert not bad, bad
