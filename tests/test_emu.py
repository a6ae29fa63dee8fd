"""The HIP kernels, built for the host by tests/emu (g++ with
AddressSanitizer and a bounds-checked HIP emulation), run one full step of
a small particle case, of a wall-bounded case and of the precursor domain
on the CPU and agree with the oracle, and so does the mixed form's
two-sweep inner solve.  An
out-of-bounds index, a host pointer passed to a kernel or an invalid
launch shape aborts the child process with a report, so this guards every
kernel change before it reaches a GPU.  The full GPU suite
runs the same way with BB_EMU=1 (see tests/conftest.py)."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EMU = os.path.join(ROOT, "tests", "emu")

SCRIPT = r"""
import sys
sys.path.insert(0, %(root)r)
sys.path.insert(0, %(tests)r)
import conftest  # points the loader at the emulation build
import numpy as np
from bluebottle_amd import configs, sim
from oracle.pyoracle import OracleState

def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))

which = sys.argv[1]
cases = {"particles": lambda: configs.small_particle_case(
             n=12, nparts=2, radius_dx=2.5, lamb_max_iter=2),
         "channel": lambda: configs.channel_case(n=(12, 8, 10), nparts=1,
                                                 init="quiescent")}
for name, make in cases.items():
    if name != which:
        continue
    case = make()
    s = sim.Simulation(case)
    s.to_device()
    s.prepare()
    s.pull()
    o = OracleState.from_sim(s)
    ia = s.step()
    ib = o.call("step")
    s.pull()
    assert ia == ib, (name, ia, ib)
    for f in ("u", "v", "w", "p", "phi"):
        e = rel(s.array(f), o.array(f))
        assert e < 1e-8, (name, f, e)
    for f in ("phase", "phase_shell", "flag_u", "flag_v", "flag_w"):
        assert np.array_equal(s.array(f), o.array(f)), (name, f)
    # the mixed form's two-sweep inner solve, cold-started, against the
    # FP64 oracle's solve of the same system
    s.hip.cuda_U_star_2()
    s.hip.cuda_dom_BC_star()
    s.pull()
    s.array("phi")[:] = 0.0
    s.push()
    o = OracleState.from_sim(s)
    s.set_solver(1)
    s.hip.bb_set_inner_sweeps(2)
    s.hip.cuda_PP_bicgstab(0)
    s.hip.bb_set_inner_sweeps(3)
    s.set_solver(0)
    s.pull()
    o.call("PP_bicgstab")
    e = rel(s.array("phi"), o.array("phi"))
    assert e < 1e-8, (name, "two-sweep phi", e)
    s.close()
    print("emulated step", name, "ok", flush=True)

# the precursor domain: its set-up and one step with the linear forcing
if which != "precursor":
    sys.exit(0)
import dataclasses
from test_oracle_turb import turb_case
s = sim.Simulation(dataclasses.replace(turb_case(n=(8, 6, 5), turbA=0.5, seed=4),
                                       pp_residual=1e-11))
s.to_device()
o = OracleState.from_sim(s)
s.prepare()
o.lib.orc_prepare_turb(o.ptr)
s.step()
o.lib.orc_step_turb(o.ptr)
s.pull()
for f in ("u", "v", "w", "p", "f_x", "f_y", "f_z"):
    e = rel(s.array(f), o.array(f))
    assert e < 1e-8, ("precursor", f, e)
s.close()
print("emulated step precursor ok", flush=True)
"""


def _asan():
    out = subprocess.run(["gcc", "-print-file-name=libasan.so"],
                         capture_output=True, text=True, check=True).stdout
    return out.strip()


def test_emulated_step_matches_oracle():
    asan = _asan()
    if not os.path.isabs(asan):
        pytest.skip("AddressSanitizer runtime not available")
    subprocess.run(["make", "-s", "-j8", "-C", EMU], check=True)
    env = dict(os.environ)
    env.update(BB_EMU="1", LD_PRELOAD=asan, ASAN_OPTIONS="detect_leaks=0")
    code = SCRIPT % {"root": ROOT, "tests": os.path.join(ROOT, "tests")}
    # the three cases in processes of their own, side by side
    procs = {w: subprocess.Popen([sys.executable, "-c", code, w], env=env,
                                 stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                                 text=True)
             for w in ("particles", "channel", "precursor")}
    for w, p in procs.items():
        out, err = p.communicate(timeout=900)
        assert p.returncode == 0, (w, out[-2000:] + err[-4000:])
        assert out.count("ok") == 1, (w, out)
