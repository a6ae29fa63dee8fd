"""Restart continuation on the device: a run that writes a restart file
(src/domain.c:3223-3320, after cuda_dom_pull/cuda_part_pull as in
src/bluebottle.c:786-791) and keeps stepping must agree bit for bit with a
fresh set-up that reads the file back (src/bluebottle.c:453-476) and takes
the same steps.  The file carries the whole time-stepping state (fields,
AB2 history, phase and flags, particle records and Lamb coefficients), and
every kernel is deterministic, so nothing may differ."""
import ctypes as C

import numpy as np
import pytest

from bluebottle_amd import abi, configs, runtime, sim

pytestmark = pytest.mark.gpu

FIELDS = ("u", "v", "w", "u0", "v0", "w0", "conv_u", "conv_v", "conv_w",
          "diff_u", "diff_v", "diff_w", "p", "p0", "phi", "phase",
          "phase_shell", "flag_u", "flag_v", "flag_w")

CASES = {
    "settling": lambda: configs.small_particle_case(
        n=24, nparts=3, radius_dx=3.0, taylor_green=False, lamb_max_iter=2),
    "channel": lambda: configs.channel_case(n=(16, 12, 16), nparts=2,
                                            init="quiescent"),
    "channel_free": lambda: configs.channel_case(n=(16, 12, 16), nparts=0),
    "tg_free": lambda: configs.periodic_case(16, 0.01, CFL=0.3,
                                             pp_max_iter=500,
                                             pp_residual=1e-11),
}


def _snapshot(s):
    s.pull()
    out = {nm: s.array(nm).copy() for nm in FIELDS}
    n = len(s.case.particles)
    if n:
        ptr = C.POINTER(abi.part_struct).in_dll(s.host, "parts")
        out["parts"] = bytes(C.string_at(C.addressof(ptr[0]),
                                         n * C.sizeof(abi.part_struct)))
        for nm in ("pnm_re", "pnm_im", "phinm_re", "phinm_im", "chinm_re",
                   "chinm_im"):
            out[nm] = s.array(nm).copy()
    out["ttime"] = runtime.get_scalar("ttime")
    out["dt"] = runtime.get_scalar("dt")
    out["stepnum"] = runtime.get_scalar("stepnum", "i")
    return out


@pytest.mark.parametrize("which", sorted(CASES))
def test_restart_continues_bit_exact(which, tmp_path):
    before, after = 3, 2
    path = tmp_path / "restart.config"
    s = sim.Simulation(CASES[which]())
    s.set_solver(0)
    s.to_device()
    s.prepare()
    for _ in range(before):
        s.step()
    s.write_restart(path)
    for _ in range(after):
        s.step()
    a = _snapshot(s)
    s.close()

    r = sim.Simulation(CASES[which]())
    r.set_solver(0)
    r.to_device()
    r.restart(path)
    for _ in range(after):
        r.step()
    b = _snapshot(r)
    r.close()

    assert a["stepnum"] == b["stepnum"] == before + after
    for k in a:
        if isinstance(a[k], np.ndarray):
            assert np.array_equal(a[k], b[k]), k
        else:
            assert a[k] == b[k], k


def test_vtk_after_device_steps(tmp_path):
    """ParaView output of a stepped device state: the written pressure is
    the pulled field, and every particle is drawn at its pulled centre."""
    import xml.etree.ElementTree as ET
    s = sim.Simulation(CASES["settling"]())
    s.set_solver(0)
    s.to_device()
    s.prepare()
    s.step()
    s.write_vtk(tmp_path, 0)
    gc = s.dom.Gcc
    p3 = s.array("p").reshape(gc.knb, gc.jnb, gc.inb)[1:-1, 1:-1, 1:-1]
    root = ET.parse(tmp_path / "out_0_0.vtr").getroot()
    arrays = {da.get("Name"): da.text.split() for da in root.iter("DataArray")}
    assert [float(x) for x in arrays["p"]] == [float("%e" % x)
                                               for x in p3.ravel()]
    proot = ET.parse(tmp_path / "out_0.vtp").getroot()
    piece = proot.find("PolyData/Piece")
    pts = np.array([float(x) for x in piece.find("Points/DataArray").text.split()])
    ind = [int(x) for x in piece.find("PointData/DataArray").text.split()]
    pts = pts.reshape(-1, 3)
    for q, P in enumerate(s.parts()):
        # the last point drawn for a particle is the particle itself
        row = max(r for r, i in enumerate(ind) if i == q)
        assert np.allclose(pts[row], (P.x, P.y, P.z), atol=1e-6)
    s.close()
