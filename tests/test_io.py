"""The driver side's files in the reference's formats (bb_io.c), without a
GPU:

* flow.config and part.config: the reference's own sample inputs
  (tests/golden/ref_input, copied from its sim/input) parse to the values
  their text states, including files that leave out optional lines; every
  benchmark configuration written by configs.write_case reads back equal,
  bit for bit;
* the restart dump: its size is the reference's layout
  (src/domain.c:3223-3413), values sit at the offsets that layout gives
  (checked with an independent numpy reader of the documented field
  order), gradP.z is written with a count of 2 and read with a count of 1,
  and a dump read back restores every array bit for bit;
* the Lamb log (src/recorder.c:1955-2001): one block per particle with
  its own expansion length.
"""
import ctypes as C
import dataclasses
import os

import numpy as np
import pytest

from bluebottle_amd import abi, configs, runtime, sim

HERE = os.path.dirname(os.path.abspath(__file__))
REF = os.path.join(HERE, "golden", "ref_input")


def _read_flow(name, tmp_path):
    d = tmp_path / "case" / "input"
    d.mkdir(parents=True)
    (d / "flow.config").write_bytes(open(os.path.join(REF, name), "rb").read())
    (d / "part.config").write_text("n 0\n(l/a) 1.\n")
    host, _ = runtime.load()
    assert host.bb_read_flow_config(str(d / "flow.config").encode()) == 0
    return host


def test_reference_flow_config(tmp_path):
    _read_flow("flow.config", tmp_path)
    D = runtime.gvar("Dom", abi.dom_struct)
    assert (D.xs, D.xe, D.xn) == (-2.0, 2.0, 32)
    b = runtime.bc()
    assert b.pW == abi.PERIODIC and b.pB == abi.NEUMANN and b.dsB == 0.0
    assert b.uB == abi.DIRICHLET and b.uBDm == 0.0 and b.uBDa == 0.0
    assert runtime.get_scalar("pp_max_iter", "i") == 2000
    assert runtime.get_scalar("pp_residual") == 1e-6
    assert runtime.get_scalar("lamb_cut") == 1e-4
    assert runtime.get_scalar("init_cond", "i") == abi.QUIESCENT
    assert runtime.get_scalar("out_plane", "i") == abi.BOTTOM
    gp = runtime.gvar("gradP", abi.gradP_struct)
    assert (gp.zm, gp.za, gp.z) == (-10.0, 0.0, 0.0)
    assert runtime.get_scalar("dev_start", "i") == 0


def test_reference_sedimentation_config_without_optional_lines(tmp_path):
    """flow.config-sedimentation has no vel_tDelay, p_tDelay or g_tDelay
    lines: they keep their previous values, and the rest still parses --
    except the g lines: the attempt at "g_tDelay" consumes the "g" of
    "g.x" before failing on "_", so the g.x, g.y, g.z and gain lines that
    follow do not match either and g keeps its previous value, exactly as
    the reference's sequence of fscanf calls leaves it."""
    runtime.set_scalar("vel_tDelay", 0.25)
    runtime.set_scalar("p_tDelay", 0.5)
    g = runtime.gvar("g", abi.g_struct)
    g.zm, g.za = 7.0, 3.0
    _read_flow("flow.config-sedimentation", tmp_path)
    assert runtime.get_scalar("vel_tDelay") == 0.25
    assert runtime.get_scalar("p_tDelay") == 0.5
    D = runtime.gvar("Dom", abi.dom_struct)
    assert (D.xs, D.xe, D.xn) == (-4.0, 4.0, 64)
    b = runtime.bc()
    assert b.pW == abi.NEUMANN and b.pB == abi.PERIODIC
    assert b.uW == abi.NEUMANN and b.uB == abi.PERIODIC
    assert runtime.get_scalar("init_cond", "i") == abi.SHEAR
    assert runtime.get_scalar("out_plane", "i") == abi.NORTH
    assert runtime.get_scalar("lamb_relax") == 0.5
    assert runtime.get_scalar("lamb_max_iter", "i") == 2000
    assert (g.zm, g.za) == (7.0, 3.0)
    assert runtime.get_scalar("dev_start", "i") == 1


def test_reference_duct_config(tmp_path):
    _read_flow("flow.config-grad_P_duct", tmp_path)
    b = runtime.bc()
    assert b.pS == abi.NEUMANN and b.uS == abi.DIRICHLET
    assert b.wT == abi.DIRICHLET
    assert runtime.get_scalar("out_plane", "i") == abi.EAST
    gp = runtime.gvar("gradP", abi.gradP_struct)
    assert (gp.xm, gp.xa) == (-10.0, -10.0)


def test_reference_part_configs():
    host, _ = runtime.load()
    assert host.bb_read_part_config(
        os.path.join(REF, "part.config").encode()) == 0
    P = C.POINTER(abi.part_struct).in_dll(host, "parts")
    assert runtime.get_scalar("nparts", "i") == 1
    p = P[0]
    assert (p.r, p.x, p.rho, p.E, p.sigma, p.order, p.rs) == (
        1.0, 0.0, 2.0, 1e6, 0.5, 2, 1.2)
    assert (p.e_dry, p.coeff_fric, p.translating, p.rotating) == (1.0, 0.5,
                                                                  1, 1)
    # billiards: 11 particles, an older format without e_dry, coeff_fric
    # and rs/r lines
    assert host.bb_read_part_config(
        os.path.join(REF, "part-billiards.config").encode()) == 0
    assert runtime.get_scalar("nparts", "i") == 11
    p0, p1 = P[0], P[1]
    assert (p0.r, p0.z, p0.aFz, p0.E, p0.order) == (0.48, -3.0, 100.0, 1e7, 2)
    assert p1.z == -0.8164965809
    assert (p0.e_dry, p0.rs, p0.translating) == (0.0, 0.0, 1)


@pytest.mark.parametrize("which", ["config1", "config2", "channel",
                                   "config5_small"])
def test_case_round_trip(which, tmp_path):
    case = {"config1": lambda: configs.config1(),
            "config2": lambda: configs.config2(),
            "channel": lambda: configs.channel_case(n=(16, 12, 16), nparts=2),
            "config5_small": lambda: configs.config5(nx=48, ny=24, nz=48,
                                                     nparts=20)}[which]()
    configs.write_case(case, str(tmp_path))
    back = configs.read_case(str(tmp_path), noise=case.noise, seed=case.seed)
    a, b = dataclasses.asdict(case), dataclasses.asdict(back)
    for k in a:
        if k in ("implicit",):
            continue
        assert a[k] == b[k] or (isinstance(a[k], tuple) and
                                tuple(a[k]) == tuple(b[k])), k


@pytest.mark.parametrize("which", ["config1", "config2", "config3",
                                   "config5"])
def test_committed_inputs_are_the_seeded_configurations(which, tmp_path):
    """bluebottle_amd/inputs/<config>/input holds the seeded set-up of each
    benchmark configuration in the reference's grammar, byte for byte what
    `python -m bluebottle_amd.configs <config> ROOT` writes, and the
    driver's readers parse it back to the same case."""
    import os
    root = os.path.join(os.path.dirname(configs.__file__), "inputs", which)
    assert configs.main([which, str(tmp_path)]) == 0
    for name in ("flow.config", "part.config"):
        with open(os.path.join(root, "input", name), "rb") as f:
            committed = f.read()
        assert committed == (tmp_path / "input" / name).read_bytes(), name
    case = configs.CONFIGS[which]()
    back = configs.read_case(root, noise=case.noise, seed=case.seed)
    assert len(back.particles) == len(case.particles)
    for p, q in zip(case.particles, back.particles):
        assert (p.x, p.y, p.z, p.r) == (q.x, q.y, q.z, q.r)
    assert back.n == case.n and back.bc_u == case.bc_u


# ---------------------------------------------------------------------------
# restart
# ---------------------------------------------------------------------------
def _fill(s, rng):
    arrays = {}
    for name in ("u", "u0", "diff0_u", "conv0_u", "diff_u", "conv_u", "u_star",
                 "v", "v0", "diff0_v", "conv0_v", "diff_v", "conv_v", "v_star",
                 "w", "w0", "diff0_w", "conv0_w", "diff_w", "conv_w", "w_star",
                 "p", "phi", "p0"):
        a = s.array(name)
        a[:] = rng.standard_normal(a.size)
        arrays[name] = a.copy()
    for name in ("phase", "phase_shell", "flag_u", "flag_v", "flag_w"):
        a = s.array(name)
        a[:] = rng.integers(-5, 5, a.size)
        arrays[name] = a.copy()
    for c in ("pnm", "phinm", "chinm"):
        for part in ("re", "im"):
            for suf in ("", "0", "00"):
                nm = "%s_%s%s" % (c, part, suf)
                a = s.array(nm)
                a[:] = rng.standard_normal(a.size)
                arrays[nm] = a.copy()
    return arrays


def test_restart_layout_and_round_trip(tmp_path):
    case = configs.small_particle_case(n=8, nparts=2, radius_dx=1.5)
    s = sim.Simulation(case)
    rng = np.random.default_rng(5)
    arrays = _fill(s, rng)
    for k, v in (("ttime", 1.25), ("dt0", 0.01), ("dt", 0.02),
                 ("pid_int", 0.3), ("pid_back", -0.7)):
        runtime.set_scalar(k, v)
    runtime.set_scalar("stepnum", 17)
    gp = runtime.gvar("gradP", abi.gradP_struct)
    gp.z, gp.zm, gp.za = -3.5, 9.25, 4.0
    parts0 = bytes(bytearray(C.string_at(
        C.addressof(s.parts()[0]), C.sizeof(abi.part_struct) * 2)))
    path = str(tmp_path / "restart.config")
    assert s.host.bb_out_restart(path.encode()) == 0
    raw = open(path, "rb").read()
    d = s.dom
    gx, gy, gz, gc = d.Gfx.s3b, d.Gfy.s3b, d.Gfz.s3b, d.Gcc.s3b
    nc = 2 * runtime.get_scalar("coeff_stride", "i")
    size = (3 * 8 + 2 * 4 + 7 * 8 * (gx + gy + gz) + 3 * 8 * gc + 2 * 4 * gc
            + 4 * (gx + gy + gz) + 9 * 8 + 2 * C.sizeof(abi.part_struct) + 4
            + 18 * 8 * nc + 7 * 8 + 2 * 8)
    assert len(raw) == size
    # an independent reader of the documented order
    off = 0

    def take(dtype, count):
        nonlocal off
        a = np.frombuffer(raw, dtype=dtype, count=count, offset=off)
        off += a.nbytes
        return a
    assert tuple(take(np.float64, 3)) == (1.25, 0.01, 0.02)
    assert take(np.int32, 2)[0] == 17
    for comp, G in (("u", gx), ("v", gy), ("w", gz)):
        for nm in ("%s", "%s0", "diff0_%s", "conv0_%s", "diff_%s", "conv_%s",
                   "%s_star"):
            assert np.array_equal(take(np.float64, G), arrays[nm % comp])
    for nm in ("p", "phi", "p0"):
        assert np.array_equal(take(np.float64, gc), arrays[nm])
    for nm, G in (("phase", gc), ("phase_shell", gc), ("flag_u", gx),
                  ("flag_v", gy), ("flag_w", gz)):
        assert np.array_equal(take(np.int32, G), arrays[nm])
    take(np.float64, 9)
    assert bytes(take(np.uint8, 2 * C.sizeof(abi.part_struct))) == parts0
    assert take(np.int32, 1)[0] == nc // 2
    order = [("pnm_re", ""), ("pnm_im", ""), ("pnm_re", "0"), ("pnm_im", "0"),
             ("pnm_re", "00"), ("pnm_im", "00"), ("phinm_re", ""),
             ("phinm_im", ""), ("phinm_re", "0"), ("phinm_im", "0"),
             ("phinm_re", "00"), ("phinm_im", "00"), ("chinm_re", ""),
             ("chinm_im", ""), ("chinm_re", "0"), ("chinm_im", "0"),
             ("chinm_re", "00"), ("chinm_im", "00")]
    for nm, suf in order:
        assert np.array_equal(take(np.float64, nc), arrays[nm + suf])
    tail = take(np.float64, 7 + 2)
    assert tuple(tail[5:7]) == (0.3, -0.7)
    # gradP.z written with a count of 2: gradP.z then gradP.zm
    assert tuple(tail[7:9]) == (-3.5, 9.25)
    assert off == len(raw)
    # read back into cleared state
    for nm in arrays:
        s.array(nm)[:] = 0
    runtime.set_scalar("stepnum", 0)
    gp.z, gp.zm = 0.0, 1.0
    assert s.host.bb_in_restart(path.encode()) == 0
    assert runtime.get_scalar("stepnum", "i") == 17
    assert (gp.z, gp.zm) == (-3.5, 1.0)      # zm is not read back
    for nm, a in arrays.items():
        assert np.array_equal(s.array(nm), a), nm
    s.close()


def test_lamb_log(tmp_path):
    case = configs.small_particle_case(n=8, nparts=2, radius_dx=1.5, order=2)
    s = sim.Simulation(case)
    host = s.host
    buf = C.c_char * 256
    rd = buf.in_dll(host, "bb_root_dir")
    rd.value = str(tmp_path).encode()
    (tmp_path / "record").mkdir()
    for nm in ("pnm_re", "pnm_im", "phinm_re", "phinm_im", "chinm_re",
               "chinm_im"):
        a = s.array(nm)
        a[:] = np.arange(a.size) + 0.5
    runtime.set_scalar("ttime", 2.0)
    host.recorder_lamb(b"lamb.rec", 3)
    rd.value = b"."
    text = (tmp_path / "record" / "lamb.rec").read_text().splitlines()
    assert text[0] == "ttime = 2.000000e+00, iter = 3"
    assert text[1] == "particle 0"
    # order 2: (n, m) = (0,0) (1,0) (1,1) (2,0) (2,1) (2,2)
    rows = text[3:9]
    assert [tuple(int(v) for v in r.split()[:2]) for r in rows] == [
        (0, 0), (1, 0), (1, 1), (2, 0), (2, 1), (2, 2)]
    assert float(rows[1].split()[2]) == 1.5
    assert text[9] == "particle 1"
    s.close()


# ---------------------------------------------------------------------------
# record.config and the recorder files
# ---------------------------------------------------------------------------
def _rec(name, kind="d"):
    return runtime.get_scalar(name, kind)


def test_reference_record_config():
    """sim/input/record.config of the reference: every output interval."""
    import os
    host, _ = runtime.load()
    path = os.path.join(os.path.dirname(__file__), "golden", "ref_input",
                        "record.config")
    assert host.bb_read_record_config(path.encode()) == 0
    assert _rec("rec_flow_field_dt") == -1
    assert _rec("rec_particle_dt") == -0.01
    assert _rec("rec_paraview_dt") == 0.01
    assert _rec("rec_prec_flow_field_dt") == -0.01     # PRECURSOR
    assert _rec("rec_prec_dt") == -0.01                # PRECURSOR_VTK
    assert _rec("rec_restart_dt") == 1.0
    assert _rec("rec_restart_stop", "i") == 1
    assert _rec("rec_flow_field_vel", "i") == 0


def test_record_config_options(tmp_path):
    host, _ = runtime.load()
    p = tmp_path / "record.config"
    p.write_text("FLOW_FIELD 0.5\nvelocity\npressure\n\nPARTICLE 0.25\n"
                 "position\n\nRESTART 2\n\nPARAVIEW -1\n")
    assert host.bb_read_record_config(str(p).encode()) == 0
    assert (_rec("rec_flow_field_dt"), _rec("rec_flow_field_vel", "i"),
            _rec("rec_flow_field_p", "i")) == (0.5, 1, 1)
    assert (_rec("rec_particle_dt"), _rec("rec_particle_pos", "i")) == (0.25, 1)
    assert (_rec("rec_restart_dt"), _rec("rec_paraview_dt")) == (2.0, -1.0)
    # untouched outputs take their defaults (off)
    assert _rec("rec_prec_dt") == -1 and _rec("rec_restart_stop", "i") == 0
    # the reference accepts only "position" after PARTICLE
    p.write_text("PARTICLE 0.25\nvelocity\n\n")
    assert host.bb_read_record_config(str(p).encode()) == -1
    p.write_text("NOSUCH 1\n")
    assert host.bb_read_record_config(str(p).encode()) == -1
    # an option list running to the end of the file
    p.write_text("FLOW_FIELD 1\nvelocity")
    assert host.bb_read_record_config(str(p).encode()) == 0
    assert _rec("rec_flow_field_vel", "i") == 1


def test_recorder_init_files(tmp_path):
    host, _ = runtime.load()
    buf = C.c_char * 256
    buf.in_dll(host, "bb_root_dir").value = str(tmp_path).encode()
    (tmp_path / "record").mkdir()
    host.recorder_bicgstab_init(b"solver_expd.rec")
    host.recorder_lamb_init(b"lamb.rec")
    head = (tmp_path / "record" / "solver_expd.rec").read_text()
    assert head == "%-12s%-15s%-15s%-8s%-15s" % ("stepnum", "ttime", "dt",
                                                 "niter", "resid")
    assert (tmp_path / "record" / "lamb.rec").read_text() == ""
    buf.in_dll(host, "bb_root_dir").value = b"."


def test_reference_lid_driven_cavity_config(tmp_path):
    """flow.config-lid_driven_cavity: Dirichlet walls on four faces with the
    lid (north) moving at u = 10, Neumann pressure there, periodic z; like
    the sedimentation file it has no *_tDelay lines, so the g lines are
    lost to the "g_tDelay" attempt and the PID gains after them do not
    parse either: those keep their previous values."""
    runtime.set_scalar("Kp", 0.25)
    g = runtime.gvar("g", abi.g_struct)
    g.ym = 4.0
    _read_flow("flow.config-lid_driven_cavity", tmp_path)
    D = runtime.gvar("Dom", abi.dom_struct)
    assert (D.xs, D.xe, D.xn) == (-4.0, 4.0, 64)
    b = runtime.bc()
    for f in "WESN":
        assert getattr(b, "p" + f) == abi.NEUMANN
        for c in "uvw":
            assert getattr(b, c + f) == abi.DIRICHLET
    assert (b.pB, b.uT, b.wB) == (abi.PERIODIC,) * 3
    assert (b.uNDm, b.uNDa, b.uND) == (10.0, 0.0, 0.0)
    assert b.uSDm == 0.0 and b.vNDm == 0.0
    assert runtime.get_scalar("nu") == 1.0
    assert runtime.get_scalar("lamb_max_iter", "i") == 2000
    assert runtime.get_scalar("init_cond", "i") == abi.SHEAR
    assert runtime.get_scalar("out_plane", "i") == abi.NORTH
    gp = runtime.gvar("gradP", abi.gradP_struct)
    assert (gp.xm, gp.xa, gp.zm) == (0.0, 0.0, 0.0)
    assert g.ym == 4.0
    assert runtime.get_scalar("Kp") == 0.25
