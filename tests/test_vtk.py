"""ParaView output in the reference's XML VTK layout (src/vtk.c), read back
with an XML parser:
* out_N.vtr: p and phase in k, j, i order over the interior, the face
  velocities averaged to cell centres, node coordinates of the grid;
* out_N.vtp: one point per particle plus its periodic images, in the
  reference's order (single faces, xy/zy/xz edges, corners, the particle
  itself), and F = Fh + Fi + Fk + buoyant weight;
* out_nodes_N.vtp: the 26 Lebedev directions at r = rs around every point,
  wrapped back into the box;
* out.pvd: a valid collection that grows by three DataSets per output.
"""
import math
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from bluebottle_amd import abi, runtime, sim


def _case():
    parts = [sim.Particle(x=1.0, y=0.75, z=0.5, r=0.2),
             sim.Particle(x=0.1, y=0.75, z=0.5, r=0.2),
             sim.Particle(x=1.9, y=1.4, z=0.05, r=0.2, rho=3.0)]
    return sim.Case(n=(8, 6, 4), hi=(2.0, 1.5, 1.0), nu=0.01, init="taylor_green",
                    noise=1e-3, seed=4, particles=parts,
                    g=(0.0, 0.0, 0.0, 0.0, -2.0, 0.0))


def _arrays(root):
    out = {}
    for da in root.iter("DataArray"):
        vals = da.text.split()
        out[da.get("Name")] = vals
    return out


@pytest.fixture()
def written(tmp_path):
    s = sim.Simulation(_case())
    rng = np.random.default_rng(11)
    for nm in ("p", "u", "v", "w"):
        a = s.array(nm)
        a[:] = rng.standard_normal(a.size)
    ph = s.array("phase")
    ph[:] = rng.integers(-1, 3, ph.size)
    parts = s.parts()
    for q, P in enumerate(parts):
        for k, fld in enumerate(("u", "v", "w", "udot", "ox", "Fx", "Fy", "Fz",
                                 "iFx", "iFz", "kFz", "Lx", "iLx")):
            setattr(P, fld, 0.125 * (q + 1) + 0.01 * k)
        for m in range(26):
            P.nodes[m] = (m % 3) - 1 + 10 * q
    gg = runtime.gvar("g", abi.g_struct)
    gg.z = -2.0
    runtime.set_scalar("ttime", 0.5)
    d = str(tmp_path)
    assert s.host.bb_vtk_init(d.encode()) == 0
    assert s.host.bb_vtk_out(d.encode(), 0) == 0
    runtime.set_scalar("ttime", 0.75)
    assert s.host.bb_vtk_out(d.encode(), 1) == 0
    yield s, tmp_path
    s.close()


def test_grid_file(written):
    s, d = written
    root = ET.parse(d / "out_1_0.vtr").getroot()
    assert root.get("type") == "RectilinearGrid"
    piece = root.find("RectilinearGrid/Piece")
    assert piece.get("Extent") == "0 8 0 6 0 4"
    A = _arrays(root)
    D = s.dom
    gc = D.Gcc
    p3 = s.array("p").reshape(gc.knb, gc.jnb, gc.inb)[1:-1, 1:-1, 1:-1]
    assert [float(x) for x in A["p"]] == [float("%e" % x) for x in p3.ravel()]
    ph3 = s.array("phase").reshape(gc.knb, gc.jnb, gc.inb)[1:-1, 1:-1, 1:-1]
    assert [int(x) for x in A["phase"]] == list(ph3.ravel())
    u = s.field3("u")[1:-1, 1:-1, :]
    v = s.field3("v")[1:-1, :, 1:-1]
    w = s.field3("w")[:, 1:-1, 1:-1]
    # cell (i) of the interior averages faces i and i+1 of the ghosted grid
    uc = 0.5 * (u[:, :, 1:-2] + u[:, :, 2:-1])
    vc = 0.5 * (v[:, 1:-2, :] + v[:, 2:-1, :])
    wc = 0.5 * (w[1:-2] + w[2:-1])
    vel = np.array([float(x) for x in A["vel"]]).reshape(-1, 3)
    for c, ref in enumerate((uc, vc, wc)):
        assert np.array_equal(vel[:, c], [float("%e" % x) for x in ref.ravel()])
    assert np.allclose([float(x) for x in A["x"]], np.arange(9) * 0.25)
    assert np.allclose([float(x) for x in A["y"]], np.arange(7) * 0.25)
    assert np.allclose([float(x) for x in A["z"]], np.arange(5) * 0.25)
    hdr = ET.parse(d / "out_1.pvtr").getroot()
    pc = hdr.find("PRectilinearGrid/Piece")
    assert (pc.get("Extent"), pc.get("Source")) == ("0 8 0 6 0 4",
                                                    "out_1_0.vtr")


# the points of out_N.vtp for _case(): particle index, centre
EXPECTED = [
    (0, (1.0, 0.75, 0.5)),
    (1, (2.1, 0.75, 0.5)),          # x low face: +Lx
    (1, (0.1, 0.75, 0.5)),
    (2, (-0.1, 1.4, 0.05)),         # x high: -Lx
    (2, (1.9, -0.1, 0.05)),         # y high: -Ly
    (2, (1.9, 1.4, 1.05)),          # z low: +Lz
    (2, (-0.1, -0.1, 0.05)),        # xy edge
    (2, (1.9, -0.1, 1.05)),         # zy edge
    (2, (-0.1, 1.4, 1.05)),         # xz edge
    (2, (-0.1, -0.1, 1.05)),        # corner
    (2, (1.9, 1.4, 0.05)),
]


def test_particle_file(written):
    s, d = written
    root = ET.parse(d / "out_0.vtp").getroot()
    piece = root.find("PolyData/Piece")
    assert int(piece.get("NumberOfPoints")) == len(EXPECTED)
    pts = np.array([float(x) for x in
                    piece.find("Points/DataArray").text.split()]).reshape(-1, 3)
    A = _arrays(piece.find("PointData"))
    assert [int(x) for x in A["n"]] == [e[0] for e in EXPECTED]
    assert np.allclose(pts, [e[1] for e in EXPECTED], atol=1e-6)
    parts = s.parts()
    F = np.array([float(x) for x in A["F"]]).reshape(-1, 3)
    Fh = np.array([float(x) for x in A["Fh"]]).reshape(-1, 3)
    for row, (q, _) in enumerate(EXPECTED):
        P = parts[q]
        m = 4.0 / 3.0 * math.pi * (P.rho - 1.0) * P.r ** 3
        assert F[row, 2] == pytest.approx(P.Fz + P.iFz + P.kFz + m * -2.0,
                                          abs=1e-6)
        assert F[row, 0] == pytest.approx(P.Fx + P.iFx + P.kFx, abs=1e-6)
        assert tuple(Fh[row]) == tuple(float("%f" % x)
                                       for x in (P.Fx, P.Fy, P.Fz))
    assert [float(x) for x in A["r"]] == [0.2] * len(EXPECTED)


def _lebedev_directions():
    dirs = []
    for a in range(3):
        for sgn in (1, -1):
            e = np.zeros(3)
            e[a] = sgn
            dirs.append(e)
    for a, b in ((0, 1), (0, 2), (1, 2)):
        for sa in (1, -1):
            for sb in (1, -1):
                e = np.zeros(3)
                e[a], e[b] = sa, sb
                dirs.append(e / math.sqrt(2))
    for sx in (1, -1):
        for sy in (1, -1):
            for sz in (1, -1):
                dirs.append(np.array([sx, sy, sz]) / math.sqrt(3))
    return np.array(dirs)


def test_node_file(written):
    s, d = written
    nroot = ET.parse(d / "out_nodes_0.vtp").getroot()
    piece = nroot.find("PolyData/Piece")
    assert int(piece.get("NumberOfPoints")) == 26 * len(EXPECTED)
    pts = np.array([float(x) for x in
                    piece.find("Points/DataArray").text.split()]).reshape(-1, 26, 3)
    A = _arrays(piece.find("PointData"))
    n = np.array([int(x) for x in A["n"]]).reshape(-1, 26)
    status = np.array([int(x) for x in A["status"]]).reshape(-1, 26)
    dirs = _lebedev_directions()
    L = np.array([2.0, 1.5, 1.0])
    parts = s.parts()
    for row, (q, c) in enumerate(EXPECTED):
        assert (n[row] == q).all()
        assert list(status[row]) == [(m % 3) - 1 + 10 * q for m in range(26)]
        rs = parts[q].rs
        # undo the wrap into the box, then every node is c + rs * direction
        rel = pts[row] - np.array(c)
        rel -= L * np.round(rel / L)
        got = rel / rs
        assert np.allclose(np.linalg.norm(got, axis=1), 1.0, atol=1e-4)
        # the 26 directions of the degree-7 rule, each once
        match = np.abs(got[:, None, :] - dirs[None]).max(axis=2) < 1e-4
        assert (match.sum(axis=1) == 1).all() and (match.sum(axis=0) == 1).all()
        # wrapped into [start, end] on each axis (z tested against x's end)
        assert (pts[row][:, 0] >= -1e-6).all() and (pts[row][:, 0] <= 2 + 1e-6).all()
        assert (pts[row][:, 1] >= -1e-6).all() and (pts[row][:, 1] <= 1.5 + 1e-6).all()
        assert (pts[row][:, 2] >= -1e-6).all()
    # node order: +x, +y, -x, -y, +z, -z first (theta = pi/2 at phi = 0,
    # pi/2, pi, 3pi/2, then the poles)
    rel = pts[0] - np.array(EXPECTED[0][1])
    first = rel[:6] / parts[0].rs
    assert np.allclose(first, [[1, 0, 0], [0, 1, 0], [-1, 0, 0], [0, -1, 0],
                               [0, 0, 1], [0, 0, -1]], atol=1e-5)


def test_collection(written):
    s, d = written
    text = (d / "out.pvd").read_text()
    root = ET.fromstring(text)
    ds = root.findall("Collection/DataSet")
    assert [(x.get("timestep"), x.get("part"), x.get("file")) for x in ds] == [
        ("5.000000e-01", "0", "out_0.pvtr"), ("5.000000e-01", "1", "out_0.vtp"),
        ("5.000000e-01", "2", "out_nodes_0.vtp"),
        ("7.500000e-01", "0", "out_1.pvtr"), ("7.500000e-01", "1", "out_1.vtp"),
        ("7.500000e-01", "2", "out_nodes_1.vtp")]
    assert text.endswith("</Collection>\n</VTKFile>")


def test_precursor_domain_files(tmp_path):
    """out_turb_N.pvtr/.vtr carry p and vel only, the header's piece extent
    truncates (0.9 / (0.9 / 7) = 6.999... gives 6 cells in x, where the
    main header rounds to 7), the piece file truncates as the main one
    does, and out_turb.pvd gains one DataSet per output."""
    case = sim.Case(n=(7, 5, 4), hi=(0.9, 1.0, 0.8), nu=0.01,
                    init="taylor_green", noise=1e-3, seed=2)
    s = sim.Simulation(case)
    try:
        rng = np.random.default_rng(3)
        for nm in ("p", "u", "v", "w"):
            a = s.array(nm)
            a[:] = rng.standard_normal(a.size)
        d = str(tmp_path).encode()
        runtime.set_scalar("ttime", 0.25)
        assert s.host.bb_vtk_turb_init(d) == 0
        assert s.host.bb_vtk_turb_out(d, 3) == 0
        runtime.set_scalar("ttime", 0.5)
        assert s.host.bb_vtk_turb_out(d, 4) == 0
        assert s.host.bb_vtk_grid(d, 4) == 0
        root = ET.parse(tmp_path / "out_turb_4_0.vtr").getroot()
        cd = root.find("RectilinearGrid/Piece/CellData")
        assert cd.get("Scalars") == "p"
        assert root.find("RectilinearGrid/Piece").get("Extent") == "0 6 0 5 0 4"
        A = _arrays(root)
        assert sorted(A) == ["p", "vel", "x", "y", "z"]
        gc = s.dom.Gcc
        p3 = s.array("p").reshape(gc.knb, gc.jnb, gc.inb)[1:-1, 1:-1, 1:7]
        assert [float(x) for x in A["p"]] == [float("%e" % x)
                                              for x in p3.ravel()]
        u = s.field3("u")[1:-1, 1:-1, :]
        uc = 0.5 * (u[:, :, 1:7] + u[:, :, 2:8])
        vel = np.array([float(x) for x in A["vel"]]).reshape(-1, 3)
        assert np.array_equal(vel[:, 0], [float("%e" % x) for x in uc.ravel()])
        assert len(A["x"]) == 7
        hdr = ET.parse(tmp_path / "out_turb_4.pvtr").getroot()
        names = [x.get("Name") for x in hdr.iter("PDataArray")]
        assert names == ["p", "vel", "x", "y", "z"]
        pc = hdr.find("PRectilinearGrid/Piece")
        assert pc.get("Extent") == "0 6 0 5 0 4"
        assert pc.get("Source") == "out_turb_4_0.vtr"
        main = ET.parse(tmp_path / "out_4.pvtr").getroot()
        assert main.find("PRectilinearGrid/Piece").get("Extent") == "0 7 0 5 0 4"
        ds = ET.parse(tmp_path / "out_turb.pvd").getroot().findall(
            "Collection/DataSet")
        assert [(x.get("timestep"), x.get("part"), x.get("file"))
                for x in ds] == [("2.500000e-01", "0", "out_turb_3.pvtr"),
                                 ("5.000000e-01", "0", "out_turb_4.pvtr")]
    finally:
        s.close()


def test_ghost_cell_output(tmp_path):
    """out_ghost_N_0.vtr (src/vtk.c:612-933) covers every cell of the
    ghosted grid in k, j, i order: p, phase, phase_shell, the face values of
    u, u* and the face flags averaged to the cell centres (ghost faces
    included), coordinates from one cell below the domain; out_ghost.pvd
    gains the grid and the particle file."""
    case = sim.Case(n=(5, 4, 3), hi=(1.25, 1.0, 0.75), nu=0.01,
                    init="taylor_green", noise=1e-3, seed=2,
                    particles=[sim.Particle(x=0.6, y=0.5, z=0.4, r=0.2)])
    s = sim.Simulation(case)
    try:
        rng = np.random.default_rng(5)
        for nm in ("p", "u", "v", "w", "u_star", "v_star", "w_star"):
            a = s.array(nm)
            a[:] = rng.standard_normal(a.size)
        for nm in ("phase", "phase_shell", "flag_u", "flag_v", "flag_w"):
            a = s.array(nm)
            a[:] = rng.integers(-1, 3, a.size)
        d = str(tmp_path).encode()
        runtime.set_scalar("ttime", 0.125)
        assert s.host.bb_vtk_ghost_init(d) == 0
        assert s.host.bb_vtk_ghost_out(d, 2) == 0
        root = ET.parse(tmp_path / "out_ghost_2_0.vtr").getroot()
        assert root.find("RectilinearGrid").get("WholeExtent") == "0 7 0 6 0 5"
        assert root.find("RectilinearGrid/Piece").get("Extent") == "0 7 0 6 0 5"
        A = _arrays(root)
        assert [float(x) for x in A["p"]] == [float("%e" % x)
                                              for x in s.array("p")]
        assert [int(x) for x in A["phase_shell"]] == list(s.array("phase_shell"))
        D = s.dom
        gc = D.Gcc
        nc = gc.knb * gc.jnb * gc.inb

        def centres(xn, yn, zn):
            x = s.field3(xn)
            y = s.field3(yn)
            z = s.field3(zn)
            return (0.5 * (x[:, :, :-1] + x[:, :, 1:]),
                    0.5 * (y[:, :-1, :] + y[:, 1:, :]),
                    0.5 * (z[:-1] + z[1:]))

        for name, comps, fmt in (("vel", ("u", "v", "w"), "%e"),
                                 ("vel_star", ("u_star", "v_star", "w_star"),
                                  "%e")):
            got = np.array([float(x) for x in A[name]]).reshape(nc, 3)
            for c, ref in enumerate(centres(*comps)):
                assert ref.shape == (gc.knb, gc.jnb, gc.inb)
                assert np.array_equal(got[:, c], [float(fmt % x)
                                                  for x in ref.ravel()])
        fl = np.array([float(x) for x in A["flag"]]).reshape(nc, 3)
        fu = s.array("flag_u").reshape(gc.knb, gc.jnb, gc.inb + 1)
        assert np.array_equal(fl[:, 0], (0.5 * (fu[:, :, :-1] + fu[:, :, 1:])).ravel())
        assert np.allclose([float(x) for x in A["x"]], -0.25 + 0.25 * np.arange(8))
        assert np.allclose([float(x) for x in A["z"]], -0.25 + 0.25 * np.arange(6))
        ds = ET.parse(tmp_path / "out_ghost.pvd").getroot().findall(
            "Collection/DataSet")
        assert [(x.get("part"), x.get("file")) for x in ds] == [
            ("0", "out_ghost_2.pvtr"), ("1", "out_2.vtp")]
        assert (tmp_path / "out_2.vtp").exists()
        hdr = ET.parse(tmp_path / "out_ghost_2.pvtr").getroot()
        assert [x.get("Name") for x in hdr.iter("PDataArray")] == [
            "p", "phase", "phase_shell", "vel", "vel_star", "flag", "x", "y", "z"]
    finally:
        s.close()
