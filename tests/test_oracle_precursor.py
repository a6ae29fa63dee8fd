"""Pins of the oracle's precursor-inflow pieces (the PRECURSOR ghost fill
of src/bluebottle_kernel.cu:228-1247 and the plane extraction of
src/bluebottle_kernel.cu:3619-3851).

* Ghost fill: every plane entry is tagged with its two transverse grid
  indices, so the checks fix each plane's layout (W/E j + k*jnb, S/N
  k + i*knb, B/T i + j*inb) and which values each face writes: the normal
  component's odd reflection about the plane value, a tangential
  component's ghost and first interior value (low face) or last interior
  value and ghost (high face) straight from its two planes; and the
  reference's particulars (cuda_dom_BC's v_WE_W for both west planes of
  v, cuda_dom_BC_star's south kernel on the north face of v).
* Plane extraction: on fields linear in every coordinate, linear
  interpolation is exact, so each plane must hold the field at the plane
  position (normal component, plus the given velocity) or half a cell
  either side of it (tangential components), except for the S/N w plane,
  which holds the slope term alone and writes it back into w, as the
  reference's `plane_n = w[M] = ...` does.
"""
import ctypes as C

import numpy as np
import pytest

from bluebottle_amd import abi, configs, sim
from oracle.pyoracle import OracleState

P, D, N, T = abi.PERIODIC, abi.DIRICHLET, abi.NEUMANN, abi.PRECURSOR


def precursor_case(n=(12, 10, 8), faces=(1, 1, 1, 1, 1, 1), hi=(1.2, 1.0, 0.8)):
    """Velocity faces flagged in `faces` (W, E, S, N, B, T) are PRECURSOR,
    the others periodic; p is Neumann on precursor faces."""
    bc_v = tuple((T, 0.0) if f else (P, 0.0) for f in faces)
    bc_p = tuple((N, 0.0) if f else (P, 0.0) for f in faces)
    return sim.Case(n=n, hi=hi, nu=0.01, init="quiescent", bc_p=bc_p,
                    bc_u=bc_v, bc_v=bc_v, bc_w=bc_v)


def configs18(case):
    c = [t for t, _ in case.bc_u] + [t for t, _ in case.bc_v] + \
        [t for t, _ in case.bc_w]
    return (C.c_int * 18)(*c)


def tagged(s, name):
    """plane value 1000 * (first transverse index) + (second), the indices
    in x, y, z order"""
    comp, axis = s.PLANES[name]
    G = s.grid(comp)
    ni, nj, nk = G.inb, G.jnb, G.knb
    if axis == 0:       # j + k*jnb
        j, k = np.meshgrid(np.arange(nj), np.arange(nk), indexing="xy")
        return (1000.0 * j + k).ravel() + 0.25
    if axis == 1:       # k + i*knb
        k, i = np.meshgrid(np.arange(nk), np.arange(ni), indexing="xy")
        return (1000.0 * i + k).ravel() + 0.5
    i, j = np.meshgrid(np.arange(ni), np.arange(nj), indexing="xy")
    return (1000.0 * i + j).ravel() + 0.75


@pytest.fixture()
def filled():
    case = precursor_case()
    s = sim.Simulation(case)
    rng = np.random.default_rng(3)
    for nm in ("u", "v", "w", "u_star", "v_star", "w_star", "p"):
        a = s.array(nm)
        a[:] = rng.standard_normal(a.size)
    for nm in s.PLANES:
        s.array(nm)[:] = tagged(s, nm)
    o = OracleState.from_sim(s)
    yield s, o
    s.close()


def _face_checks(s, before, after, comp, star):
    """the six faces of one component after the precursor fill, checked
    away from the edges (later faces overwrite the rows the earlier ones
    wrote there)"""
    G = s.grid(comp)
    F0 = before.reshape(G.knb, G.jnb, G.inb)
    F = after.reshape(G.knb, G.jnb, G.inb)
    ni, nj, nk = G.inb, G.jnb, G.knb
    I, J, K = slice(2, ni - 2), slice(2, nj - 2), slice(2, nk - 2)
    k, j, i = np.meshgrid(np.arange(nk), np.arange(nj), np.arange(ni),
                          indexing="ij")
    tWE = (1000.0 * j + k + 0.25)[:, :, 0]    # (k, j)
    tSN = (1000.0 * i + k + 0.5)[:, 0, :]     # (k, i)
    tBT = (1000.0 * i + j + 0.75)[0]          # (j, i)
    # x faces, rows J x K
    x = lambda a, c: a[K, J, c]
    if comp == "u":
        assert np.array_equal(x(F, 1), tWE[K, J])
        assert np.array_equal(x(F, 0), 2 * tWE[K, J] - x(F0, 2))
        assert np.array_equal(x(F, ni - 2), tWE[K, J])
        assert np.array_equal(x(F, ni - 1), 2 * tWE[K, J] - x(F0, ni - 3))
    else:
        for c in (0, 1, ni - 2, ni - 1):
            assert np.array_equal(x(F, c), tWE[K, J])
    # y faces, rows I x K
    y = lambda a, r: a[K, r, I]
    if comp == "v":
        assert np.array_equal(y(F, 1), tSN[K, I])
        assert np.array_equal(y(F, 0), 2 * tSN[K, I] - y(F0, 2))
        if star:
            # the north face gets the south kernel: the north ghost and the
            # last face keep their values
            assert np.array_equal(y(F, nj - 1), y(F0, nj - 1))
            assert np.array_equal(y(F, nj - 2), y(F0, nj - 2))
        else:
            assert np.array_equal(y(F, nj - 2), tSN[K, I])
            assert np.array_equal(y(F, nj - 1), 2 * tSN[K, I] - y(F0, nj - 3))
    else:
        for r in (0, 1, nj - 2, nj - 1):
            assert np.array_equal(y(F, r), tSN[K, I])
    # z faces, rows I x J
    z = lambda a, l: a[l, J, I]
    if comp == "w":
        assert np.array_equal(z(F, 1), tBT[J, I])
        assert np.array_equal(z(F, 0), 2 * tBT[J, I] - z(F0, 2))
        assert np.array_equal(z(F, nk - 2), tBT[J, I])
        assert np.array_equal(z(F, nk - 1), 2 * tBT[J, I] - z(F0, nk - 3))
    else:
        for l in (0, 1, nk - 2, nk - 1):
            assert np.array_equal(z(F, l), tBT[J, I])


def test_ghost_fill_layouts(filled):
    s, o = filled
    # the two planes of a tangential face carry the same tags, so the
    # west-face quirk of v does not show here (test_west_planes_of_v)
    before = {c: o.array(c).copy() for c in ("u", "v", "w")}
    o.call("dom_BC")
    for c in ("u", "v", "w"):
        _face_checks(s, before[c], o.array(c), c, star=False)
    before = {c: o.array(c + "_star").copy() for c in ("u", "v", "w")}
    o.call("dom_BC_star")
    for c in ("u", "v", "w"):
        _face_checks(s, before[c], o.array(c + "_star"), c, star=True)


def test_west_planes_of_v():
    """cuda_dom_BC gives v's west face v_WE_W for both planes,
    cuda_dom_BC_star gives it v_WE_W and v_WE_E."""
    case = precursor_case()
    s = sim.Simulation(case)
    s.array("v_WE_W")[:] = 1.0
    s.array("v_WE_E")[:] = 2.0
    o = OracleState.from_sim(s)
    s.close()
    o.call("dom_BC")
    o.call("dom_BC_star")
    G = case.n
    nk, nj, ni = G[2] + 2, G[1] + 3, G[0] + 2
    v = o.array("v").reshape(nk, nj, ni)
    vs = o.array("v_star").reshape(nk, nj, ni)
    # the interior j range: other faces overwrite the ghost rows
    assert (v[2:-2, 2:-2, 0] == 1.0).all() and (v[2:-2, 2:-2, 1] == 1.0).all()
    assert (vs[2:-2, 2:-2, 0] == 1.0).all() and (vs[2:-2, 2:-2, 1] == 2.0).all()
    assert (v[2:-2, 2:-2, -2] == 1.0).all() and (v[2:-2, 2:-2, -1] == 2.0).all()


def _linear(X, Y, Z, c):
    return c[0] + c[1] * X + c[2] * Y + c[3] * Z


def test_yank_linear_fields():
    n = (16, 16, 16)
    case = precursor_case(n=n, hi=(1.0, 1.0, 1.0))
    s = sim.Simulation(case)
    d = s.dom
    coef = {"u": (0.3, 1.1, -0.7, 0.4), "v": (-0.2, 0.5, 0.9, -1.3),
            "w": (0.1, -0.6, 0.2, 0.8)}
    off = {"u": (1.0, 0.5, 0.5), "v": (0.5, 1.0, 0.5), "w": (0.5, 0.5, 1.0)}

    def coords(comp):
        G = s.grid(comp)
        o_ = off[comp]
        x = d.xs + (np.arange(G.inb) - o_[0]) * d.dx
        y = d.ys + (np.arange(G.jnb) - o_[1]) * d.dy
        z = d.zs + (np.arange(G.knb) - o_[2]) * d.dz
        return np.meshgrid(z, y, x, indexing="ij")[::-1]

    for c in "uvw":
        X, Y, Z = coords(c)
        s.array(c)[:] = _linear(X, Y, Z, coef[c]).ravel()
    o = OracleState.from_sim(s)
    s.close()
    lib = o.lib
    lib.orc_yank_turb_planes.argtypes = [C.c_void_p, C.POINTER(C.c_int),
                                         C.POINTER(C.c_double),
                                         C.POINTER(C.c_double)]
    # W, S and B faces precursor (the low face of each pair)
    cfg = [T, P, T, P, T, P] * 3
    pos = (C.c_double * 9)(0.37, 0, 0, 0, 0.61, 0, 0, 0, 0.43)
    vel = (C.c_double * 18)(*([0.0] * 18))
    vel[0], vel[8], vel[16] = 0.125, -0.25, 0.5
    w_before = o.array("w").copy()
    lib.orc_yank_turb_planes(o.ptr, (C.c_int * 18)(*cfg), pos, vel)
    h = d.dx

    def plane(name):
        comp, axis = sim.Simulation.PLANES[name]
        G = {"u": d.Gfx, "v": d.Gfy, "w": d.Gfz}[comp]
        shape = ((G.knb, G.jnb), (G.inb, G.knb), (G.jnb, G.inb))[axis]
        return o.array(name).reshape(shape)

    def along(comp, axis, q):
        """the field's linear function at coordinate q on `axis`, over the
        plane's (row, column) grid"""
        X, Y, Z = coords(comp)
        if axis == 0:
            x = np.full_like(X[:, :, 0], q)
            return _linear(x, Y[:, :, 0], Z[:, :, 0], coef[comp])
        if axis == 1:
            y = np.full_like(X[:, 0, :], q)
            return _linear(X[:, 0, :], y, Z[:, 0, :], coef[comp]).T
        z = np.full_like(X[0], q)
        return _linear(X[0], Y[0], z, coef[comp])

    tol = 1e-12
    assert np.allclose(plane("u_WE"), along("u", 0, 0.37) + 0.125, atol=tol)
    assert np.allclose(plane("v_WE_W"), along("v", 0, 0.37 - h / 2), atol=tol)
    assert np.allclose(plane("v_WE_E"), along("v", 0, 0.37 + h / 2), atol=tol)
    assert np.allclose(plane("w_WE_W"), along("w", 0, 0.37 - h / 2), atol=tol)
    assert np.allclose(plane("w_WE_E"), along("w", 0, 0.37 + h / 2), atol=tol)
    assert np.allclose(plane("u_SN_S"), along("u", 1, 0.61 - h / 2), atol=tol)
    assert np.allclose(plane("u_SN_N"), along("u", 1, 0.61 + h / 2), atol=tol)
    assert np.allclose(plane("v_SN"), along("v", 1, 0.61) - 0.25, atol=tol)
    assert np.allclose(plane("w_SN_S"), along("w", 1, 0.61 - h / 2), atol=tol)
    assert np.allclose(plane("u_BT_B"), along("u", 2, 0.43 - h / 2), atol=tol)
    assert np.allclose(plane("u_BT_T"), along("u", 2, 0.43 + h / 2), atol=tol)
    assert np.allclose(plane("v_BT_B"), along("v", 2, 0.43 - h / 2), atol=tol)
    assert np.allclose(plane("v_BT_T"), along("v", 2, 0.43 + h / 2), atol=tol)
    # w's row j of the S/N plane is overwritten before the B/T extraction
    # reads it (see below)
    j = int(np.floor(0.61 / d.dy)) + 1
    rows = np.arange(d.Gfz.jnb) != j
    assert np.allclose(plane("w_BT")[rows], (along("w", 2, 0.43) + 0.5)[rows],
                       atol=tol)
    # the S/N w plane: the slope term alone, (dw/dy) (pos + dy/2 - y_n),
    # also stored into w at the plane's cell, whose index is that of pos
    y_n = (j - 1 + 0.5) * d.dy
    slope = coef["w"][2] * (0.61 + d.dy / 2 - y_n)
    assert np.allclose(plane("w_SN_N"), slope, atol=tol)
    w = o.array("w").reshape(d.Gfz.knb, d.Gfz.jnb, d.Gfz.inb)
    wb = w_before.reshape(w.shape)
    assert np.allclose(w[:, j, :], slope, atol=tol)
    changed = np.ones(w.shape[1], bool)
    changed[j] = False
    assert np.array_equal(w[:, changed, :], wb[:, changed, :])
