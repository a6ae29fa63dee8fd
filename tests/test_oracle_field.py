"""Pins of the oracle's field kernels against closed forms.

Each check fixes a piece of the discretisation by what the mathematics
says it must give, so that a dropped term, a wrong sign or a shifted index
in the oracle fails:
  * ghost-cell BCs (src/bluebottle_kernel.cu:26-1248): periodic copies,
    the odd reflection of the normal Dirichlet face, the quadratic wall
    extrapolation 8/3 bc - 2 u1 + u2/3 of tangential Dirichlet faces and the
    Neumann copy reproduce analytic profiles with those symmetries exactly;
  * the predictor (src/bluebottle_kernel.cu:1504-2003): its divergence-form
    convection is exact on linear velocity fields, its Laplacian on
    quadratic ones, and the Euler / Adams-Bashforth 2 combination with
    the incremental pressure gradient follows the textbook formulas;
  * the Poisson operator (src/bicgstab_kernel.cu:1654-1856) has the
    discrete eigenvalue -(4/h^2) sum sin^2(pi k h) on periodic modes and is
    symmetric; BiCGSTAB meets its tolerance on the true residual;
  * pressure solve + projection leave a discretely divergence-free field
    and a zero-mean phi; update_p, find_dt and solvability follow their
    closed forms.
"""
import math

import numpy as np
import pytest

from bluebottle_amd import abi, configs, sim
from oracle.pyoracle import OracleState

DOM_BUF = 1


def coords(d, G, which):
    """Positions of every (ghosted) point of grid G, as (k, j, i) arrays.
    Cell centres sit at xs + (i - 1/2) dx, faces at xs + (i - 1) dx."""
    off = {"c": (0.5, 0.5, 0.5), "u": (1.0, 0.5, 0.5), "v": (0.5, 1.0, 0.5),
           "w": (0.5, 0.5, 1.0)}[which]
    x = d.xs + (np.arange(G.inb) - off[0]) * d.dx
    y = d.ys + (np.arange(G.jnb) - off[1]) * d.dy
    z = d.zs + (np.arange(G.knb) - off[2]) * d.dz
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    return X, Y, Z


def lim(G):
    """(is, ie, js, je, ks, ke) of a grid (`is` is a Python keyword)."""
    return (getattr(G, "is"), G.ie, G.js, G.je, G.ks, G.ke)


def interior(G, a):
    a3 = a.reshape(G.knb, G.jnb, G.inb)
    i0, i1, j0, j1, k0, k1 = lim(G)
    return a3[k0:k1, j0:j1, i0:i1]


def state(case):
    """Oracle state of a freshly set-up case, with its cages built (the
    set-up marks every cell phase 1, src/particle.c:559-571; build_cages
    gives the fluid its -1)."""
    s = sim.Simulation(case)
    o = OracleState.from_sim(s)
    d = s.dom
    s.close()
    o.call("build_cages")
    return o, d


# ---------------------------------------------------------------------------
# boundary conditions
# ---------------------------------------------------------------------------
def test_bc_channel_profiles():
    case = configs.channel_case(n=(12, 10, 8))
    o, d = state(case)
    uS, uN = case.bc_u[2][1], case.bc_u[3][1]
    Lx, Ly, Lz = d.xl, d.yl, d.zl

    def tang(X, Y, Z, lo, hi):
        eta = (Y - d.ys) / Ly
        yy = Y - d.ys
        return lo + (hi - lo) * eta + yy * (yy - Ly) * (
            0.7 + 0.3 * np.sin(2 * np.pi * (X - d.xs) / Lx)
            * np.cos(2 * np.pi * (Z - d.zs) / Lz))

    exact = {}
    X, Y, Z = coords(d, d.Gfx, "u")
    exact["u"] = tang(X, Y, Z, uS, uN)
    X, Y, Z = coords(d, d.Gfz, "w")
    exact["w"] = tang(X, Y, Z, 0.0, 0.0)
    X, Y, Z = coords(d, d.Gfy, "v")
    exact["v"] = np.sin(np.pi * (Y - d.ys) / Ly) * (
        1 + 0.5 * np.cos(2 * np.pi * (X - d.xs) / Lx)) * np.sin(
        2 * np.pi * (Z - d.zs) / Lz + 0.3)
    X, Y, Z = coords(d, d.Gcc, "c")
    exact["p"] = np.cos(np.pi * (Y - d.ys) / Ly) * np.cos(
        2 * np.pi * (X - d.xs) / Lx + 0.2) * np.cos(2 * np.pi * (Z - d.zs) / Lz)
    grids = {"u": d.Gfx, "v": d.Gfy, "w": d.Gfz, "p": d.Gcc}
    for name, G in grids.items():
        a = o.array(name)
        a[:] = np.nan
        interior(G, a)[...] = interior(G, exact[name].ravel())
        if name == "v":
            # the wall faces are interior indices; the BC overwrites them
            a3 = a.reshape(G.knb, G.jnb, G.inb)
            a3[:, G.js, :] = 12.0
            a3[:, G.je - 1, :] = -7.0
    o.call("dom_BC")
    for name, G in grids.items():
        got = o.array(name).reshape(G.knb, G.jnb, G.inb)
        want = exact[name]
        # every point of the ghosted array, except the 12 edge lines where
        # two ghost layers meet (filled in W,E,S,N,B,T order from ghosts)
        k, j, i = np.meshgrid(np.arange(G.knb), np.arange(G.jnb),
                              np.arange(G.inb), indexing="ij")
        ghost = ((i == 0) | (i == G.inb - 1)).astype(int) + \
            ((j == 0) | (j == G.jnb - 1)).astype(int) + \
            ((k == 0) | (k == G.knb - 1)).astype(int)
        sel = ghost <= 1
        err = np.abs(got[sel] - want[sel]).max()
        assert err < 1e-12, (name, err)


def test_bc_periodic_profiles():
    case = configs.periodic_case(10, 0.01)
    o, d = state(case)
    L = d.xl
    for name, G, which in (("u", d.Gfx, "u"), ("v", d.Gfy, "v"),
                           ("w", d.Gfz, "w"), ("p", d.Gcc, "c")):
        X, Y, Z = coords(d, G, which)
        ex = np.sin(2 * np.pi * X / L + 0.1) * np.cos(2 * np.pi * Y / L) \
            + np.cos(2 * np.pi * Z / L + 0.4)
        a = o.array(name)
        a[:] = np.nan
        interior(G, a)[...] = interior(G, ex.ravel())
        if name in ("u", "v", "w"):
            # on a periodic axis the low wall face is rebuilt from the
            # high one (the same plane of the periodic box)
            a3 = a.reshape(G.knb, G.jnb, G.inb)
            i0, _, j0, _, k0, _ = lim(G)
            if name == "u":
                a3[:, :, i0] = 99.0
            elif name == "v":
                a3[:, j0, :] = 99.0
            else:
                a3[k0, :, :] = 99.0
    o.call("dom_BC")
    for name, G, which in (("u", d.Gfx, "u"), ("v", d.Gfy, "v"),
                           ("w", d.Gfz, "w"), ("p", d.Gcc, "c")):
        X, Y, Z = coords(d, G, which)
        ex = np.sin(2 * np.pi * X / L + 0.1) * np.cos(2 * np.pi * Y / L) \
            + np.cos(2 * np.pi * Z / L + 0.4)
        got = o.array(name).reshape(G.knb, G.jnb, G.inb)
        assert np.abs(got - ex).max() < 1e-12, name


# ---------------------------------------------------------------------------
# predictor
# ---------------------------------------------------------------------------
def _set_fields(o, d, fu, fv, fw, fp):
    for name, G, which, f in (("u0", d.Gfx, "u", fu), ("v0", d.Gfy, "v", fv),
                              ("w0", d.Gfz, "w", fw), ("p0", d.Gcc, "c", fp)):
        X, Y, Z = coords(d, G, which)
        o.array(name)[:] = f(X, Y, Z).ravel()


LIN = {
    "u": (0.3, 0.7, -0.4, 0.2), "v": (-0.1, 0.25, 0.5, -0.6),
    "w": (0.2, -0.3, 0.15, 0.45),
}


def _lin(c):
    a0, ax, ay, az = LIN[c]
    return lambda X, Y, Z: a0 + ax * X + ay * Y + az * Z


def _grad(c):
    return np.array(LIN[c][1:])


def test_predictor_euler_on_linear_fields():
    case = configs.periodic_case(8, 0.02)
    o, d = state(case)
    o.s.dt = 1e-3
    o.s.dt0 = -1.0
    pg = (0.11, -0.07, 0.05)
    _set_fields(o, d, _lin("u"), _lin("v"), _lin("w"),
                lambda X, Y, Z: 1.0 + pg[0] * X + pg[1] * Y + pg[2] * Z)
    for fname, G in (("f_x", d.Gfx), ("f_y", d.Gfy), ("f_z", d.Gfz)):
        o.array(fname)[:] = 0.01
    o.call("U_star_2")
    vel = {c: _lin(c) for c in "uvw"}
    grads = {c: _grad(c) for c in "uvw"}
    for comp, G, which, axis in (("u", d.Gfx, "u", 0), ("v", d.Gfy, "v", 1),
                                 ("w", d.Gfz, "w", 2)):
        X, Y, Z = coords(d, G, which)
        U = {c: vel[c](X, Y, Z) for c in "uvw"}
        # d(u_a u_b)/dx_b = u_b d u_a/dx_b + u_a div u
        div = grads["u"][0] + grads["v"][1] + grads["w"][2]
        conv = sum(U[b] * grads[comp][ib] for ib, b in enumerate("uvw")) \
            + U[comp] * div
        got_c = interior(G, o.array("conv_" + comp))
        assert np.abs(got_c - interior(G, conv.ravel())).max() < 1e-12
        assert np.abs(interior(G, o.array("diff_" + comp))).max() < 1e-9
        ustar = U[comp] + o.s.dt * (-pg[axis] / o.s.rho_f - conv + 0.01)
        got = interior(G, o.array(comp + "_star"))
        assert np.abs(got - interior(G, ustar.ravel())).max() < 1e-13


def test_predictor_diffusion_on_quadratic_fields():
    case = configs.periodic_case(8, 0.03)
    o, d = state(case)
    o.s.dt = 1e-3
    o.s.dt0 = -1.0
    q = {"u": (0.5, -0.2, 0.3), "v": (0.1, 0.4, -0.6), "w": (-0.3, 0.2, 0.7)}
    fs = {c: (lambda a: (lambda X, Y, Z: a[0] * X * X + a[1] * Y * Y
                          + a[2] * Z * Z + 0.3 * X * Y))(q[c]) for c in "uvw"}
    _set_fields(o, d, fs["u"], fs["v"], fs["w"], lambda X, Y, Z: 0 * X)
    o.call("U_star_2")
    for comp, G in (("u", d.Gfx), ("v", d.Gfy), ("w", d.Gfz)):
        lap = 2.0 * sum(q[comp])
        got = interior(G, o.array("diff_" + comp))
        assert np.abs(got - o.s.nu * lap).max() < 1e-9 * max(1, o.s.nu * lap)


def test_predictor_ab2_combination():
    """Variable-step AB2: u* = u0 + dt [ -dp/dx / rho
       - (1 + dt/(2 dt0)) c + dt/(2 dt0) c0 + (1 + dt/(2 dt0)) d
       - dt/(2 dt0) d0 + f ]"""
    case = configs.periodic_case(8, 0.02)
    o, d = state(case)
    o.s.dt = 2e-3
    o.s.dt0 = 1.5e-3
    rng = np.random.default_rng(5)
    _set_fields(o, d, _lin("u"), _lin("v"), _lin("w"), lambda X, Y, Z: 0 * X)
    for comp, G in (("u", d.Gfx), ("v", d.Gfy), ("w", d.Gfz)):
        o.array("conv0_" + comp)[:] = rng.normal(size=G.s3b)
        o.array("diff0_" + comp)[:] = rng.normal(size=G.s3b)
    o.call("U_star_2")
    r = o.s.dt / (2 * o.s.dt0)
    for comp, G in (("u", d.Gfx), ("v", d.Gfy), ("w", d.Gfz)):
        c, df = o.array("conv_" + comp), o.array("diff_" + comp)
        c0, d0 = o.array("conv0_" + comp), o.array("diff0_" + comp)
        u0 = o.array(comp + "0")
        want = u0 + o.s.dt * (-(1 + r) * c + r * c0 + (1 + r) * df - r * d0)
        got = interior(G, o.array(comp + "_star"))
        assert np.abs(got - interior(G, want)).max() < 1e-12


def test_predictor_masks_solid_faces():
    case = configs.small_particle_case(n=16, nparts=1, radius_dx=3.0)
    o, d = state(case)
    o.call("build_cages")
    o.s.dt = 1e-3
    o.s.dt0 = -1.0
    _set_fields(o, d, _lin("u"), _lin("v"), _lin("w"), lambda X, Y, Z: 0 * X)
    o.call("U_star_2")
    ph = o.array("phase").reshape(d.Gcc.knb, d.Gcc.jnb, d.Gcc.inb)
    us = o.array("u_star").reshape(d.Gfx.knb, d.Gfx.jnb, d.Gfx.inb)
    G = d.Gfx
    nsolid = 0
    for k in range(G.ks, G.ke):
        for j in range(G.js, G.je):
            for i in range(lim(G)[0], G.ie):
                solid = ph[k, j, i - 1] >= 0 or ph[k, j, i] >= 0
                if solid:
                    nsolid += 1
                    assert us[k, j, i] == 0.0
                else:
                    assert us[k, j, i] != 0.0
    assert nsolid > 0


# ---------------------------------------------------------------------------
# pressure Poisson operator and solver
# ---------------------------------------------------------------------------
def _matvec(o, x):
    import ctypes as C
    y = np.zeros_like(x)
    o.lib.orc_pp_matvec(o.ptr, x.ctypes.data_as(C.POINTER(C.c_double)),
                        y.ctypes.data_as(C.POINTER(C.c_double)))
    return y


def test_poisson_operator_eigenmode():
    n = 12
    case = configs.periodic_case(n, 0.01)
    o, d = state(case)
    o.call("build_cages")
    h = d.dx
    xc = (np.arange(n) + 0.5) * h
    Z, Y, X = np.meshgrid(xc, xc, xc, indexing="ij")
    for kx, ky, kz in ((1, 1, 1), (2, 0, 3), (0, 5, 1)):
        phi = np.cos(2 * np.pi * kx * X) * np.cos(2 * np.pi * ky * Y) * \
            np.cos(2 * np.pi * kz * Z)
        lam = -(4 / h ** 2) * (math.sin(math.pi * kx * h) ** 2
                               + math.sin(math.pi * ky * h) ** 2
                               + math.sin(math.pi * kz * h) ** 2)
        y = _matvec(o, phi.ravel().copy())
        assert np.abs(y - lam * phi.ravel()).max() < 1e-9 * abs(lam)


def test_poisson_operator_symmetric_and_constant_null_space():
    case = configs.channel_case(n=(10, 8, 6))
    o, d = state(case)
    o.call("build_cages")
    n = d.xn * d.yn * d.zn
    rng = np.random.default_rng(2)
    x, y = rng.normal(size=n), rng.normal(size=n)
    assert abs(np.dot(_matvec(o, x.copy()), y) - np.dot(x, _matvec(o, y.copy()))) \
        < 1e-9 * np.abs(_matvec(o, x.copy())).max() * math.sqrt(n)
    # Neumann in y, periodic in x and z: constants are in the null space
    assert np.abs(_matvec(o, np.ones(n))).max() < 1e-9


def test_bicgstab_true_residual():
    import ctypes as C
    case = configs.channel_case(n=(12, 10, 8))
    o, d = state(case)
    o.call("build_cages")
    n = d.xn * d.yn * d.zn
    rng = np.random.default_rng(9)
    b = rng.normal(size=n)
    b -= b.mean()
    b /= np.linalg.norm(b)
    x = np.zeros(n)
    resid = C.c_double()
    conv = C.c_int()
    it = o.lib.orc_bicgstab(o.ptr, b.ctypes.data_as(C.POINTER(C.c_double)),
                            x.ctypes.data_as(C.POINTER(C.c_double)), 2000,
                            1e-10, C.byref(resid), C.byref(conv))
    assert conv.value == 1 and it > 5
    true = np.linalg.norm(b - _matvec(o, x.copy()))
    assert true < 1e-8, true
    assert resid.value <= 1e-10


@pytest.mark.parametrize("which", ["channel", "particles"])
def test_bicgstab_iterates_match_textbook_implementation(which):
    """The iterate sequence, not just the converged answer: after each of
    the first eight iterations, orc_bicgstab's x equals that of SciPy's
    BiCGSTAB (van der Vorst's right-preconditioned method, r-hat = r0, the
    algorithm CUSP implements) run on the same matrix -- assembled here
    column by column from the oracle's operator -- with the same Jacobi
    preconditioner, b and x0.  Different summation orders and update
    spellings leave differences at rounding level."""
    import ctypes as C
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    case = (configs.channel_case(n=(8, 6, 6)) if which == "channel" else
            configs.small_particle_case(n=8, nparts=1, radius_dx=2.0))
    o, d = state(case)
    o.call("build_cages")
    n = d.xn * d.yn * d.zn
    cols = [_matvec(o, e.copy()) for e in np.eye(n)]
    A = sp.csr_matrix(np.array(cols).T)
    dinv = 1.0 / A.diagonal()
    M = spla.LinearOperator((n, n), matvec=lambda v: dinv * v)
    rng = np.random.default_rng(21)
    b = rng.normal(size=n)
    x0 = rng.normal(size=n) * 0.1
    ref = []
    spla.bicgstab(A, b, x0=x0.copy(), rtol=1e-300, atol=0.0, maxiter=8, M=M,
                  callback=lambda xk: ref.append(xk.copy()))
    assert len(ref) == 8
    P = C.POINTER(C.c_double)
    for k in range(1, 9):
        x = x0.copy()
        resid, conv = C.c_double(), C.c_int()
        it = o.lib.orc_bicgstab(o.ptr, b.ctypes.data_as(P), x.ctypes.data_as(P),
                                k, 1e-300, C.byref(resid), C.byref(conv))
        assert it == k
        err = np.abs(x - ref[k - 1]).max() / np.abs(ref[k - 1]).max()
        assert err < 1e-11, (k, err)
        # the recursively updated residual is the true one, to rounding
        assert resid.value == pytest.approx(np.linalg.norm(b - A @ x),
                                            rel=1e-8)


@pytest.mark.parametrize("which", ["tg", "channel", "particles"])
def test_projection_is_divergence_free(which):
    if which == "tg":
        case = configs.periodic_case(12, 0.01, pp_residual=1e-12,
                                     pp_max_iter=2000)
    elif which == "channel":
        case = configs.channel_case(n=(12, 10, 8))
        case.pp_residual = 1e-12
    else:
        case = configs.small_particle_case(n=16, nparts=1, radius_dx=3.0,
                                           pp_residual=1e-12)
    s = sim.Simulation(case)
    o = OracleState.from_sim(s)
    d = s.dom
    s.close()
    o.call("build_cages")
    o.call("dom_BC")
    o.s.dt = 2e-3
    o.s.dt0 = -1.0
    for c in "uvw":
        np.copyto(o.array(c + "0"), o.array(c))
    o.call("U_star_2")
    if o.nparts:
        o.call("part_BC_star")
    o.call("dom_BC_star")
    o.call("solvability")
    o.call("dom_BC_star")
    o.call("PP_bicgstab")
    assert o.s.last_converged
    o.call("dom_BC_phi")
    o.call("project")
    G = d.Gcc
    phi = interior(G, o.array("phi"))
    assert abs(phi.mean()) < 1e-12 * max(1.0, np.abs(phi).max())
    u = o.array("u").reshape(d.Gfx.knb, d.Gfx.jnb, d.Gfx.inb)
    v = o.array("v").reshape(d.Gfy.knb, d.Gfy.jnb, d.Gfy.inb)
    w = o.array("w").reshape(d.Gfz.knb, d.Gfz.jnb, d.Gfz.inb)
    div = (u[1:-1, 1:-1, 2:-1] - u[1:-1, 1:-1, 1:-2]) / d.dx \
        + (v[1:-1, 2:-1, 1:-1] - v[1:-1, 1:-2, 1:-1]) / d.dy \
        + (w[2:-1, 1:-1, 1:-1] - w[1:-2, 1:-1, 1:-1]) / d.dz
    ph = interior(G, o.array("phase"))
    shell = interior(G, o.array("phase_shell"))
    fluid = (ph < 0) & (shell > 0)
    us = o.array("u_star").reshape(d.Gfx.knb, d.Gfx.jnb, d.Gfx.inb)
    vs = o.array("v_star").reshape(d.Gfy.knb, d.Gfy.jnb, d.Gfy.inb)
    ws = o.array("w_star").reshape(d.Gfz.knb, d.Gfz.jnb, d.Gfz.inb)
    div_star = (us[1:-1, 1:-1, 2:-1] - us[1:-1, 1:-1, 1:-2]) / d.dx \
        + (vs[1:-1, 2:-1, 1:-1] - vs[1:-1, 1:-2, 1:-1]) / d.dy \
        + (ws[2:-1, 1:-1, 1:-1] - ws[1:-2, 1:-1, 1:-1]) / d.dz
    scale = np.abs(div_star[fluid]).max()
    assert scale > 0
    assert np.abs(div[fluid]).max() < 1e-8 * scale, (
        np.abs(div[fluid]).max(), scale)


def test_update_p_closed_form():
    case = configs.periodic_case(8, 0.01)
    o, d = state(case)
    o.call("build_cages")
    rng = np.random.default_rng(4)
    G = d.Gcc
    o.array("p0")[:] = rng.normal(size=G.s3b)
    o.array("phi")[:] = rng.normal(size=G.s3b)
    p0 = o.array("p0").copy()
    phi = o.array("phi").copy()
    o.call("update_p")
    want = interior(G, p0 + phi)
    want = want - want.mean()
    got = interior(G, o.array("p"))
    assert np.abs(got - want).max() < 1e-12


def test_find_dt_closed_form():
    case = configs.channel_case(n=(12, 10, 8))
    o, d = state(case)
    rng = np.random.default_rng(8)
    for c, G in (("u", d.Gfx), ("v", d.Gfy), ("w", d.Gfz)):
        o.array(c)[:] = rng.uniform(-1, 0.5, size=G.s3b)
    o.array("w")[:] = -np.abs(o.array("w"))   # all negative: max is 0
    um, vm = o.array("u").max(), o.array("v").max()
    nu = o.s.nu
    want = o.s.CFL / (um / d.dx + 2 * nu / d.dx ** 2 + vm / d.dy
                      + 2 * nu / d.dy ** 2 + 0.0 / d.dz + 2 * nu / d.dz ** 2)
    got = o.lib.orc_find_dt(o.ptr)
    assert abs(got - want) < 1e-14 * want


def test_solvability_zeroes_net_flux():
    case = configs.channel_case(n=(12, 10, 8))
    o, d = state(case)
    rng = np.random.default_rng(6)
    for c, G in (("u_star", d.Gfx), ("v_star", d.Gfy), ("w_star", d.Gfz)):
        o.array(c)[:] = rng.normal(size=G.s3b)
    o.call("solvability")
    us = o.array("u_star").reshape(d.Gfx.knb, d.Gfx.jnb, d.Gfx.inb)
    vs = o.array("v_star").reshape(d.Gfy.knb, d.Gfy.jnb, d.Gfy.inb)
    ws = o.array("w_star").reshape(d.Gfz.knb, d.Gfz.jnb, d.Gfz.inb)
    fx = (us[1:-1, 1:-1, d.Gfx.ie - 1] - us[1:-1, 1:-1, lim(d.Gfx)[0]]).sum() * d.dy * d.dz
    fy = (vs[1:-1, d.Gfy.je - 1, 1:-1] - vs[1:-1, d.Gfy.js, 1:-1]).sum() * d.dx * d.dz
    fz = (ws[d.Gfz.ke - 1, 1:-1, 1:-1] - ws[d.Gfz.ks, 1:-1, 1:-1]).sum() * d.dx * d.dy
    assert abs(fx + fy + fz) < 1e-11


def test_forcing_ramps():
    """f = -gradP / rho_f with gradP ramping at rate `a` from the delay
    until it reaches `m` (a = 0: m at once); before the delay gradP keeps
    its value."""
    case = configs.periodic_case(8, 0.01)
    o, d = state(case)
    o.s.rho_f = 2.0
    o.s.p_tDelay = 0.1
    gp = o.s.gradP
    gp.x, gp.y, gp.z = 9.0, 9.0, 9.0
    gp.xm, gp.xa = 2.0, 0.5
    gp.ym, gp.ya = -3.0, 0.0
    gp.zm, gp.za = -1.0, -4.0
    for t, want in ((0.05, (9.0, 9.0, 9.0)), (1.1, (0.5, -3.0, -1.0)),
                    (0.2, (0.05, -3.0, -0.4)), (20.0, (2.0, -3.0, -1.0))):
        gp.x, gp.y, gp.z = 9.0, 9.0, 9.0
        o.s.ttime = t
        o.call("compute_forcing")
        got = (o.s.gradP.x, o.s.gradP.y, o.s.gradP.z)
        assert np.allclose(got, want, rtol=1e-14, atol=1e-15), (t, got)
        for c, w in (("f_x", want[0]), ("f_y", want[1]), ("f_z", want[2])):
            assert np.allclose(o.array(c), -w / 2.0, rtol=1e-14, atol=0)


def test_history_shift_copies():
    case = configs.small_particle_case(n=8, nparts=1, radius_dx=2.0)
    o, d = state(case)
    rng = np.random.default_rng(7)
    for name in ("conv_u", "conv_v", "conv_w", "diff_u", "diff_v", "diff_w",
                 "u", "v", "w", "p"):
        o.array(name)[:] = rng.normal(size=o.array(name).size)
    o.call("store_u")
    for a, b in (("conv_u", "conv0_u"), ("conv_v", "conv0_v"),
                 ("conv_w", "conv0_w"), ("diff_u", "diff0_u"),
                 ("diff_v", "diff0_v"), ("diff_w", "diff0_w"), ("u", "u0"),
                 ("v", "v0"), ("w", "w0"), ("p", "p0")):
        assert np.array_equal(o.array(a), o.array(b)), b
    from oracle.pyoracle import COEFF_NAMES
    for nm in COEFF_NAMES[:6]:
        o.array(nm)[:] = rng.normal(size=o.array(nm).size)
    o.call("store_coeffs")
    for nm, nm00 in zip(COEFF_NAMES[:6], COEFF_NAMES[12:18]):
        assert np.array_equal(o.array(nm), o.array(nm00))


def test_pid_forcing_one_step():
    """The PID controller on gradP.z (src/cuda_bluebottle.cu:3784-3806),
    one step computed by hand: two spheres of radius 0.1 and density 3 in a
    unit box with rho_f = 1 make a suspension of volume fraction
    phi = 2 (4/3) pi 0.1^3 and mixture density phi*3 + (1 - phi)*1; their
    mean vertical acceleration a = (0.2 + 0.6)/2 = 0.4 feeds
    gradP.z += (Kp a + Ki I/t + Kd (a - a_prev)) rho_mix with the running
    integral I += a dt; the ramp of gradP.z is off while the controller
    runs."""
    case = configs.small_particle_case(n=8, nparts=2, radius_dx=0.8)
    o, d = state(case)
    for P, wdot in zip(o.s.parts[:2], (0.2, 0.6)):
        P.r, P.rho, P.wdot = 0.1, 3.0, wdot
    o.s.rho_f = 1.0
    o.s.Kp, o.s.Ki, o.s.Kd = 2.0, 0.5, 0.25
    o.s.pid_int, o.s.pid_back = 0.01, 0.3
    o.s.dt, o.s.ttime, o.s.p_tDelay = 0.02, 0.8, 0.0
    gp = o.s.gradP
    gp.z, gp.zm, gp.za = -0.7, 5.0, 0.0
    o.call("compute_forcing")
    phi = 2 * (4.0 / 3.0) * math.pi * 0.1 ** 3
    rho_mix = phi * 3.0 + (1 - phi) * 1.0
    integ = 0.01 + 0.4 * 0.02                       # 0.018
    ctrl = 2.0 * 0.4 + 0.5 * integ / 0.8 + 0.25 * (0.4 - 0.3)
    want = -0.7 + ctrl * rho_mix
    assert abs(o.s.gradP.z - want) < 1e-14
    assert abs(o.s.pid_int - 0.018) < 1e-16
    assert o.s.pid_back == 0.4
    assert np.allclose(o.array("f_z"), -want, rtol=1e-14, atol=0)
