"""Pins of the oracle's Crank-Nicolson velocity solves (the reference's
IMPLICIT build, src/cuda_bicgstab.cu:44-714) against closed forms.

  * the 13-diagonal operator I - (nu dt / 2) L, applied as a DIA product
    with the periodic wrap diagonals, has the eigenvalue
    1 + (nu dt / 2)(4/h^2) sum sin^2(pi k h) on periodic Fourier modes of
    every component (the duplicated periodic face included), and on the
    wall-normal cosine modes of the tangential components of a channel,
    where the zero wall-face flags make the implicit part a homogeneous
    Neumann Laplacian (eigenvalue (4/h^2) sin^2(pi m / 2n));
  * faces inside particles and wall-normal Dirichlet faces are identity rows
    and keep the value the boundary conditions gave them;
  * the conjugate-gradient solution agrees with a dense direct solve of
    the same matrix, and an iteration-capped solve reports the capped count,
    not converged, and its true residual;
  * the whole update multiplies a shear mode by the Crank-Nicolson factor
    (1 + nu dt lam / 2) / (1 - nu dt lam / 2);
  * with nu = 0 the operator is the identity and u* is the explicit
    Euler / Adams-Bashforth update of the convection, pressure gradient and
    forcing, exact on linear fields, with conv stored;
  * find_dt drops the viscous limit and caps growth at 1.1 dt; store_u
    leaves no diffusion history.
"""
import ctypes as C
import math

import numpy as np

from bluebottle_amd import configs
from test_oracle_field import _lin, _grad, _set_fields, coords, interior, \
    state

_PR = C.POINTER(C.c_double)


def _p(a):
    return a.ctypes.data_as(_PR)


def _grid(d, comp):
    return (d.Gfx, d.Gfy, d.Gfz)[comp]


def _shape(d, comp):
    G = _grid(d, comp)
    return getattr(G, "in"), G.jn, G.kn


def _coeffs(o, comp):
    n = _grid(o.s.dom, comp).s3
    vals = np.zeros(13 * n)
    o.lib.orc_helmholtz_coeffs(o.ptr, comp, _p(vals))
    return vals


def _matvec(o, comp, vals, x):
    y = np.zeros_like(x)
    o.lib.orc_helmholtz_matvec(o.ptr, comp, _p(vals), _p(x), _p(y))
    return y


def _mode(d, comp, k, neumann_y=False):
    """A separable Fourier mode on the unknowns of `comp` (x fastest) and
    the eigenvalue of the discrete Laplacian on it."""
    ni, nj, nk = _shape(d, comp)
    h = (d.dx, d.dy, d.dz)
    n = (ni, nj, nk)
    fac, lam = [], 0.0
    for a in range(3):
        idx = np.arange(n[a])
        if a == 1 and neumann_y:
            fac.append(np.cos(math.pi * k[a] * (idx + 0.5) / n[a]))
            lam -= 4 / h[a] ** 2 * math.sin(math.pi * k[a] / (2 * n[a])) ** 2
        else:
            pos = idx * h[a] + (0.0 if a == comp else 0.5 * h[a])
            fac.append(np.cos(2 * math.pi * k[a] * pos + 0.3 * (a + 1)))
            lam -= 4 / h[a] ** 2 * math.sin(math.pi * k[a] * h[a]) ** 2
    x = fac[2][:, None, None] * fac[1][None, :, None] * fac[0][None, None, :]
    return x.ravel().copy(), lam


def test_operator_periodic_eigenmodes():
    o, d = state(configs.periodic_case(10, 0.05))
    o.s.dt = 0.02
    for comp in range(3):
        vals = _coeffs(o, comp)
        for k in ((1, 2, 0), (3, 1, 2), (0, 0, 4)):
            x, lam = _mode(d, comp, k)
            y = _matvec(o, comp, vals, x)
            want = (1 - 0.5 * o.s.nu * o.s.dt * lam) * x
            assert np.abs(y - want).max() < 1e-13, (comp, k)


def test_operator_channel_tangential_modes_are_neumann():
    o, d = state(configs.channel_case(n=(8, 6, 10)))
    o.s.nu = 0.04
    o.s.dt = 0.03
    for comp in (0, 2):
        vals = _coeffs(o, comp)
        for k in ((1, 1, 2), (2, 3, 1), (0, 5, 3)):
            x, lam = _mode(d, comp, k, neumann_y=True)
            y = _matvec(o, comp, vals, x)
            want = (1 - 0.5 * o.s.nu * o.s.dt * lam) * x
            assert np.abs(y - want).max() < 1e-13, (comp, k)


def test_identity_rows_on_walls_and_particles():
    o, d = state(configs.channel_case(n=(8, 6, 8)))
    o.s.dt = 0.01
    vals = _coeffs(o, 1).reshape(13, -1)
    ni, nj, nk = _shape(d, 1)
    row = vals.reshape(13, nk, nj, ni)
    e6 = np.zeros(13)
    e6[6] = 1.0
    # the z-periodic kernels run after the y-wall ones and touch the wall
    # rows of the first and last z planes again (src/cuda_bicgstab.cu:
    # 400-456 launch order), so only the rows between them are identities
    for j in (0, nj - 1):
        blk = row[:, 1:-1, j, :].reshape(13, -1)
        assert np.array_equal(blk, np.repeat(e6[:, None], blk.shape[1], 1))
    o, d = state(configs.small_particle_case(n=12, nparts=1, radius_dx=3.0))
    o.s.dt = 0.01
    nid = 0
    for comp, name in ((0, "flag_u"), (1, "flag_v"), (2, "flag_w")):
        G = _grid(d, comp)
        own = interior(G, o.array(name)).ravel()
        vals = _coeffs(o, comp).reshape(13, -1)
        solid = own < 0
        nid += int(solid.sum())
        assert np.array_equal(vals[:, solid],
                              np.repeat(e6[:, None], solid.sum(), 1))
        assert np.all(vals[6, ~solid] > 1.0)
    assert nid > 0


def _dense(o, comp, vals):
    """Dense matrix of the DIA arrays: row C holds vals[dg, C] at column
    C + offset[dg] when that column exists."""
    d = o.s.dom
    G = _grid(d, comp)
    n, s1, s2, s3 = G.s3, G.s1, G.s2, G.s3
    w = [2 if comp == a else 1 for a in range(3)]
    off = [-s3 + w[2] * s2, -s2, -s2 + w[1] * s1, -s1, -s1 + w[0], -1, 0, 1,
           s1 - w[0], s1, s2 - w[1] * s1, s2, s3 - w[2] * s2]
    v = vals.reshape(13, n)
    A = np.zeros((n, n))
    for dg in range(13):
        for r in range(n):
            c = r + off[dg]
            if 0 <= c < n:
                A[r, c] += v[dg, r]
    return A


def _pcg(o, comp, vals, b, max_iter, tol):
    x = np.zeros_like(b)
    resid = np.zeros(1)
    conv = np.zeros(1, dtype=np.int32)
    it = o.lib.orc_pcg(o.ptr, comp, _p(vals), _p(b), _p(x), max_iter, tol,
                       _p(resid), conv.ctypes.data_as(C.POINTER(C.c_int)))
    return x, it, float(resid[0]), int(conv[0])


def test_pcg_matches_dense_solve():
    o, d = state(configs.channel_case(n=(6, 4, 6)))
    o.s.nu = 0.05
    o.s.dt = 0.05
    rng = np.random.default_rng(11)
    for comp in (0, 2):
        vals = _coeffs(o, comp)
        A = _dense(o, comp, vals)
        b = rng.normal(size=A.shape[0])
        b /= np.linalg.norm(b)
        x, it, resid, conv = _pcg(o, comp, vals, b, 500, 1e-14)
        assert conv == 1 and 0 < it < 500
        xd = np.linalg.solve(A, b)
        assert np.abs(x - xd).max() < 1e-12 * np.abs(xd).max()
        assert np.linalg.norm(b - A @ x) <= 1.01e-14


def test_pcg_iteration_cap_reports_true_residual():
    o, d = state(configs.periodic_case(8, 0.3))
    o.s.dt = 0.05
    rng = np.random.default_rng(3)
    vals = _coeffs(o, 0)
    A = _dense(o, 0, vals)
    b = rng.normal(size=A.shape[0])
    b /= np.linalg.norm(b)
    x, it, resid, conv = _pcg(o, 0, vals, b, 3, 1e-14)
    assert it == 3 and conv == 0
    true = np.linalg.norm(b - A @ x)
    assert abs(resid - true) < 1e-12 and resid > 1e-6
    x0, it0, _, conv0 = _pcg(o, 0, vals, b, 0, 1e-14)
    assert it0 == 0 and conv0 == 0 and not x0.any()


def test_crank_nicolson_factor_on_shear_mode():
    o, d = state(configs.periodic_case(16, 0.05))
    o.s.implicit = 1
    o.s.dt = 0.01
    o.s.dt0 = -1.0
    o.s.pp_residual = 1e-14
    o.s.pp_max_iter = 1000
    m = 2
    _set_fields(o, d, lambda X, Y, Z: np.sin(2 * math.pi * m * Y),
                lambda X, Y, Z: 0 * X, lambda X, Y, Z: 0 * X,
                lambda X, Y, Z: 0 * X)
    for c in ("f_x", "f_y", "f_z"):
        o.array(c)[:] = 0.0
    lam = -4 / d.dy ** 2 * math.sin(math.pi * m * d.dy) ** 2
    q = 0.5 * o.s.nu * o.s.dt * lam
    g = (1 + q) / (1 - q)
    for comp in range(3):
        o.lib.orc_star_helmholtz(o.ptr, comp)
        assert o.s.last_converged == 1
    want = g * interior(d.Gfx, o.array("u0"))
    got = interior(d.Gfx, o.array("u_star"))
    assert np.abs(got - want).max() < 1e-12
    for c, G in (("v_star", d.Gfy), ("w_star", d.Gfz)):
        assert np.abs(interior(G, o.array(c))).max() < 1e-14
    for c in ("conv_u", "conv_v", "conv_w"):
        assert np.abs(o.array(c)).max() < 1e-12


def _own_inner(G, a, comp):
    """Interior faces less the two boundary faces along the component's own
    axis (the periodic boundary condition rewrites the first of them)."""
    x = interior(G, a)
    sl = [slice(None)] * 3
    sl[2 - comp] = slice(1, -1)
    return x[tuple(sl)]


def test_inviscid_update_is_explicit_on_linear_fields():
    case = configs.periodic_case(8, 0.0)
    for ab2 in (False, True):
        o, d = state(case)
        o.s.implicit = 1
        o.s.nu = 0.0
        o.s.dt = 2e-3
        o.s.dt0 = 1.5e-3 if ab2 else -1.0
        pg = (0.11, -0.07, 0.05)
        _set_fields(o, d, _lin("u"), _lin("v"), _lin("w"),
                    lambda X, Y, Z: 1.0 + pg[0] * X + pg[1] * Y + pg[2] * Z)
        rng = np.random.default_rng(4)
        for comp, G in (("u", d.Gfx), ("v", d.Gfy), ("w", d.Gfz)):
            o.array("conv0_" + comp)[:] = rng.normal(size=G.s3b)
        for c in ("f_x", "f_y", "f_z"):
            o.array(c)[:] = 0.01
        for comp in range(3):
            o.lib.orc_star_helmholtz(o.ptr, comp)
            assert o.s.last_converged == 1 and o.s.last_iters <= 1
        r = o.s.dt / (2 * o.s.dt0) if ab2 else 0.0
        for ci, (comp, G) in enumerate((("u", d.Gfx), ("v", d.Gfy),
                                        ("w", d.Gfz))):
            X, Y, Z = coords(d, G, comp)
            U = {c: _lin(c)(X, Y, Z) for c in "uvw"}
            gr = {c: _grad(c) for c in "uvw"}
            div = gr["u"][0] + gr["v"][1] + gr["w"][2]
            conv = sum(U[b] * gr[comp][ib] for ib, b in enumerate("uvw")) \
                + U[comp] * div
            got_c = interior(G, o.array("conv_" + comp))
            assert np.abs(got_c - interior(G, conv.ravel())).max() < 1e-12
            c0 = o.array("conv0_" + comp).reshape(conv.shape)
            cterm = -(1 + r) * conv + r * c0 if ab2 else -conv
            want = U[comp] + o.s.dt * (-pg[ci] / o.s.rho_f + cterm + 0.01)
            got = _own_inner(G, o.array(comp + "_star"), ci)
            assert np.abs(got - _own_inner(G, want.ravel(), ci)).max() < 1e-13


def test_wall_normal_faces_keep_boundary_values():
    o, d = state(configs.channel_case(n=(8, 6, 8)))
    o.s.implicit = 1
    o.s.dt = 0.01
    o.s.dt0 = -1.0
    o.s.pp_residual = 1e-13
    rng = np.random.default_rng(2)
    for c in ("u0", "v0", "w0"):
        o.array(c)[:] = rng.normal(size=o.array(c).size)
    o.call("dom_BC")
    for comp in range(3):
        o.lib.orc_star_helmholtz(o.ptr, comp)
        assert o.s.last_converged == 1
    G = d.Gfy
    vs = o.array("v_star").reshape(G.knb, G.jnb, G.inb)
    assert np.abs(vs[1:-1, G.js, 1:-1]).max() < 1e-14
    assert np.abs(vs[1:-1, G.je - 1, 1:-1]).max() < 1e-14


def test_find_dt_implicit_closed_form_and_cap():
    o, d = state(configs.channel_case(n=(12, 10, 8)))
    o.s.implicit = 1
    rng = np.random.default_rng(9)
    for c, G in (("u", d.Gfx), ("v", d.Gfy), ("w", d.Gfz)):
        o.array(c)[:] = rng.uniform(-1, 0.5, size=G.s3b)
    um, vm, wm = (max(o.array(c).max(), 0.0) for c in "uvw")
    free = o.s.CFL / (um / d.dx + vm / d.dy + wm / d.dz)
    o.s.dt = 10 * free
    assert abs(o.lib.orc_find_dt(o.ptr) - free) < 1e-14 * free
    o.s.dt = 0.5 * free
    assert o.lib.orc_find_dt(o.ptr) == 1.1 * o.s.dt


def test_store_u_implicit_keeps_no_diffusion_history():
    o, d = state(configs.periodic_case(8, 0.01))
    o.s.implicit = 1
    rng = np.random.default_rng(1)
    for name in ("conv_u", "diff_u", "u", "p"):
        o.array(name)[:] = rng.normal(size=o.array(name).size)
    o.array("diff0_u")[:] = 7.0
    o.call("store_u")
    assert np.array_equal(o.array("conv0_u"), o.array("conv_u"))
    assert np.array_equal(o.array("u0"), o.array("u"))
    assert np.all(o.array("diff0_u") == 7.0)


def test_pcg_iterates_match_textbook_implementation():
    """The iterate sequence: after each of the first six iterations,
    orc_pcg's x equals that of SciPy's preconditioned conjugate gradient
    (the method CUSP's cg implements) on the same assembled matrix, with
    the same Jacobi preconditioner and b, from x0 = 0, for a tangential
    and a wall-normal component of a channel."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    o, d = state(configs.channel_case(n=(6, 4, 6), nparts=0))
    o.s.nu = 0.05
    o.s.dt = 0.05
    rng = np.random.default_rng(12)
    for comp in (0, 1):
        vals = _coeffs(o, comp)
        A = sp.csr_matrix(_dense(o, comp, vals))
        n = A.shape[0]
        dinv = 1.0 / A.diagonal()
        M = spla.LinearOperator((n, n), matvec=lambda v: dinv * v)
        b = rng.normal(size=n)
        ref = []
        spla.cg(A, b, x0=np.zeros(n), rtol=1e-300, atol=0.0, maxiter=6, M=M,
                callback=lambda xk: ref.append(xk.copy()))
        assert len(ref) == 6
        for k in range(1, 7):
            x, it, resid, conv = _pcg(o, comp, vals, b, k, 1e-300)
            assert it == k
            err = np.abs(x - ref[k - 1]).max() / np.abs(ref[k - 1]).max()
            assert err < 1e-12, (comp, k, err)
