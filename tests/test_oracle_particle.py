"""Pins of the oracle's particle kernels against their definitions and
closed forms (brute force on small grids):
  * cages and phases (src/cuda_particle.cu:357-896,
    src/particle_kernel.cu:25-533): a cell belongs to particle p when its
    centre lies within r of p or of a periodic image (the highest index
    wins where spheres overlap); phase_shell is 0 exactly on solid cells
    with a fluid face neighbour; a face flag is -1 on a fluid/solid
    interface or between two shell cells, 0 on non-periodic domain faces
    and 1 elsewhere;
  * particle motion (src/bluebottle_kernel.cu:2247-3617): free fall under
    buoyancy-corrected gravity and the mean pressure gradient gives a
    linear velocity ramp (after the first, half-weighted, trapezoid step)
    and the matching positions; a constant angular velocity turns the body
    axes by |omega| dt per step about omega;
  * lamb_err (src/quadrature_kernel.cu:578-654): the largest relative
    change among coefficients above the lamb_cut threshold, per particle,
    maximised over particles; coeff0 takes the new coefficients.
"""
import ctypes as C
import math

import numpy as np
import pytest

from bluebottle_amd import abi, configs, sim
from oracle.pyoracle import OracleState, COEFF_NAMES

FLT_MIN = 1.1754943508222875e-38


def _state(case):
    s = sim.Simulation(case)
    o = OracleState.from_sim(s)
    d = s.dom
    s.close()
    return o, d


def brute_phase(d, parts, per):
    xn, yn, zn = d.xn, d.yn, d.zn
    xc = d.xs + (np.arange(xn) + 0.5) * d.dx
    yc = d.ys + (np.arange(yn) + 0.5) * d.dy
    zc = d.zs + (np.arange(zn) + 0.5) * d.dz
    Z, Y, X = np.meshgrid(zc, yc, xc, indexing="ij")
    phase = -np.ones((zn, yn, xn), dtype=np.int32)
    L = (d.xl, d.yl, d.zl)
    for p, P in enumerate(parts):
        dx = X - P.x
        dy = Y - P.y
        dz = Z - P.z
        for arr, ax in ((dx, 0), (dy, 1), (dz, 2)):
            if per[ax]:
                arr -= L[ax] * np.round(arr / L[ax])
        inside = dx * dx + dy * dy + dz * dz < P.r * P.r
        phase[inside] = p
    return phase


def brute_shell_flags(phase, per):
    """Shell and face flags from the interior phase array.  Neighbours
    across a periodic axis wrap; across a wall they are fluid."""
    solid = phase > -1

    def nb(a, ax, shift, fill):
        if per[2 - ax]:
            return np.roll(a, shift, axis=ax)
        out = np.full_like(a, fill)
        sl_dst = [slice(None)] * 3
        sl_src = [slice(None)] * 3
        if shift == 1:
            sl_dst[ax] = slice(1, None)
            sl_src[ax] = slice(None, -1)
        else:
            sl_dst[ax] = slice(None, -1)
            sl_src[ax] = slice(1, None)
        out[tuple(sl_dst)] = a[tuple(sl_src)]
        return out

    shell = np.ones_like(phase)
    for ax in range(3):
        for sh in (1, -1):
            fluid_nb = ~nb(solid, ax, sh, False)
            shell[solid & fluid_nb] = 0
    flags = []
    # faces normal to x (axis 2 of the (k, j, i) arrays), then y, then z
    for ax in (2, 1, 0):
        lo_solid = nb(solid, ax, 1, False)      # cell on the low side
        lo_shell = nb(shell, ax, 1, 1)
        f = 1 - 2 * ((lo_solid != solid) | ((lo_shell < 1) & (shell < 1)))
        flags.append(f.astype(np.int32))        # face at the low side of each cell
    return shell, flags


@pytest.mark.parametrize("which", ["periodic", "channel", "odd"])
def test_cages_phase_shell_flags_brute_force(which):
    if which == "periodic":
        case = configs.small_particle_case(n=20, nparts=6, radius_dx=2.7,
                                           seed=11)
        # push two spheres across the periodic faces
        case.particles[0].x = 0.02
        case.particles[1].y = 0.985
        case.particles[1].z = 0.01
    elif which == "odd":
        case = configs.small_particle_case(n=17, nparts=4, radius_dx=3.1,
                                           seed=4)
        case.particles[2].x = 0.99
    else:
        case = configs.channel_case(n=(20, 16, 18), nparts=5)
    o, d = _state(case)
    o.call("build_cages")
    b = case.bc_u
    per = (b[0][0] == abi.PERIODIC, b[2][0] == abi.PERIODIC,
           b[4][0] == abi.PERIODIC)
    parts = o.parts()
    phase = brute_phase(d, parts, per)
    Gc = d.Gcc
    got = o.array("phase").reshape(Gc.knb, Gc.jnb, Gc.inb)[1:-1, 1:-1, 1:-1]
    assert np.array_equal(got, phase), np.argwhere(got != phase)[:5]
    assert (phase > -1).sum() > 50
    shell, flags = brute_shell_flags(phase, per)
    gots = o.array("phase_shell").reshape(Gc.knb, Gc.jnb, Gc.inb)[1:-1, 1:-1, 1:-1]
    assert np.array_equal(gots, shell)
    for name, G, f, ax in (("flag_u", d.Gfx, flags[0], 0),
                           ("flag_v", d.Gfy, flags[1], 1),
                           ("flag_w", d.Gfz, flags[2], 2)):
        a = o.array(name).reshape(G.knb, G.jnb, G.inb)
        # faces at the low side of every interior cell
        if ax == 0:
            low = a[1:-1, 1:-1, 1:-2]
            high = a[1:-1, 1:-1, -2]
            fhigh = f[:, :, 0] if per[0] else np.zeros_like(f[:, :, 0])
        elif ax == 1:
            low = a[1:-1, 1:-2, 1:-1]
            high = a[1:-1, -2, 1:-1]
            fhigh = f[:, 0, :] if per[1] else np.zeros_like(f[:, 0, :])
        else:
            low = a[1:-2, 1:-1, 1:-1]
            high = a[-2, 1:-1, 1:-1]
            fhigh = f[0, :, :] if per[2] else np.zeros_like(f[0, :, :])
        want = f.copy()
        if not per[ax]:
            # the domain's own faces are 0 on a wall
            if ax == 0:
                want[:, :, 0] = 0
            elif ax == 1:
                want[:, 0, :] = 0
            else:
                want[0, :, :] = 0
        assert np.array_equal(low, want), (name, np.argwhere(low != want)[:5])
        assert np.array_equal(high, fhigh), name


def _one_sphere(rho=2.0, **kw):
    case = configs.small_particle_case(n=16, nparts=1, radius_dx=3.0,
                                       taylor_green=False)
    P = case.particles[0]
    P.x, P.y, P.z, P.rho = 0.5, 0.5, 0.5, rho
    for k, v in kw.items():
        setattr(P, k, v)
    return case


def test_free_fall_linear_ramp():
    case = _one_sphere(rho=2.5)
    case.g = (0.0, 0.0, 0.0, 0.0, -1.0, 0.0)
    case.gradP = (0.0, 0.0, 0.0, 0.0, 0.3, 0.0)
    o, d = _state(case)
    o.s.g.z = -1.0
    o.s.gradP.z = 0.3
    o.s.dt = 1e-3
    rho_p, rho_f = 2.5, o.s.rho_f
    a = (1 - rho_f / rho_p) * (-1.0) - 0.3 / rho_p
    ws, zs = [0.0], [0.5]
    for _ in range(5):
        o.call("move_parts")
        P = o.parts()[0]
        ws.append(P.w)
        zs.append(P.z)
    dw = np.diff(ws)
    assert abs(dw[0] - 0.5 * o.s.dt * a) < 1e-15
    assert np.allclose(dw[1:], o.s.dt * a, rtol=1e-12, atol=0)
    for n in range(1, 6):
        assert abs((zs[n] - zs[n - 1]) - 0.5 * o.s.dt * (ws[n] + ws[n - 1])) < 1e-15
    P = o.parts()[0]
    assert P.u == 0.0 and P.v == 0.0


def test_constant_spin_rotates_body_axes():
    case = _one_sphere()
    o, d = _state(case)
    o.s.dt = 2e-3
    om = np.array([0.3, -1.1, 2.0])
    P = o.s.parts[0]
    P.ox, P.oy, P.oz = om
    P.ox0, P.oy0, P.oz0 = om
    steps = 7
    for _ in range(steps):
        o.call("move_parts")
    P = o.parts()[0]
    theta = np.linalg.norm(om) * o.s.dt * steps
    k = om / np.linalg.norm(om)

    def rod(v):
        return v * math.cos(theta) + np.cross(k, v) * math.sin(theta) \
            + k * np.dot(k, v) * (1 - math.cos(theta))

    for got, e in (((P.axx, P.axy, P.axz), (1, 0, 0)),
                   ((P.ayx, P.ayy, P.ayz), (0, 1, 0)),
                   ((P.azx, P.azy, P.azz), (0, 0, 1))):
        assert np.allclose(got, rod(np.array(e, float)), atol=1e-12)
    assert (P.ox, P.oy, P.oz) == tuple(om)


def lamb_err_brute(o):
    stride, n = o.coeff_stride, o.nparts
    best = FLT_MIN
    for p in range(n):
        c = np.concatenate([o.array(nm)[p * stride:(p + 1) * stride]
                            for nm in COEFF_NAMES[:6]])
        c0 = np.concatenate([o.array(nm)[p * stride:(p + 1) * stride]
                             for nm in COEFF_NAMES[6:12]])
        big = c * c > FLT_MIN
        if not big.any():
            continue
        cmax = np.abs(c[big]).max()
        sel = big & (np.abs(c) > o.s.lamb_cut * cmax)
        err = np.abs(c[sel] - c0[sel]) / np.maximum(np.abs(c[sel]), 1e-16)
        best = max(best, err.max())
    return best


def test_lamb_err_brute_force():
    case = configs.small_particle_case(n=16, nparts=4, radius_dx=2.5)
    o, d = _state(case)
    rng = np.random.default_rng(12)
    for nm in COEFF_NAMES[:12]:
        a = o.array(nm)
        a[:] = rng.normal(size=a.size) * 10.0 ** rng.integers(-8, 1, size=a.size)
    # particle 2: all coefficients tiny except one; particle 3: unchanged
    st = o.coeff_stride
    for nm in COEFF_NAMES[:6]:
        o.array(nm)[2 * st:3 * st] = 1e-21
    o.array("pnm_re")[2 * st] = 0.5
    for nm, nm0 in zip(COEFF_NAMES[:6], COEFF_NAMES[6:12]):
        o.array(nm0)[3 * st:4 * st] = o.array(nm)[3 * st:4 * st]
    want = lamb_err_brute(o)
    new = {nm: o.array(nm).copy() for nm in COEFF_NAMES[:6]}
    got = o.lib.orc_lamb_err(o.ptr)
    assert got == want
    for nm, nm0 in zip(COEFF_NAMES[:6], COEFF_NAMES[6:12]):
        assert np.array_equal(o.array(nm0), new[nm])


# ---------------------------------------------------------------------------
# pair forces: lubrication and contact
# ---------------------------------------------------------------------------
def _pair(gap, U, a=0.1, E=1.0e6, sigma=0.5):
    """Two equal spheres on the x axis, gap `gap` between surfaces; the
    first moves with +U towards the second (neither is integrated, so both
    collision passes see the same velocities)."""
    case = configs.small_particle_case(n=16, nparts=2, radius_dx=1.0,
                                       taylor_green=False)
    x0 = 0.3
    for P, x in zip(case.particles, (x0, x0 + 2 * a + gap)):
        P.x, P.y, P.z, P.r = x, 0.5, 0.5, a
        P.translating, P.rotating = 0, 0
        P.E, P.sigma = E, sigma
    o, d = _state(case)
    o.s.dt = 1e-4
    o.s.parts[0].u = U
    o.s.parts[0].u0 = U
    o.call("move_parts_sub")
    return o, o.parts()


def test_lubrication_reynolds_limit_and_third_law():
    a, U = 0.1, 0.05
    o, (Pi, Pj) = _pair(0.02 * a, U, a)
    mu = o.s.mu
    reynolds = -1.5 * math.pi * mu * a * a * U / (0.02 * a)
    # the squeeze force opposes the approach and tends to Reynolds' law
    assert Pi.iFx < 0 < Pj.iFx
    assert abs(Pi.iFx / reynolds - 1) < 0.1
    assert abs(Pi.iFx + Pj.iFx) < 1e-12 * abs(Pi.iFx)
    for f in ("iFy", "iFz", "iLx", "iLy", "iLz"):
        assert abs(getattr(Pi, f)) < 1e-12 * abs(Pi.iFx)
    # closer gaps push harder, and the force is linear in the velocity
    o2, (Qi, _) = _pair(0.01 * a * 2, 2 * U, a)
    assert abs(Qi.iFx / Pi.iFx - 2.0) < 1e-12


def test_no_pair_force_beyond_interaction_length():
    a = 0.1
    o, (Pi, Pj) = _pair(1.05 * a, 0.05, a)   # hN = r (ratio 1)
    for P in (Pi, Pj):
        for f in ("iFx", "iFy", "iFz", "iLx", "iLy", "iLz"):
            assert getattr(P, f) == 0.0


def test_hertz_contact_at_rest():
    a, E, sigma = 0.1, 2.0e5, 0.3
    delta = 0.01 * a
    o, (Pi, Pj) = _pair(-delta, 0.0, a, E=E, sigma=sigma)
    estar = 1.0 / (2 * (1 - sigma ** 2) / E)
    rstar = 1.0 / (2.0 / a)
    hertz = 4.0 / 3.0 * estar * math.sqrt(rstar) * delta ** 1.5
    # repulsive along the line of centres, equal and opposite
    assert abs(-Pi.iFx / hertz - 1) < 1e-12
    assert abs(Pj.iFx / hertz - 1) < 1e-12
    assert abs(Pi.iFy) < 1e-12 * hertz and abs(Pi.iFz) < 1e-12 * hertz


# ---------------------------------------------------------------------------
# rigid-body fills inside particles
# ---------------------------------------------------------------------------
def _rigid_case():
    case = _one_sphere()
    P = case.particles[0]
    P.x, P.y, P.z = 0.47, 0.52, 0.5
    return case


def _min_image(v, L):
    return v - L * np.round(v / L)


def test_parts_internal_is_rigid_body_motion():
    o, d = _state(_rigid_case())
    o.call("build_cages")
    P = o.s.parts[0]
    U = np.array([0.1, -0.2, 0.05])
    Om = np.array([0.7, 0.3, -1.2])
    P.u, P.v, P.w = U
    P.ox, P.oy, P.oz = Om
    rng = np.random.default_rng(1)
    before = {}
    for c, G in (("u", d.Gfx), ("v", d.Gfy), ("w", d.Gfz)):
        o.array(c)[:] = rng.normal(size=G.s3b)
        before[c] = o.array(c).copy()
    o.call("parts_internal")
    Gc = d.Gcc
    ph = o.array("phase").reshape(Gc.knb, Gc.jnb, Gc.inb)
    changed = 0
    for ax, (c, G, off) in enumerate((("u", d.Gfx, (1.0, 0.5, 0.5)),
                                      ("v", d.Gfy, (0.5, 1.0, 0.5)),
                                      ("w", d.Gfz, (0.5, 0.5, 1.0)))):
        fl = o.array("flag_" + c).reshape(G.knb, G.jnb, G.inb)
        got = o.array(c).reshape(G.knb, G.jnb, G.inb)
        old = before[c].reshape(G.knb, G.jnb, G.inb)
        for k in range(1, G.knb - 1):
            for j in range(1, G.jnb - 1):
                for i in range(1, G.inb - 1):
                    cell = [k, j, i]
                    lo = list(cell)
                    lo[2 - ax] -= 1
                    if not (0 <= lo[2 - ax]) or cell[2 - ax] > (Gc.knb, Gc.jnb, Gc.inb)[2 - ax] - 2:
                        continue
                    inside = ph[tuple(lo)] > -1 and ph[tuple(cell)] > -1 \
                        and fl[k, j, i] != -1
                    if not inside:
                        assert got[k, j, i] == old[k, j, i]
                        continue
                    pos = np.array([d.xs + (i - off[0]) * d.dx,
                                    d.ys + (j - off[1]) * d.dy,
                                    d.zs + (k - off[2]) * d.dz])
                    r = _min_image(pos - np.array([P.x, P.y, P.z]), d.xl)
                    want = (U + np.cross(Om, r))[ax]
                    assert abs(got[k, j, i] - want) < 1e-14
                    changed += 1
    assert changed > 20


def test_p_fill_is_rigid_body_pressure():
    """Inside a particle the pressure is the n = 0 Lamb term plus the
    pressure of a rigid body's frame: centrifugal 1/2 rho |Omega x r|^2 and
    the hydrostatic (-gradP - rho a_P) . r."""
    o, d = _state(_rigid_case())
    o.call("build_cages")
    P = o.s.parts[0]
    Om = np.array([0.4, -0.8, 1.1])
    acc = np.array([0.3, -0.1, -0.6])
    P.ox, P.oy, P.oz = Om
    P.udot, P.vdot, P.wdot = acc
    gp = np.array([0.05, 0.2, -0.7])
    o.s.gradP.x, o.s.gradP.y, o.s.gradP.z = gp
    c00 = 0.37
    o.array("pnm_re")[:] = 0.0
    o.array("pnm_re")[0] = c00
    rng = np.random.default_rng(2)
    Gc = d.Gcc
    o.array("p")[:] = rng.normal(size=Gc.s3b)
    old = o.array("p").copy().reshape(Gc.knb, Gc.jnb, Gc.inb)
    o.call("part_p_fill")
    got = o.array("p").reshape(Gc.knb, Gc.jnb, Gc.inb)
    ph = o.array("phase").reshape(Gc.knb, Gc.jnb, Gc.inb)
    rho, mu, nu, a = o.s.rho_f, o.s.mu, o.s.nu, P.r
    n00 = 1.0 / (2.0 * math.sqrt(math.pi))
    nsolid = 0
    for k in range(1, Gc.knb - 1):
        for j in range(1, Gc.jnb - 1):
            for i in range(1, Gc.inb - 1):
                if ph[k, j, i] < 0:
                    assert got[k, j, i] == old[k, j, i]
                    continue
                nsolid += 1
                pos = np.array([d.xs + (i - 0.5) * d.dx, d.ys + (j - 0.5) * d.dy,
                                d.zs + (k - 0.5) * d.dz])
                r = _min_image(pos - np.array([P.x, P.y, P.z]), d.xl)
                want = c00 * n00 * mu * nu / a ** 2 \
                    + 0.5 * rho * np.dot(np.cross(Om, r), np.cross(Om, r)) \
                    + np.dot(-gp - rho * acc, r)
                assert abs(got[k, j, i] - want) < 1e-13
    assert nsolid > 50


def test_part_bc_p_masks_rhs():
    """The particle pass on the Poisson right-hand side keeps it on fluid
    cells away from particles and zeroes it on solid cells and on the
    fluid side of the solid shell."""
    case = configs.small_particle_case(n=16, nparts=2, radius_dx=3.0)
    o, d = _state(case)
    o.call("build_cages")
    rng = np.random.default_rng(3)
    for c, G in (("u_star", d.Gfx), ("v_star", d.Gfy), ("w_star", d.Gfz)):
        o.array(c)[:] = rng.normal(size=G.s3b)
    o.s.dt = 1e-3
    o.call("PP_rhs")
    rhs = o.array("rhs_p").copy()
    o.call("part_BC_p")
    Gc = d.Gcc
    ph = o.array("phase").reshape(Gc.knb, Gc.jnb, Gc.inb)[1:-1, 1:-1, 1:-1].ravel()
    sh = o.array("phase_shell").reshape(Gc.knb, Gc.jnb, Gc.inb)[1:-1, 1:-1, 1:-1].ravel()
    keep = (ph < 0) & (sh > 0)
    got = o.array("rhs_p")
    assert np.array_equal(got[keep], rhs[keep])
    assert np.all(got[~keep] == 0.0)
    assert (~keep).sum() > 50


# ---------------------------------------------------------------------------
# sphere-wall forces (src/bluebottle_kernel.cu:2875-3590): lubrication
# against the asymptotic sphere-plane results, contact against Hertz
# ---------------------------------------------------------------------------
def _wall_sphere(h, a=None, wall="S", E=1.0e6, sigma=0.5, vel=(0.0, 0.0, 0.0),
                 omega=(0.0, 0.0, 0.0), dt=1e-4):
    """One sphere in the wall-bounded box (no-slip y walls, p Neumann), its
    surface a distance h from the south (y = 0) or north (y = 1) wall, with
    the given velocity and angular velocity (both time levels).  Returns the
    particle record after move_parts_sub's collision pass (not integrated:
    translating = rotating = 0)."""
    case = configs.channel_case(n=(16, 16, 16), nparts=1)
    P = case.particles[0]
    if a is None:
        a = P.r
    P.r = a
    P.x, P.z = 0.5, 0.5
    P.y = a + h if wall == "S" else 1.0 - a - h
    P.translating, P.rotating = 0, 0
    P.E, P.sigma = E, sigma
    # wall velocities zero, so that only the sphere's motion counts
    zero = tuple((t, 0.0) for t, _ in case.bc_u)
    case.bc_u = zero
    o, d = _state(case)
    o.s.dt = dt
    Q = o.s.parts[0]
    Q.u, Q.v, Q.w = vel
    Q.u0, Q.v0, Q.w0 = vel
    Q.ox, Q.oy, Q.oz = omega
    Q.ox0, Q.oy0, Q.oz0 = omega
    o.call("move_parts_sub")
    Q = o.parts()[0]
    rec = _Rec({f: getattr(Q, f) for f in ("y", "iFx", "iFy", "iFz", "iLx",
                                          "iLy", "iLz")})
    return o, rec, a


class _Rec(dict):
    """A copy of a particle record's fields (the oracle state's memory is
    released with the state)."""
    __getattr__ = dict.__getitem__


def test_wall_squeeze_leading_order():
    """Normal approach to a plane wall: F = -6 pi mu a^2 U / h at leading
    order in h/a (Brenner 1961; Cox and Brenner 1967), repulsive on
    approach for either wall."""
    U = 0.04
    for wall, sgn in (("S", -1.0), ("N", 1.0)):
        h = 0.02 * 0.125
        o, P, a = _wall_sphere(h, wall=wall, vel=(0.0, sgn * U, 0.0))
        mu = o.s.mu
        lead = -6.0 * math.pi * mu * a * a * (sgn * U) / h
        # the next order is O(h/a) (the interaction length is a)
        assert abs(P.iFy / lead - 1.0) < 2.0 * h / a
        assert P.iFy * sgn < 0   # opposes the approach
        assert P.iFx == 0.0 and P.iFz == 0.0
        assert P.iLx == 0.0 and P.iLy == 0.0 and P.iLz == 0.0
    # gaps below 1 % of the radius are treated as 1 % (the regularisation
    # of the reference: the force stays finite at contact)
    _, P1, a = _wall_sphere(0.01 * 0.125, vel=(0.0, -U, 0.0))
    _, P2, _ = _wall_sphere(0.002 * 0.125, vel=(0.0, -U, 0.0))
    assert abs(P1.iFy / P2.iFy - 1.0) < 1e-14


def test_wall_tangential_log_coefficients_and_reciprocity():
    """Tangential translation and rotation next to a plane wall: the force
    and torque grow like ln(a/h) with the Goldman, Cox and Brenner (1967)
    coefficients 6 pi mu a U (8/15) for the drag, 8 pi mu a^3 Omega (2/5)
    for the rotational torque, and the coupling 8 pi mu a^2 (1/10) in both
    directions (force from rotation equals torque from translation, the
    symmetry of the resistance matrix).  A sphere translating along +x above
    the south wall is turned as if rolling forward (torque about -z)."""
    U, Om = 0.03, 0.7
    h1, h2 = 0.02 * 0.125, 0.08 * 0.125
    res = {}
    for h in (h1, h2):
        o, Pt, a = _wall_sphere(h, vel=(U, 0.0, 0.0))
        _, Pr, _ = _wall_sphere(h, omega=(0.0, 0.0, Om))
        res[h] = (Pt, Pr)
    mu = o.s.mu
    dlog = math.log(h2 / h1)
    (Pt1, Pr1), (Pt2, Pr2) = res[h1], res[h2]
    # drag from translation
    assert abs((Pt2.iFx - Pt1.iFx) / dlog
               / (6 * math.pi * mu * a * U * 8 / 15) - 1) < 1e-12
    assert Pt1.iFx < 0
    # rotational torque
    assert abs((Pr2.iLz - Pr1.iLz) / dlog
               / (8 * math.pi * mu * a ** 3 * Om * 2 / 5) - 1) < 1e-12
    assert Pr1.iLz < 0
    # coupling: torque from translation (rolling sense) and force from rotation
    c_t = (Pt2.iLz - Pt1.iLz) / dlog / U
    c_r = (Pr2.iFx - Pr1.iFx) / dlog / Om
    assert abs(c_t / (8 * math.pi * mu * a * a / 10) - 1) < 1e-12
    assert abs(c_r / c_t - 1) < 1e-12
    assert Pt1.iLz < 0 and Pr1.iFx < 0
    # nothing out of the plane of the motion
    for P in (Pt1, Pr1):
        assert P.iFz == 0.0 and P.iLx == 0.0 and P.iLy == 0.0


def test_wall_hertz_contact_at_rest():
    """A sphere at rest pressed into either wall by delta: the Hertz force
    (4/3) E* sqrt(a) delta^(3/2) of a sphere on a flat of the same material,
    1/E* = 2 (1 - sigma^2) / E, pushing it away from the wall."""
    E, sigma = 3.0e5, 0.25
    delta_rel = 0.02
    for wall, away in (("S", 1.0), ("N", -1.0)):
        o, P, a = _wall_sphere(-delta_rel * 0.125, wall=wall, E=E, sigma=sigma)
        delta = delta_rel * 0.125
        estar = E / (2.0 * (1.0 - sigma * sigma))
        hertz = 4.0 / 3.0 * estar * math.sqrt(a) * delta ** 1.5
        assert abs(P.iFy * away / hertz - 1.0) < 1e-12, (wall, P.iFy, hertz)
        assert P.iFx == 0.0 and P.iFz == 0.0


def test_spring_is_hookes_law():
    """spring_parts (src/bluebottle_kernel.cu:3592-3617): a spring of rest
    length l anchored at X0 pulls with -k (|d| - l) d/|d|, d = x - X0."""
    case = _one_sphere()
    P = case.particles[0]
    P.x, P.y, P.z = 0.55, 0.47, 0.52
    P.spring_k, P.spring_x, P.spring_y, P.spring_z = 40.0, 0.5, 0.5, 0.5
    for l in (0.0, 0.03):
        P.spring_l = l
        o, d = _state(case)
        o.s.dt = 1e-4
        o.call("move_parts_sub")
        Q = o.parts()[0]
        got = np.array([Q.kFx, Q.kFy, Q.kFz])
        dv = np.array([P.x - 0.5, P.y - 0.5, P.z - 0.5])
        dn = np.linalg.norm(dv)
        want = -P.spring_k * (dn - l) * dv / dn
        assert np.allclose(got, want, rtol=1e-13, atol=0), (l, got, want)
