"""Synthetic cases: the benchmark configurations and small test cases, and
their sim/input files (flow.config, part.config) in the reference's
grammar.

Configurations: 1 `config1` (48^3 Taylor-Green box, one fixed sphere),
2 `config2` (128^3, 64 settling spheres), 3 `config3` (256^3, 1000
settling spheres), 4 the learned preconditioner on configuration 3's flow
(no case of its own), 5 `config5` (384 x 192 x 384 wall-bounded shear box,
8000 spheres).

Physical scales (the configurations state Re but not nu or |g|):
  * Taylor-Green cases: unit periodic box, U = L = 1, nu = 1/Re; the
    Taylor-Green field uses 2 pi x so that it is periodic on the unit box.
  * Sedimentation cases: |g| = 1 along -z, rho_f = 1, and nu from the
    particle Reynolds number Re_p = (4/9)(rho_p/rho_f - 1) g a^3 / nu^2
    (Stokes settling velocity times diameter over nu).  The mean pressure
    gradient balances the suspension's excess weight:
    gradP.z = (rho_p - rho_f) phi g_z, so the box does not accelerate.
Particle positions are drawn by random sequential addition with a minimum
surface gap, from a splitmix64 stream with the configuration's seed.
"""
import math

import numpy as np

from . import abi
from .sim import Case, Particle, splitmix64


def uniform01(seed, count):
    u = splitmix64(seed, count)
    return (u >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def place_spheres(n_parts, radius, lo, hi, min_gap, seed, periodic=(1, 1, 1),
                  wall_gap=None, max_tries=2_000_000):
    """Seeded random sequential addition of non-overlapping spheres."""
    lo = np.asarray(lo, float)
    hi = np.asarray(hi, float)
    L = hi - lo
    stream = uniform01(seed, 3 * max_tries)
    pos = np.zeros((0, 3))
    dmin = 2.0 * radius + min_gap
    tries = 0
    out = []
    while len(out) < n_parts:
        if tries >= max_tries:
            raise RuntimeError("could not place %d spheres" % n_parts)
        c = lo + L * stream[3 * tries:3 * tries + 3]
        tries += 1
        if wall_gap is not None:
            ok = True
            for a in range(3):
                if not periodic[a] and (c[a] - lo[a] < radius + wall_gap
                                        or hi[a] - c[a] < radius + wall_gap):
                    ok = False
            if not ok:
                continue
        if len(out):
            d = pos - c
            for a in range(3):
                if periodic[a]:
                    d[:, a] -= L[a] * np.round(d[:, a] / L[a])
            if np.min(np.einsum("ij,ij->i", d, d)) < dmin * dmin:
                continue
        out.append(c)
        pos = np.vstack([pos, c[None, :]])
    return pos


def periodic_case(n, nu, **kw):
    return Case(n=(n, n, n), nu=nu, **kw)


def config1(n=48, lamb_max_iter=3, pp_residual=1e-10, nparts=1, order=4):
    """48^3 Taylor-Green, Re = 100, one fixed sphere of radius 4 dx at the
    centre, L = 4, Jacobi BiCGSTAB at rel 1e-10, max 500."""
    dx = 1.0 / n
    parts = []
    if nparts:
        parts = [Particle(x=0.5, y=0.5, z=0.5, r=4 * dx, rho=2.0, order=order,
                          translating=0, rotating=0)]
    return Case(n=(n, n, n), nu=1.0 / 100.0, CFL=0.3, pp_max_iter=500,
                pp_residual=pp_residual, lamb_max_iter=lamb_max_iter,
                init="taylor_green", noise=1e-3, seed=1, particles=parts)


def taylor_green(n=256, pp_residual=1e-8, pp_max_iter=20000, re=100.0):
    """Sphere-free periodic Taylor-Green box (configuration 1's flow and
    noise) at n^3, for timing the fluid path alone."""
    return Case(n=(n, n, n), nu=1.0 / re, CFL=0.3, pp_max_iter=pp_max_iter,
                pp_residual=pp_residual, lamb_max_iter=1,
                init="taylor_green", noise=1e-3, seed=1, particles=[])


def sedimentation(n, nparts, radius_dx, density_ratio, re_p, seed,
                  pp_residual=1e-8, lamb_max_iter=25, min_gap_dx=1.0,
                  order=4, pp_max_iter=20000):
    dx = 1.0 / n
    a = radius_dx * dx
    g = 1.0
    nu = math.sqrt((4.0 / 9.0) * (density_ratio - 1.0) * g * a ** 3 / re_p)
    pos = place_spheres(nparts, a, (0, 0, 0), (1, 1, 1), min_gap_dx * dx, seed)
    phi = nparts * 4.0 / 3.0 * math.pi * a ** 3
    parts = [Particle(x=float(p[0]), y=float(p[1]), z=float(p[2]), r=a,
                      rho=density_ratio, order=order) for p in pos]
    gz = -g
    gradPz = (density_ratio - 1.0) * phi * gz
    return Case(n=(n, n, n), nu=nu, CFL=0.3, pp_max_iter=pp_max_iter,
                pp_residual=pp_residual, lamb_max_iter=lamb_max_iter,
                init="quiescent", seed=seed, particles=parts,
                gradP=(0.0, 0.0, 0.0, 0.0, gradPz, 0.0),
                g=(0.0, 0.0, 0.0, 0.0, gz, 0.0))


def config2(n=128, nparts=64):
    """128^3 periodic, 64 spheres of radius 5 dx, density ratio 2, gravity
    -z, L = 4.  Re = 200 is read as the particle Reynolds number scale of
    the settling problem (Re_p = 200 gives nu from the Stokes velocity)."""
    return sedimentation(n, nparts, 5.0, 2.0, 200.0, seed=2)


def config3(n=256, nparts=1000, lamb_max_iter=25):
    """256^3 periodic, 1000 spheres of radius 6 dx (phi = 5.4 %), density
    ratio 2, Re_p = 10, L = 4, 26 nodes, rel tol 1e-8."""
    return sedimentation(n, nparts, 6.0, 2.0, 10.0, seed=3,
                         lamb_max_iter=lamb_max_iter)


def small_particle_case(n=24, nparts=3, radius_dx=3.0, seed=7, order=4,
                        lamb_max_iter=2, pp_residual=1e-11, taylor_green=True):
    """A few moving spheres in a small periodic Taylor-Green box."""
    dx = 1.0 / n
    a = radius_dx * dx
    pos = place_spheres(nparts, a, (0, 0, 0), (1, 1, 1), 2 * dx, seed)
    parts = [Particle(x=float(p[0]), y=float(p[1]), z=float(p[2]), r=a,
                      rho=2.0, order=order) for p in pos]
    return Case(n=(n, n, n), nu=0.01, CFL=0.3, pp_max_iter=2000,
                pp_residual=pp_residual, lamb_max_iter=lamb_max_iter,
                init="taylor_green" if taylor_green else "quiescent",
                noise=1e-3, seed=seed, particles=parts,
                g=(0.0, 0.0, 0.0, 0.0, -1.0, 0.0),
                gradP=(0.0, 0.0, 0.0, 0.0, -0.05, 0.0))


def wide_case(n=(72, 32, 16), nparts=2, seed=9):
    """A periodic box of cubic cells, longer in x than one tile of the
    Poisson sweeps (64 cells), with a few spheres of radius 3 dx."""
    nx, ny, nz = n
    dx = 1.0 / nx
    hi = (1.0, ny * dx, nz * dx)
    a = 3.0 * dx
    pos = place_spheres(nparts, a, (0, 0, 0), hi, 2 * dx, seed)
    parts = [Particle(x=float(p[0]), y=float(p[1]), z=float(p[2]), r=a,
                      rho=2.0) for p in pos]
    return Case(n=(nx, ny, nz), hi=hi, nu=0.01, CFL=0.3, pp_max_iter=4000,
                pp_residual=1e-11, lamb_max_iter=2, init="taylor_green",
                noise=1e-3, seed=seed, particles=parts,
                g=(0.0, 0.0, 0.0, 0.0, -1.0, 0.0),
                gradP=(0.0, 0.0, 0.0, 0.0, -0.05, 0.0))


def config5(nx=384, ny=192, nz=384, nparts=8000, lamb_max_iter=25,
              pp_residual=1e-8):
    """Dense sheared channel: x in [0, 2], y in [0, 1], z in [0, 2] with
    cubic cells (384 x 192 x 384), no-slip y walls moving at -U/2 and +U/2
    (U = 1), p Neumann at the walls, periodic x and z; Re_bulk = U H / nu
    = 2000 (H = 1, the wall distance); 8000 spheres of radius 4 dx at seeded
    random positions at least 1 dx apart and 1 dx from the walls (volume
    fraction 7.6 %), density ratio 1.5, no gravity; L = 4; the initial
    flow is the linear shear between the walls (init_cond SHEAR)."""
    dx = 1.0 / ny
    lo, hi = (0.0, 0.0, 0.0), (nx * dx, ny * dx, nz * dx)
    a = 4.0 * dx
    uw = 0.5
    D, N, P = abi.DIRICHLET, abi.NEUMANN, abi.PERIODIC
    pos = place_spheres(nparts, a, lo, hi, dx, 5, periodic=(1, 0, 1),
                        wall_gap=dx)
    parts = [Particle(x=float(p[0]), y=float(p[1]), z=float(p[2]), r=a,
                      rho=1.5, order=4) for p in pos]
    return Case(n=(nx, ny, nz), lo=lo, hi=hi, nu=1.0 / 2000.0, CFL=0.3,
                pp_max_iter=20000, pp_residual=pp_residual,
                lamb_max_iter=lamb_max_iter,
                bc_p=((P, 0.), (P, 0.), (N, 0.), (N, 0.), (P, 0.), (P, 0.)),
                bc_u=((P, 0.), (P, 0.), (D, -uw), (D, uw), (P, 0.), (P, 0.)),
                bc_v=((P, 0.), (P, 0.), (D, 0.), (D, 0.), (P, 0.), (P, 0.)),
                bc_w=((P, 0.), (P, 0.), (D, 0.), (D, 0.), (P, 0.), (P, 0.)),
                out_plane=abi.NORTH, init="shear", seed=5, particles=parts)


def channel_case(n=(16, 12, 16), nparts=0, seed=5, init="taylor_green"):
    """Wall-bounded shear box: no-slip y walls (Dirichlet u = +-U_w), p
    Neumann, periodic x and z.  Spheres of radius 2 dx are placed at least
    one radius from either wall: nearer, or started in a Taylor-Green or
    linear shear field at this resolution, the coupled step diverges
    within a few steps (dt falls below 1e-10), in the oracle as on the
    device, so multi-step runs of this case start quiescent."""
    nx, ny, nz = n
    uw = 0.5
    D, N, P = abi.DIRICHLET, abi.NEUMANN, abi.PERIODIC
    bc_p = ((P, 0.), (P, 0.), (N, 0.), (N, 0.), (P, 0.), (P, 0.))
    bc_u = ((P, 0.), (P, 0.), (D, -uw), (D, uw), (P, 0.), (P, 0.))
    bc_v = ((P, 0.), (P, 0.), (D, 0.), (D, 0.), (P, 0.), (P, 0.))
    bc_w = ((P, 0.), (P, 0.), (D, 0.), (D, 0.), (P, 0.), (P, 0.))
    parts = []
    if nparts:
        a = 2.0 / nx
        pos = place_spheres(nparts, a, (0, 0, 0), (1, 1, 1), 1.0 / nx, seed,
                            periodic=(1, 0, 1), wall_gap=a)
        parts = [Particle(x=float(p[0]), y=float(p[1]), z=float(p[2]), r=a,
                          rho=1.5) for p in pos]
    return Case(n=(nx, ny, nz), nu=0.01, CFL=0.3, pp_max_iter=2000,
                pp_residual=1e-11, lamb_max_iter=2, bc_p=bc_p, bc_u=bc_u,
                bc_v=bc_v, bc_w=bc_w, out_plane=abi.NORTH,
                init=init, noise=1e-3, seed=seed, particles=parts)


# ---------------------------------------------------------------------------
# sim/input files (the reference's grammar: src/domain.c:30-730,
# src/particle.c:79-159)
# ---------------------------------------------------------------------------
_FACES = "WESNBT"
_BC_NAME = {abi.PERIODIC: "PERIODIC", abi.DIRICHLET: "DIRICHLET",
            abi.NEUMANN: "NEUMANN", abi.PRECURSOR: "PRECURSOR"}
_INIT_NAME = {"quiescent": "QUIESCENT", "shear": "SHEAR",
              "taylor_green": "TAYLOR_GREEN"}
_INIT_CODE = {abi.QUIESCENT: "quiescent", abi.SHEAR: "shear",
              100: "taylor_green"}
_PLANE_NAME = {abi.WEST: "WEST", abi.EAST: "EAST", abi.SOUTH: "SOUTH",
               abi.NORTH: "NORTH", abi.BOTTOM: "BOTTOM", abi.TOP: "TOP",
               abi.HOMOGENEOUS: "HOMOGENEOUS"}


def _g(x):
    return "%.17g" % x


def write_case(case, root):
    """Write root/input/flow.config and root/input/part.config for `case`.
    The Taylor-Green noise amplitude and seed have no place in the grammar
    (init_cond TAYLOR_GREEN is this build's extension)."""
    import os
    os.makedirs(os.path.join(root, "input"), exist_ok=True)
    c = case
    L = []
    L.append("DOMAIN")
    ext = []
    for ax, name in enumerate("XYZ"):
        ext.append("(%ss, %se, %sn) %s %s %d" % (name, name, name, _g(c.lo[ax]),
                                                 _g(c.hi[ax]), c.n[ax]))
    L += ext + ["", "GPU DOMAIN DECOMPOSITION", "DEV RANGE 0 0", "n 1"] + ext
    L += ["E -1 W -1 N -1 S -1 T -1 B -1", "", "PHYSICAL PARAMETERS",
          "rho_f " + _g(c.rho_f), "nu " + _g(c.nu), "", "SIMULATION PARAMETERS",
          "duration " + _g(c.duration), "CFL " + _g(c.CFL),
          "pp_max_iter %d" % c.pp_max_iter,
          "pp_residual " + _g(c.pp_residual),
          "lamb_max_iter %d" % c.lamb_max_iter,
          "lamb_residual " + _g(c.lamb_residual),
          "lamb_relax " + _g(c.lamb_relax), "lamb_cut " + _g(c.lamb_cut), "",
          "BOUNDARY CONDITIONS", "vel_tDelay " + _g(c.vel_tDelay)]
    for title, comp, spec in (("PRESSURE", "p", c.bc_p),
                              ("X-VELOCITY", "u", c.bc_u),
                              ("Y-VELOCITY", "v", c.bc_v),
                              ("Z-VELOCITY", "w", c.bc_w)):
        L.append(title)
        for f, face in enumerate(_FACES):
            t, val = spec[f]
            line = "bc.%s%s %s" % (comp, face, _BC_NAME[t])
            if comp == "p" and t == abi.NEUMANN:
                line += " 0"
            elif comp != "p" and t in (abi.DIRICHLET, abi.PRECURSOR):
                line += " %s 0" % _g(val)
            L.append(line)
    gp, gg = c.gradP, c.g
    L += ["", "INITIAL CONDITION", "init_cond " + _INIT_NAME[c.init], "",
          "SOLVABILITY ENFORCEMENT PLANE",
          "out_plane " + _PLANE_NAME[c.out_plane], "",
          "SIMULATION DRIVING CONDITIONS", "p_tDelay " + _g(c.p_tDelay),
          "gradP.x %s %s" % (_g(gp[0]), _g(gp[1])),
          "gradP.y %s %s" % (_g(gp[2]), _g(gp[3])),
          "gradP.z %s %s" % (_g(gp[4]), _g(gp[5])),
          "g_tDelay " + _g(c.g_tDelay),
          "g.x %s %s" % (_g(gg[0]), _g(gg[1])),
          "g.y %s %s" % (_g(gg[2]), _g(gg[3])),
          "g.z %s %s" % (_g(gg[4]), _g(gg[5])), "",
          "PID CONTROLLER GAINS", "Kp 0", "Ki 0", "Kd 0"]
    with open(os.path.join(root, "input", "flow.config"), "w") as f:
        f.write("\n".join(L) + "\n")
    P = ["n %d" % len(c.particles), "(l/a) " + _g(c.interaction_length_ratio)]
    for p in c.particles:
        P += ["", "r " + _g(p.r),
              "(x, y, z) %s %s %s" % (_g(p.x), _g(p.y), _g(p.z)),
              "(aFx, aFy, aFz) %s %s %s" % tuple(_g(v) for v in p.aF),
              "(aLx, aLy, aLz) %s %s %s" % tuple(_g(v) for v in p.aL),
              "rho " + _g(p.rho), "E " + _g(p.E), "sigma " + _g(p.sigma),
              "e_dry " + _g(p.e_dry), "coeff_fric " + _g(p.coeff_fric),
              "order %d" % p.order, "rs/r " + _g(p.rs_r),
              "spring_k " + _g(p.spring_k),
              "spring (x, y, z) %s %s %s" % (_g(p.spring_x), _g(p.spring_y),
                                             _g(p.spring_z)),
              "spring_l " + _g(p.spring_l),
              "translating %d" % p.translating, "rotating %d" % p.rotating]
    with open(os.path.join(root, "input", "part.config"), "w") as f:
        f.write("\n".join(P) + "\n")


def write_turb_config(case, root):
    """Write root/input/turb.config, the precursor domain's file
    (src/domain.c:732-1287): the flow.config grammar without the Lamb
    parameters, vel_tDelay, the initial condition and the PID gains, and
    turbA as its only driving condition."""
    import os
    os.makedirs(os.path.join(root, "input"), exist_ok=True)
    c = case
    ext = ["(%ss, %se, %sn) %s %s %d" % (nm, nm, nm, _g(c.lo[ax]), _g(c.hi[ax]),
                                         c.n[ax])
           for ax, nm in enumerate("XYZ")]
    L = ["DOMAIN"] + ext + ["", "GPU DOMAIN DECOMPOSITION", "DEV RANGE 0 0",
                            "n 1"] + ext
    L += ["E -1 W -1 N -1 S -1 T -1 B -1", "", "PHYSICAL PARAMETERS",
          "rho_f " + _g(c.rho_f), "nu " + _g(c.nu), "", "SIMULATION PARAMETERS",
          "duration " + _g(c.duration), "CFL " + _g(c.CFL),
          "pp_max_iter %d" % c.pp_max_iter,
          "pp_residual " + _g(c.pp_residual), "", "BOUNDARY CONDITIONS"]
    for title, comp, spec in (("PRESSURE", "p", c.bc_p),
                              ("X-VELOCITY", "u", c.bc_u),
                              ("Y-VELOCITY", "v", c.bc_v),
                              ("Z-VELOCITY", "w", c.bc_w)):
        L.append(title)
        for f, face in enumerate(_FACES):
            t, val = spec[f]
            if t == abi.PRECURSOR:
                raise ValueError("the precursor domain has no PRECURSOR faces")
            line = "bc.%s%s %s" % (comp, face, _BC_NAME[t])
            if comp == "p" and t == abi.NEUMANN:
                line += " 0"
            elif comp != "p" and t == abi.DIRICHLET:
                line += " %s 0" % _g(val)
            L.append(line)
    L += ["", "SOLVABILITY ENFORCEMENT PLANE",
          "out_plane " + _PLANE_NAME[c.out_plane], "",
          "SIMULATION DRIVING CONDITIONS", "turbA " + _g(c.turbA)]
    with open(os.path.join(root, "input", "turb.config"), "w") as f:
        f.write("\n".join(L) + "\n")


def read_turb_case(root, seed=1):
    """A precursor-domain Case from root/input/turb.config, parsed by the
    driver library's reader; its random initial field is drawn from
    `seed` (the reference seeds from the clock)."""
    import os
    from . import runtime
    host, _ = runtime.load()
    path = os.path.join(root, "input", "turb.config").encode()
    if host.bb_read_turb_config(path) != 0:
        raise ValueError("could not read %s" % path.decode())
    D = runtime.gvar("Dom", abi.dom_struct)
    b = runtime.bc()

    def spec(comp):
        out = []
        for face in _FACES:
            t = getattr(b, comp + face)
            v = getattr(b, comp + face + "Dm") if comp != "p" and \
                t == abi.DIRICHLET else 0.0
            out.append((t, v))
        return tuple(out)

    gs = runtime.get_scalar
    return Case(n=(D.xn, D.yn, D.zn), lo=(D.xs, D.ys, D.zs),
                hi=(D.xe, D.ye, D.ze), rho_f=gs("rho_f"), nu=gs("nu"),
                CFL=gs("CFL"), pp_max_iter=gs("pp_max_iter", "i"),
                pp_residual=gs("pp_residual"), lamb_max_iter=0,
                lamb_residual=0.0, lamb_relax=0.0, lamb_cut=0.0,
                out_plane=gs("out_plane", "i"), bc_p=spec("p"),
                bc_u=spec("u"), bc_v=spec("v"), bc_w=spec("w"),
                duration=gs("duration"), init="quiescent", seed=seed,
                precursor=True, turbA=gs("turbA"))


def read_case(root, noise=1e-3, seed=1):
    """A Case from root/input/flow.config and root/input/part.config,
    parsed by the driver library's readers (the reference's grammar)."""
    import ctypes as C
    import os
    from . import runtime
    host, _ = runtime.load()
    for name, fn in (("flow.config", host.bb_read_flow_config),
                     ("part.config", host.bb_read_part_config)):
        path = os.path.join(root, "input", name).encode()
        if fn(path) != 0:
            raise ValueError("could not read %s" % path.decode())
    D = runtime.gvar("Dom", abi.dom_struct)
    b = runtime.bc()

    def spec(comp):
        # the grammar gives a value only to Dirichlet (and precursor) faces;
        # the globals keep whatever an earlier read left for the others
        out = []
        for face in _FACES:
            t = getattr(b, comp + face)
            v = getattr(b, comp + face + "Dm") if comp != "p" and t in (
                abi.DIRICHLET, abi.PRECURSOR) else 0.0
            out.append((t, v))
        return tuple(out)

    gs = runtime.get_scalar
    gp = runtime.gvar("gradP", abi.gradP_struct)
    gg = runtime.gvar("g", abi.g_struct)
    init = _INIT_CODE.get(runtime.get_scalar("init_cond", "i"))
    if init is None:
        raise ValueError("init_cond %d is not set up here"
                         % runtime.get_scalar("init_cond", "i"))
    n = gs("nparts", "i")
    ptr = C.POINTER(abi.part_struct).in_dll(host, "parts")
    parts = []
    for i in range(n):
        q = ptr[i]
        parts.append(Particle(
            x=q.x, y=q.y, z=q.z, r=q.r, rho=q.rho, E=q.E, sigma=q.sigma,
            e_dry=q.e_dry, coeff_fric=q.coeff_fric, order=q.order,
            rs_r=q.rs, spring_k=q.spring_k, spring_x=q.spring_x,
            spring_y=q.spring_y, spring_z=q.spring_z, spring_l=q.spring_l,
            translating=q.translating, rotating=q.rotating,
            aF=(q.aFx, q.aFy, q.aFz), aL=(q.aLx, q.aLy, q.aLz)))
    return Case(n=(D.xn, D.yn, D.zn), lo=(D.xs, D.ys, D.zs),
                hi=(D.xe, D.ye, D.ze), rho_f=gs("rho_f"), nu=gs("nu"),
                CFL=gs("CFL"), pp_max_iter=gs("pp_max_iter", "i"),
                pp_residual=gs("pp_residual"),
                lamb_max_iter=gs("lamb_max_iter", "i"),
                lamb_residual=gs("lamb_residual"),
                lamb_relax=gs("lamb_relax"), lamb_cut=gs("lamb_cut"),
                out_plane=gs("out_plane", "i"), bc_p=spec("p"),
                bc_u=spec("u"), bc_v=spec("v"), bc_w=spec("w"),
                gradP=(gp.xm, gp.xa, gp.ym, gp.ya, gp.zm, gp.za),
                g=(gg.xm, gg.xa, gg.ym, gg.ya, gg.zm, gg.za),
                p_tDelay=gs("p_tDelay"), g_tDelay=gs("g_tDelay"),
                vel_tDelay=gs("vel_tDelay"), init=init, noise=noise,
                seed=seed, duration=gs("duration"),
                interaction_length_ratio=gs("interactionLengthRatio"),
                particles=parts)


CONFIGS = {"config1": config1, "config2": config2, "config3": config3,
           "config5": config5}


def main(argv=None):
    """python -m bluebottle_amd.configs NAME ROOT: write the seeded inputs of
    configuration NAME (config1, config2, config3, config5) as
    ROOT/input/flow.config and ROOT/input/part.config."""
    import sys
    args = sys.argv[1:] if argv is None else argv
    if len(args) != 2 or args[0] not in CONFIGS:
        print("usage: python -m bluebottle_amd.configs {%s} ROOT"
              % ",".join(sorted(CONFIGS)), file=sys.stderr)
        return 2
    write_case(CONFIGS[args[0]](), args[1])
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
