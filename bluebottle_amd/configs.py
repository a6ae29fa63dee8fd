"""Synthetic cases: the five benchmark configurations and small test cases.

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


def channel_case(n=(16, 12, 16), nparts=0, seed=5):
    """Wall-bounded shear box: no-slip y walls (Dirichlet u = +-U_w), p
    Neumann, periodic x and z."""
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
                            periodic=(1, 0, 1), wall_gap=1.0 / nx)
        parts = [Particle(x=float(p[0]), y=float(p[1]), z=float(p[2]), r=a,
                          rho=1.5) for p in pos]
    return Case(n=(nx, ny, nz), nu=0.01, CFL=0.3, pp_max_iter=2000,
                pp_residual=1e-11, lamb_max_iter=2, bc_p=bc_p, bc_u=bc_u,
                bc_v=bc_v, bc_w=bc_w, out_plane=abi.NORTH,
                init="taylor_green", noise=1e-3, seed=seed, particles=parts)
