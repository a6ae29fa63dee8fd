"""Host-side set-up of a Bluebottle case on the C-ABI globals.

``Case`` holds what the reference reads from sim/input/flow.config and
part.config (src/domain.c:30-730, src/particle.c:79-159); ``Simulation``
writes it into the driver globals, initialises the domain and particles
(src/domain.c:1591-2347, src/particle.c:295-762), allocates and pushes the
device state, and drives the per-step loop of src/bluebottle.c:598-711
through libbluebottle_host.so.
"""
import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import abi, runtime


@dataclass
class Particle:
    """One entry of part.config (src/particle.c:96-155)."""
    x: float
    y: float
    z: float
    r: float
    rho: float = 2.0
    E: float = 1.0e6
    sigma: float = 0.5
    e_dry: float = 1.0
    coeff_fric: float = 0.5
    order: int = 4
    rs_r: float = 1.2
    spring_k: float = 0.0
    spring_x: float = 0.0
    spring_y: float = 0.0
    spring_z: float = 0.0
    spring_l: float = 0.0
    translating: int = 1
    rotating: int = 1
    aF: tuple = (0.0, 0.0, 0.0)
    aL: tuple = (0.0, 0.0, 0.0)


@dataclass
class Case:
    """A flow.config + part.config worth of parameters."""
    n: tuple = (16, 16, 16)
    lo: tuple = (0.0, 0.0, 0.0)
    hi: tuple = (1.0, 1.0, 1.0)
    rho_f: float = 1.0
    nu: float = 0.01
    CFL: float = 0.3
    pp_max_iter: int = 500
    pp_residual: float = 1e-10
    lamb_max_iter: int = 25
    lamb_residual: float = 1e-2
    lamb_relax: float = 1.0
    lamb_cut: float = 1e-4
    out_plane: int = abi.BOTTOM
    # per face W,E,S,N,B,T: (type, value) for p and u, v, w
    bc_p: tuple = ((abi.PERIODIC, 0.0),) * 6
    bc_u: tuple = ((abi.PERIODIC, 0.0),) * 6
    bc_v: tuple = ((abi.PERIODIC, 0.0),) * 6
    bc_w: tuple = ((abi.PERIODIC, 0.0),) * 6
    gradP: tuple = (0.0, 0.0, 0.0, 0.0, 0.0, 0.0)   # xm, xa, ym, ya, zm, za
    g: tuple = (0.0, 0.0, 0.0, 0.0, 0.0, 0.0)
    p_tDelay: float = 0.0
    g_tDelay: float = 0.0
    vel_tDelay: float = 0.0
    init: str = "taylor_green"       # or "quiescent", "shear"
    duration: float = 1.0
    noise: float = 1e-3
    seed: int = 1
    interaction_length_ratio: float = 1.0
    # Crank-Nicolson diffusion (the reference's -DIMPLICIT build)
    implicit: bool = False
    # the precursor domain (the reference's turbulence rank): the random
    # initial field of amplitude 3 turbA l drawn from `seed` instead of
    # `init`, the linear forcing turbA (k0/k)(u - <u>) every step, no
    # particles and no Lamb iteration
    precursor: bool = False
    turbA: float = 0.0
    particles: list = field(default_factory=list)


def splitmix64(seed, count):
    """Deterministic uint64 stream (splitmix64, Steele et al. 2014)."""
    mask = (1 << 64) - 1
    out = np.empty(count, dtype=np.uint64)
    s = seed & mask
    z = np.arange(1, count + 1, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15)
    z = z + np.uint64(s)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    out[:] = z
    return out


def normal_noise(seed, count):
    """N(0,1) samples by Box-Muller on a splitmix64 stream."""
    m = (count + 1) // 2
    u = splitmix64(seed, 2 * m)
    a = (u[:m] >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    b = (u[m:] >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    a = np.maximum(a, 1e-300)
    rad = np.sqrt(-2.0 * np.log(a))
    z = np.concatenate([rad * np.cos(2 * np.pi * b), rad * np.sin(2 * np.pi * b)])
    return z[:count]


class Simulation:
    """One case on the C-ABI state.  Only one may be alive at a time (the
    state is process-global, as in the reference)."""

    _alive = None

    def __init__(self, case: Case):
        if Simulation._alive is not None:
            Simulation._alive.close()
        self.case = case
        self.host, self.hip = runtime.load()
        self.device = False
        self._setup_host()
        Simulation._alive = self

    # ------------------------------------------------------------------
    def _setup_host(self):
        c = self.case
        rs = runtime.set_scalar
        rs("rho_f", c.rho_f)
        rs("nu", c.nu)
        rs("mu", c.nu * c.rho_f)
        rs("CFL", c.CFL)
        rs("duration", c.duration)
        rs("pp_max_iter", int(c.pp_max_iter))
        rs("pp_residual", c.pp_residual)
        rs("lamb_max_iter", int(c.lamb_max_iter))
        rs("lamb_residual", c.lamb_residual)
        rs("lamb_relax", c.lamb_relax)
        rs("lamb_cut", c.lamb_cut)
        rs("out_plane", int(c.out_plane))
        rs("init_cond", int(abi.QUIESCENT))
        rs("p_tDelay", c.p_tDelay)
        rs("g_tDelay", c.g_tDelay)
        rs("vel_tDelay", c.vel_tDelay)
        rs("interactionLengthRatio", c.interaction_length_ratio)
        rs("turbA", c.turbA)
        for k in ("Kp", "Ki", "Kd", "pid_int", "pid_back", "ttime"):
            rs(k, 0.0)
        b = runtime.bc()
        C.memset(C.addressof(b), 0, C.sizeof(b))
        for f, face in enumerate("WESNBT"):
            setattr(b, "p" + face, c.bc_p[f][0])
            setattr(b, "p" + face + "D", c.bc_p[f][1])
            for comp, spec in (("u", c.bc_u), ("v", c.bc_v), ("w", c.bc_w)):
                setattr(b, comp + face, spec[f][0])
                setattr(b, comp + face + "Dm", spec[f][1])
                setattr(b, comp + face + "D", spec[f][1])
                setattr(b, comp + face + "Da", 0.0)
        gp = runtime.gvar("gradP", abi.gradP_struct)
        gp.x, gp.y, gp.z = 0.0, 0.0, 0.0
        gp.xm, gp.xa, gp.ym, gp.ya, gp.zm, gp.za = c.gradP
        gg = runtime.gvar("g", abi.g_struct)
        gg.x, gg.y, gg.z = 0.0, 0.0, 0.0
        gg.xm, gg.xa, gg.ym, gg.ya, gg.zm, gg.za = c.g
        nx, ny, nz = c.n
        self.host.bb_host_domain(c.lo[0], c.hi[0], nx, c.lo[1], c.hi[1], ny,
                                 c.lo[2], c.hi[2], nz)
        self.dom = runtime.dom()
        if c.precursor:
            if c.particles:
                raise ValueError("the precursor domain carries no particles")
            self.host.bb_host_init_turb(int(c.seed))
        elif c.init == "taylor_green":
            self._taylor_green()
        elif c.init == "shear":
            self._shear()
        elif c.init != "quiescent":
            raise ValueError("initial condition %r is not set up here" % c.init)
        n = len(c.particles)
        self.host.bb_host_particles(n)
        if n:
            ptr = C.POINTER(abi.part_struct).in_dll(self.host, "parts")
            for i, p in enumerate(c.particles):
                P = ptr[i]
                P.x, P.y, P.z, P.r = p.x, p.y, p.z, p.r
                P.rho, P.E, P.sigma = p.rho, p.E, p.sigma
                P.e_dry, P.coeff_fric = p.e_dry, p.coeff_fric
                P.order, P.rs = p.order, p.rs_r
                P.spring_k, P.spring_l = p.spring_k, p.spring_l
                P.spring_x, P.spring_y, P.spring_z = p.spring_x, p.spring_y, p.spring_z
                P.translating, P.rotating = p.translating, p.rotating
                P.aFx, P.aFy, P.aFz = p.aF
                P.aLx, P.aLy, P.aLz = p.aL
        self.host.bb_host_particles_init()

    def _taylor_green(self):
        """u = sin(2 pi x) cos(2 pi y) cos(2 pi z),
        v = -cos(2 pi x) sin(2 pi y) cos(2 pi z), w = 0, plus seeded
        N(0, noise) on every face value (ghosts included)."""
        c, d = self.case, self.dom
        tp = 2.0 * np.pi
        # position of index i along an axis: (i - 1 + 0.5 - s) * h, with
        # s = 0.5 on the component's staggered axis (faces) and 0 across
        for comp, G, sx, sy, sz in (("u", d.Gfx, 0.5, 0.0, 0.0),
                                    ("v", d.Gfy, 0.0, 0.5, 0.0),
                                    ("w", d.Gfz, 0.0, 0.0, 0.5)):
            i = np.arange(G.inb) - 1 + 0.5 - sx
            j = np.arange(G.jnb) - 1 + 0.5 - sy
            k = np.arange(G.knb) - 1 + 0.5 - sz
            x = (d.xs + i * d.dx)[None, None, :]
            y = (d.ys + j * d.dy)[None, :, None]
            z = (d.zs + k * d.dz)[:, None, None]
            if comp == "u":
                f = np.sin(tp * x) * np.cos(tp * y) * np.cos(tp * z)
            elif comp == "v":
                f = -np.cos(tp * x) * np.sin(tp * y) * np.cos(tp * z)
            else:
                f = np.zeros((len(k), len(j), len(i)))
            f = f.reshape(-1)
            seed = c.seed * 3 + {"u": 0, "v": 1, "w": 2}[comp]
            f = f + c.noise * normal_noise(seed, f.size)
            self.array(comp)[:] = f
            self.array(comp + "0")[:] = f

    def _shear(self):
        """Linear profiles between the Dirichlet wall velocities
        (src/domain.c:2205-2241, init_cond SHEAR; as there, v's z term
        divides by ze rather than the length)."""
        c, d = self.case, self.dom
        val = lambda spec, f: spec[f][1] if spec[f][0] == abi.DIRICHLET else 0.0
        W, E, S, N, B, T = range(6)
        for comp, G in (("u", d.Gfx), ("v", d.Gfy), ("w", d.Gfz)):
            i = np.arange(G.inb)
            j = np.arange(G.jnb)
            k = np.arange(G.knb)
            x = ((i - 0.5) * d.dx + d.xs)[None, None, :]
            y = ((j - 0.5) * d.dy + d.ys)[None, :, None]
            z = ((k - 0.5) * d.dz + d.zs)[:, None, None]
            shape = (len(k), len(j), len(i))
            if comp == "u":
                sp = c.bc_u
                f = ((val(sp, N) - val(sp, S)) * (y - d.ys) / d.yl + val(sp, S)
                     + (val(sp, T) - val(sp, B)) * (z - d.zs) / d.zl
                     + val(sp, B))
            elif comp == "v":
                sp = c.bc_v
                f = ((val(sp, E) - val(sp, W)) * (x - d.xs) / d.xl + val(sp, W)
                     + (val(sp, T) - val(sp, B)) * (z - d.zs) / d.ze
                     + val(sp, B))
            else:
                sp = c.bc_w
                f = ((val(sp, E) - val(sp, W)) * (x - d.xs) / d.xl + val(sp, W)
                     + (val(sp, N) - val(sp, S)) * (y - d.ys) / d.yl
                     + val(sp, S))
            f = np.broadcast_to(f, shape).reshape(-1)
            self.array(comp)[:] = f
            self.array(comp + "0")[:] = f

    # ------------------------------------------------------------------
    # precursor inflow planes: grid and the axis each is normal to
    PLANES = {"u_WE": ("u", 0), "u_SN_S": ("u", 1), "u_SN_N": ("u", 1),
              "u_BT_B": ("u", 2), "u_BT_T": ("u", 2),
              "v_WE_W": ("v", 0), "v_WE_E": ("v", 0), "v_SN": ("v", 1),
              "v_BT_B": ("v", 2), "v_BT_T": ("v", 2),
              "w_WE_W": ("w", 0), "w_WE_E": ("w", 0), "w_SN_S": ("w", 1),
              "w_SN_N": ("w", 1), "w_BT": ("w", 2)}

    def plane_shape(self, name):
        """(rows, row length) of a plane: W/E (knb, jnb), S/N (inb, knb),
        B/T (jnb, inb) of its component's grid."""
        comp, axis = self.PLANES[name]
        G = self.grid(comp)
        return ((G.knb, G.jnb), (G.inb, G.knb), (G.jnb, G.inb))[axis]

    def count(self, name):
        if name.startswith(("pnm", "phinm", "chinm")):
            return runtime.get_scalar("coeff_stride", "i") * len(self.case.particles)
        if name in self.PLANES:
            a, b = self.plane_shape(name)
            return a * b
        return self.grid(name).s3b

    def array(self, name):
        dtype = np.int32 if name in ("phase", "phase_shell", "flag_u", "flag_v",
                                     "flag_w") else np.float64
        return runtime.array_view(name, self.count(name), dtype)

    def grid(self, name):
        d = self.dom
        if name in ("p", "p0", "phi", "phase", "phase_shell"):
            return d.Gcc
        if name in ("u", "u0", "u_star", "f_x", "flag_u", "diff_u", "diff0_u",
                    "conv_u", "conv0_u"):
            return d.Gfx
        if name in ("v", "v0", "v_star", "f_y", "flag_v", "diff_v", "diff0_v",
                    "conv_v", "conv0_v"):
            return d.Gfy
        if name in ("w", "w0", "w_star", "f_z", "flag_w", "diff_w", "diff0_w",
                    "conv_w", "conv0_w"):
            return d.Gfz
        raise KeyError(name)

    def field3(self, name):
        """(k, j, i) view of a ghosted field."""
        G = self.grid(name)
        return self.array(name).reshape(G.knb, G.jnb, G.inb)

    def parts(self):
        n = len(self.case.particles)
        ptr = C.POINTER(abi.part_struct).in_dll(self.host, "parts")
        return [ptr[i] for i in range(n)]

    # ------------------------------------------------------------------
    def to_device(self):
        h = self.hip
        h.bb_set_implicit(1 if self.case.implicit else 0)
        h.cuda_dom_malloc()
        h.cuda_part_malloc()
        h.cuda_dom_push()
        h.cuda_part_push()
        self.device = True

    def prepare(self):
        if self.case.precursor:
            self.host.bb_host_prepare_turb()
        else:
            self.host.bb_host_prepare()

    def restart(self, path):
        """Continue from a restart file instead of prepare() (the state
        must be on the device already)."""
        fn = (self.host.bb_host_restart_turb if self.case.precursor
              else self.host.bb_host_restart)
        if fn(str(path).encode()) != 0:
            raise RuntimeError("could not read restart file %s" % path)

    def write_restart(self, path):
        """Pull the device state and write it as a restart file (the
        precursor domain's layout for a precursor case)."""
        self.pull()
        fn = (self.host.bb_out_restart_turb if self.case.precursor
              else self.host.bb_out_restart)
        if fn(str(path).encode()) != 0:
            raise RuntimeError("could not write restart file %s" % path)

    def write_vtk(self, directory, stepnum):
        """Pull the device state (when there is one) and write ParaView
        output number `stepnum` into `directory` (out.pvd, out_N.pvtr,
        out_N_0.vtr, out_N.vtp, out_nodes_N.vtp)."""
        if self.device:
            self.pull()
        if self.host.bb_vtk_out(str(directory).encode(), int(stepnum)) != 0:
            raise RuntimeError("could not write VTK output to %s" % directory)

    def step(self):
        """One time step; returns the Lamb sub-iteration count.  Raises
        when the new dt fails the driver's blow-up test, as the reference
        ends the run there."""
        if self.case.precursor:
            it = self.host.bb_host_step_turb()
        else:
            it = self.host.bb_host_step()
        if it < 0:
            raise RuntimeError("the solution has diverged (dt = %g)"
                               % runtime.get_scalar("dt"))
        return it

    def pull(self):
        self.hip.cuda_dom_pull()
        self.hip.cuda_part_pull()

    def push(self):
        self.hip.cuda_dom_push()
        self.hip.cuda_part_push()

    def set_solver(self, mode):
        self.hip.bb_set_solver_mode(int(mode))

    def solve_stats(self):
        st = abi.bb_solve_stats()
        self.hip.bb_get_solve_stats(C.byref(st))
        return st

    def close(self):
        if Simulation._alive is not self:
            return
        if self.device:
            self.hip.cuda_part_free()
            self.hip.cuda_dom_free()
            self.device = False
        self.host.bb_host_free()
        if Simulation._alive is self:
            Simulation._alive = None
