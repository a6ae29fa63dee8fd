"""ctypes mirrors of the C ABI structs in include/bluebottle.h.

The layouts are those of the reference's headers (src/domain.h:39-212,
src/bluebottle.h:2061-2322, src/particle.h:190-326) with ``real == double``;
``check_sizes`` compares them with the sizes the C compiler reports.
"""
import ctypes as C

real = C.c_double

_GRID_FIELDS = [
    "is", "ie", "in", "isb", "ieb", "inb",
    "js", "je", "jn", "jsb", "jeb", "jnb",
    "ks", "ke", "kn", "ksb", "keb", "knb",
    "s1", "s1b", "s2", "s2b", "s3", "s3b",
]


class grid_info(C.Structure):
    _fields_ = ([(n, C.c_int) for n in _GRID_FIELDS]
                + [("_" + n, C.c_int) for n in _GRID_FIELDS])


class dom_struct(C.Structure):
    _fields_ = [
        ("Gcc", grid_info), ("Gfx", grid_info), ("Gfy", grid_info),
        ("Gfz", grid_info),
        ("xs", real), ("xe", real), ("xl", real), ("xn", C.c_int), ("dx", real),
        ("ys", real), ("ye", real), ("yl", real), ("yn", C.c_int), ("dy", real),
        ("zs", real), ("ze", real), ("zl", real), ("zn", C.c_int), ("dz", real),
        ("E", C.c_int), ("W", C.c_int), ("N", C.c_int), ("S", C.c_int),
        ("T", C.c_int), ("B", C.c_int),
    ]


class g_struct(C.Structure):
    _fields_ = [(n, real) for n in
                ("x", "xm", "xa", "y", "ym", "ya", "z", "zm", "za")]


class gradP_struct(C.Structure):
    _fields_ = [(n, real) for n in
                ("x", "xm", "xa", "y", "ym", "ya", "z", "zm", "za")]


def _bc_fields():
    f = []
    for face in "WESNBT":
        f += [("p" + face, C.c_int), ("p" + face + "D", real)]
    for comp in "uvw":
        for face in "WESNBT":
            f += [(comp + face, C.c_int), (comp + face + "Dm", real),
                  (comp + face + "D", real), (comp + face + "Da", real)]
    f += [("ds" + face, real) for face in "WESNBT"]
    return f


class BC(C.Structure):
    _fields_ = _bc_fields()


class cage_struct(C.Structure):
    _fields_ = [(n, C.c_int) for n in (
        "cx", "cy", "cz", "is", "ibs", "ie", "ibe", "in", "js", "jbs", "je",
        "jbe", "jn", "ks", "kbs", "ke", "kbe", "kn")]


NNODES = 26
MAX_NEIGHBORS = 12

_PART_REALS_A = [
    "r", "x", "y", "z", "x0", "y0", "z0", "u", "v", "w", "u0", "v0", "w0",
    "udot", "vdot", "wdot", "udot0", "vdot0", "wdot0",
    "axx", "axy", "axz", "ayx", "ayy", "ayz", "azx", "azy", "azz",
    "ox", "oy", "oz", "ox0", "oy0", "oz0", "oxdot", "oydot", "ozdot",
    "oxdot0", "oydot0", "ozdot0", "Fx", "Fy", "Fz", "Lx", "Ly", "Lz",
    "aFx", "aFy", "aFz", "aLx", "aLy", "aLz", "kFx", "kFy", "kFz",
    "iFx", "iFy", "iFz", "iLx", "iLy", "iLz",
]


class part_struct(C.Structure):
    _fields_ = ([(n, real) for n in _PART_REALS_A] + [
        ("nodes", C.c_int * NNODES), ("rho", real), ("E", real),
        ("sigma", real), ("cage", cage_struct), ("order", C.c_int),
        ("rs", real), ("ncoeff", C.c_int), ("spring_k", real),
        ("spring_x", real), ("spring_y", real), ("spring_z", real),
        ("spring_l", real), ("translating", C.c_int), ("rotating", C.c_int),
        ("bin", C.c_int), ("St", real * MAX_NEIGHBORS),
        ("iSt", C.c_int * MAX_NEIGHBORS), ("e_dry", real),
        ("coeff_fric", real),
    ])


class bb_solve_stats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("outer_steps", C.c_int),
                ("residual", real), ("converged", C.c_int)]


# constants of src/bluebottle.h:155-352
PERIODIC, DIRICHLET, NEUMANN, PRECURSOR = 0, 1, 2, 3
HOMOGENEOUS = 10
WEST, EAST, SOUTH, NORTH, BOTTOM, TOP = range(6)
QUIESCENT, SHEAR, CHANNEL = 0, 1, 2
BB_SOLVER_FP64, BB_SOLVER_MIXED = 0, 1

EXPECTED_SIZES = {
    "grid_info": 192, "dom_struct": 912, "BC": 720, "g_struct": 72,
    "gradP_struct": 72, "part_struct": 928,
}


def check_sizes():
    """Sizes of the mirrors against the reference ABI."""
    got = {
        "grid_info": C.sizeof(grid_info), "dom_struct": C.sizeof(dom_struct),
        "BC": C.sizeof(BC), "g_struct": C.sizeof(g_struct),
        "gradP_struct": C.sizeof(gradP_struct),
        "part_struct": C.sizeof(part_struct),
    }
    return got == EXPECTED_SIZES, got
