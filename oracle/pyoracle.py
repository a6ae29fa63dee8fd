"""ctypes wrapper of oracle/liboracle.so (test infrastructure only).

``OracleState.from_sim(sim)`` copies the driver globals of a
``bluebottle_amd.sim.Simulation`` (host arrays, scalars, structs,
particles, coefficients) into the oracle's own state, so one entry point
can be run on both sides from identical inputs.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from bluebottle_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

real = C.c_double
_PR = C.POINTER(C.c_double)
_PI = C.POINTER(C.c_int)


class orc_state(C.Structure):
    _fields_ = [
        ("dom", abi.dom_struct), ("bc", abi.BC), ("g", abi.g_struct),
        ("gradP", abi.gradP_struct),
        ("rho_f", real), ("mu", real), ("nu", real),
        ("dt", real), ("dt0", real), ("ttime", real), ("CFL", real),
        ("pp_residual", real), ("lamb_residual", real), ("lamb_relax", real),
        ("lamb_cut", real),
        ("pp_max_iter", C.c_int), ("lamb_max_iter", C.c_int),
        ("out_plane", C.c_int), ("stepnum", C.c_int), ("init_cond", C.c_int),
        ("p_tDelay", real), ("g_tDelay", real), ("vel_tDelay", real),
        ("Kp", real), ("Ki", real), ("Kd", real), ("pid_int", real),
        ("pid_back", real),
        ("p0", _PR), ("p", _PR), ("phi", _PR),
        ("phase", _PI), ("phase_shell", _PI),
        ("u", _PR), ("u0", _PR), ("v", _PR), ("v0", _PR), ("w", _PR),
        ("w0", _PR), ("f_x", _PR), ("f_y", _PR), ("f_z", _PR),
        ("diff0_u", _PR), ("diff0_v", _PR), ("diff0_w", _PR),
        ("conv0_u", _PR), ("conv0_v", _PR), ("conv0_w", _PR),
        ("diff_u", _PR), ("diff_v", _PR), ("diff_w", _PR),
        ("conv_u", _PR), ("conv_v", _PR), ("conv_w", _PR),
        ("u_star", _PR), ("v_star", _PR), ("w_star", _PR),
        ("flag_u", _PI), ("flag_v", _PI), ("flag_w", _PI),
        ("rhs_p", _PR),
        ("nparts", C.c_int), ("interactionLengthRatio", real),
        ("parts", C.POINTER(abi.part_struct)), ("binDom", abi.dom_struct),
        ("coeff_stride", C.c_int), ("coeff", _PR * 18),
        ("last_iters", C.c_int), ("last_resid", real),
        ("last_converged", C.c_int),
        ("implicit", C.c_int),
        ("plane", C.POINTER(C.c_double) * 15),
        ("turbA", real), ("turbl", real),
    ]


COEFF_NAMES = [
    "pnm_re", "pnm_im", "phinm_re", "phinm_im", "chinm_re", "chinm_im",
    "pnm_re0", "pnm_im0", "phinm_re0", "phinm_im0", "chinm_re0", "chinm_im0",
    "pnm_re00", "pnm_im00", "phinm_re00", "phinm_im00", "chinm_re00",
    "chinm_im00",
]

FIELD_NAMES = [
    "p0", "p", "phi", "phase", "phase_shell", "u", "u0", "v", "v0", "w", "w0",
    "f_x", "f_y", "f_z", "diff0_u", "diff0_v", "diff0_w", "conv0_u",
    "conv0_v", "conv0_w", "diff_u", "diff_v", "diff_w", "conv_u", "conv_v",
    "conv_w", "u_star", "v_star", "w_star", "flag_u", "flag_v", "flag_w",
]

SCALARS = ["rho_f", "mu", "nu", "dt", "dt0", "ttime", "CFL", "pp_residual",
           "lamb_residual", "lamb_relax", "lamb_cut", "p_tDelay", "g_tDelay",
           "vel_tDelay", "Kp", "Ki", "Kd", "pid_int", "pid_back",
           "interactionLengthRatio", "turbA", "turbl"]
INT_SCALARS = ["pp_max_iter", "lamb_max_iter", "out_plane", "stepnum",
               "init_cond"]

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        lib = C.CDLL(LIB)
        lib.orc_create.restype = C.POINTER(orc_state)
        lib.orc_create.argtypes = [C.POINTER(abi.dom_struct), C.c_int, C.c_int]
        lib.orc_free.argtypes = [C.POINTER(orc_state)]
        lib.orc_array.restype = C.c_void_p
        lib.orc_array.argtypes = [C.POINTER(orc_state), C.c_char_p,
                                  C.POINTER(C.c_size_t), C.POINTER(C.c_int)]
        lib.orc_sizeof_state.restype = C.c_size_t
        for fn in ("orc_find_dt", "orc_lamb_err", "orc_Nnm", "orc_Pnm",
                   "orc_compute_energy"):
            getattr(lib, fn).restype = C.c_double
        lib.orc_Nnm.argtypes = [C.c_int, C.c_int]
        lib.orc_Pnm.argtypes = [C.c_int, C.c_int, C.c_double]
        lib.orc_step.restype = C.c_int
        lib.orc_step_turb.restype = C.c_int
        lib.orc_colocate.argtypes = [C.POINTER(orc_state), C.c_int, _PR]
        lib.orc_rng_seeded.restype = C.c_ulonglong
        lib.orc_rng_seeded.argtypes = [C.c_ulonglong, C.c_ulonglong]
        lib.orc_rng_dbl.restype = C.c_double
        lib.orc_rng_dbl.argtypes = [C.POINTER(C.c_ulonglong)]
        lib.orc_init_turb.argtypes = [C.POINTER(orc_state),
                                      C.POINTER(C.c_ulonglong)]
        lib.orc_bicgstab.restype = C.c_int
        lib.orc_bicgstab.argtypes = [C.POINTER(orc_state), _PR, _PR, C.c_int,
                                     C.c_double, _PR, _PI]
        lib.orc_pp_matvec.argtypes = [C.POINTER(orc_state), _PR, _PR]
        lib.orc_pp_coeffs.argtypes = [C.POINTER(orc_state), C.c_int, _PR]
        lib.orc_star_helmholtz.argtypes = [C.POINTER(orc_state), C.c_int]
        lib.orc_helmholtz_coeffs.argtypes = [C.POINTER(orc_state), C.c_int,
                                             _PR]
        lib.orc_helmholtz_matvec.argtypes = [C.POINTER(orc_state), C.c_int,
                                             _PR, _PR, _PR]
        lib.orc_pcg.restype = C.c_int
        lib.orc_pcg.argtypes = [C.POINTER(orc_state), C.c_int, _PR, _PR, _PR,
                                C.c_int, C.c_double, _PR, _PI]
        lib.orc_lebedev_nodes.argtypes = [_PR, _PR, _PR]
        lib.orc_xyz2rtp.argtypes = [C.c_double] * 3 + [_PR] * 3
        lib.orc_lamb_vel.argtypes = [C.POINTER(orc_state), C.c_int,
                                     C.c_double, C.c_double, C.c_double,
                                     C.c_double, C.c_int, _PR, _PR, _PR]
        _lib = lib
    return _lib


class OracleState:
    def __init__(self, dom, nparts, coeff_stride):
        self.lib = load()
        self.nparts = nparts
        self.coeff_stride = coeff_stride
        self.ptr = self.lib.orc_create(C.byref(dom), nparts, coeff_stride)
        self.s = self.ptr.contents

    def __del__(self):
        try:
            self.lib.orc_free(self.ptr)
        except Exception:
            pass

    def array(self, name):
        n = C.c_size_t()
        is_int = C.c_int()
        addr = self.lib.orc_array(self.ptr, name.encode(), C.byref(n),
                                  C.byref(is_int))
        if not addr:
            raise KeyError(name)
        t = C.c_int if is_int.value else C.c_double
        return np.ctypeslib.as_array((t * n.value).from_address(addr))

    def parts(self):
        return [self.s.parts[i] for i in range(self.nparts)]

    @classmethod
    def from_sim(cls, sim):
        """Copy the (pulled) host state of a Simulation."""
        from bluebottle_amd import runtime
        host = sim.host
        dom = runtime.dom()
        n = len(sim.case.particles)
        cs = runtime.get_scalar("coeff_stride", "i")
        o = cls(dom, n, cs)
        s = o.s
        s.bc = runtime.bc()
        s.g = runtime.gvar("g", abi.g_struct)
        s.gradP = runtime.gvar("gradP", abi.gradP_struct)
        for k in SCALARS:
            setattr(s, k, runtime.get_scalar(k))
        for k in INT_SCALARS:
            setattr(s, k, runtime.get_scalar(k, "i"))
        s.implicit = 1 if sim.case.implicit else 0
        for name in FIELD_NAMES:
            np.copyto(o.array(name), sim.array(name))
        for name in getattr(sim, "PLANES", {}):
            if C.c_void_p.in_dll(host, name).value:
                np.copyto(o.array(name), sim.array(name))
        if n:
            src = C.POINTER(abi.part_struct).in_dll(host, "parts")
            C.memmove(s.parts, src, n * C.sizeof(abi.part_struct))
            s.binDom = runtime.gvar("binDom", abi.dom_struct)
            for name in COEFF_NAMES:
                np.copyto(o.array(name), sim.array(name))
        return o

    def call(self, name):
        fn = getattr(self.lib, "orc_" + name)
        fn.argtypes = [C.POINTER(orc_state)]
        return fn(self.ptr)
