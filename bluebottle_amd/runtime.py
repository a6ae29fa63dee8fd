"""Loading of the in-tree libraries and access to the C-ABI global state.

``libbluebottle_host.so`` (driver-side globals, set-up, step loop) is
loaded first; it pulls in ``libbluebottle_hip.so`` (the HIP entry points),
whose references to the globals resolve against it.  Both live in
``bluebottle_amd/lib`` and are built by ``__graft_entry__.build()``.  There
is no fallback: a missing library raises.
"""
import ctypes as C
import os

import numpy as np

from . import abi

LIB_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")
HIP_LIB = os.path.join(LIB_DIR, "libbluebottle_hip.so")
HOST_LIB = os.path.join(LIB_DIR, "libbluebottle_host.so")

_libs = None

# entry points exported by libbluebottle_hip.so: name -> (restype, argtypes)
_PD = C.POINTER(C.c_double)
ENTRY_POINTS = {
    "cuda_dom_malloc": (None, []), "cuda_dom_push": (None, []),
    "cuda_dom_pull": (None, []), "cuda_dom_free": (None, []),
    "cuda_part_malloc": (None, []), "cuda_part_push": (None, []),
    "cuda_part_pull": (None, []), "cuda_part_free": (None, []),
    "cuda_dom_BC": (None, []), "cuda_dom_BC_star": (None, []),
    "cuda_dom_BC_phi": (None, []), "cuda_dom_BC_p": (None, []),
    "cuda_U_star_2": (None, []),
    "cuda_ustar_helmholtz": (None, [C.c_int]),
    "cuda_vstar_helmholtz": (None, [C.c_int]),
    "cuda_wstar_helmholtz": (None, [C.c_int]),
    "cuda_PP_bicgstab": (None, [C.c_int]),
    "cuda_solvability": (None, []), "cuda_project": (None, []),
    "cuda_update_p": (None, []), "cuda_find_dt": (C.c_double, []),
    "cuda_store_u": (None, []),
    "cuda_compute_forcing": (None, [_PD, _PD, C.c_double, C.c_double,
                                    C.c_double]),
    "cuda_build_cages": (None, []), "cuda_part_BC": (None, []),
    "cuda_part_BC_star": (None, []), "cuda_part_BC_p": (None, [C.c_int]),
    "cuda_part_p_fill": (None, []), "cuda_store_coeffs": (None, []),
    "cuda_move_parts_sub": (None, []), "cuda_move_parts": (None, []),
    "cuda_parts_internal": (None, []),
    "cuda_dom_turb_planes_push": (None, [C.POINTER(C.c_int)]),
    "cuda_dom_turb_planes_pull": (None, [C.POINTER(C.c_int)]),
    "cuda_yank_turb_planes": (None, [C.POINTER(C.c_int), _PD, _PD]),
    "cuda_compute_turb_forcing": (None, []),
    "cuda_compute_energy": (C.c_double, []),
    # source, destination and subdomain are device pointers
    "cuda_colocate_Gfx": (None, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "cuda_colocate_Gfy": (None, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "cuda_colocate_Gfz": (None, [C.c_void_p, C.c_void_p, C.c_void_p]),
    # node tables and node values are device pointers
    "cuda_quad_check_nodes": (None, [C.c_int, C.c_void_p, C.c_void_p, C.c_int]),
    "cuda_quad_interp": (None, [C.c_int, C.c_void_p, C.c_void_p, C.c_int,
                                C.c_void_p, C.c_void_p, C.c_void_p,
                                C.c_void_p]),
    "cuda_Lamb": (None, []), "cuda_lamb_err": (C.c_double, []),
    "bb_set_solver_mode": (None, [C.c_int]),
    "bb_get_solver_mode": (C.c_int, []),
    "bb_get_solve_stats": (None, [C.POINTER(abi.bb_solve_stats)]),
    "bb_device_name": (C.c_char_p, []),
}

HOST_FUNCS = {
    "bb_host_domain": (C.c_int, [C.c_double, C.c_double, C.c_int,
                                 C.c_double, C.c_double, C.c_int,
                                 C.c_double, C.c_double, C.c_int]),
    "bb_host_particles": (C.c_int, [C.c_int]),
    "bb_host_particles_init": (C.c_int, []),
    "bb_host_free": (None, []),
    "bb_host_prepare": (C.c_int, []),
    "bb_host_restart": (C.c_int, [C.c_char_p]),
    "bb_host_step": (C.c_int, []),
    "compute_vel_BC": (None, []),
    "bb_read_flow_config": (C.c_int, [C.c_char_p]),
    "bb_read_part_config": (C.c_int, [C.c_char_p]),
    "bb_out_restart": (C.c_int, [C.c_char_p]),
    "bb_in_restart": (C.c_int, [C.c_char_p]),
    "recorder_lamb": (None, [C.c_char_p, C.c_int]),
    "bb_read_record_config": (C.c_int, [C.c_char_p]),
    "recorder_bicgstab_init": (None, [C.c_char_p]),
    "recorder_lamb_init": (None, [C.c_char_p]),
    "bb_vtk_init": (C.c_int, [C.c_char_p]),
    "bb_vtk_grid": (C.c_int, [C.c_char_p, C.c_int]),
    "bb_vtk_parts": (C.c_int, [C.c_char_p, C.c_int]),
    "bb_vtk_nodes": (C.c_int, [C.c_char_p, C.c_int]),
    "bb_vtk_out": (C.c_int, [C.c_char_p, C.c_int]),
    "bb_vtk_turb_init": (C.c_int, [C.c_char_p]),
    "bb_vtk_turb_grid": (C.c_int, [C.c_char_p, C.c_int]),
    "bb_vtk_turb_out": (C.c_int, [C.c_char_p, C.c_int]),
    "bb_vtk_ghost_init": (C.c_int, [C.c_char_p]),
    "bb_vtk_ghost_grid": (C.c_int, [C.c_char_p, C.c_int]),
    "bb_vtk_ghost_out": (C.c_int, [C.c_char_p, C.c_int]),
    "bb_read_turb_config": (C.c_int, [C.c_char_p]),
    "bb_out_restart_turb": (C.c_int, [C.c_char_p]),
    "bb_in_restart_turb": (C.c_int, [C.c_char_p]),
    "bb_host_init_turb": (C.c_int, [C.c_ulonglong]),
    "bb_host_turb_correct": (C.c_int, []),
    "bb_host_prepare_turb": (C.c_int, []),
    "bb_host_restart_turb": (C.c_int, [C.c_char_p]),
    "bb_host_step_turb": (C.c_int, []),
    "rng_init": (None, [C.c_ulonglong]),
    "rng_uint64": (C.c_ulonglong, []),
    "rng_dbl": (C.c_double, []),
    "gaussrand": (C.c_double, []),
}

EXTRA_HIP = {
    "bb_set_exit_on_nonconvergence": (None, [C.c_int]),
    "bb_synchronize": (None, []),
    "bb_stream": (C.c_void_p, []),
    "bb_prof_enable": (None, [C.c_int]),
    "bb_prof_read": (None, [C.c_int, _PD, C.POINTER(C.c_long)]),
    "bb_prof_reset": (None, []),
    "bb_set_poll_interval": (None, [C.c_int]),
    "bb_set_mixed_params": (None, [C.c_double, C.c_int, C.c_int]),
    "bb_set_inner_sweeps": (C.c_int, [C.c_int]),
    "bb_set_implicit": (None, [C.c_int]),
    "bb_learn_collect": (None, [C.c_int]),
    "bb_learn_pairs": (C.c_int, []),
    "bb_learn_train": (C.c_double, [C.c_int, C.c_double, C.c_ulong]),
    "bb_learn_enable": (None, [C.c_int]),
    "bb_learn_pair_loss": (C.c_double, [C.c_int]),
    "bb_learn_set_fallback": (None, [C.c_int, C.c_double, C.c_int]),
    "bb_learn_stats": (None, [C.POINTER(C.c_long)] * 5
                       + [C.POINTER(C.c_double)]),
    "bb_learn_reset_stats": (None, []),
    "bb_unet_num_params": (C.c_long, []),
    "bb_unet_init": (None, [C.c_ulong]),
    "bb_unet_apply": (None, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "bb_is_implicit": (C.c_int, []),
    "bb_copy_bandwidth": (C.c_double, [C.c_long, C.c_int]),
    "bb_device_alloc": (C.c_void_p, [C.c_size_t]),
    "bb_device_free": (None, [C.c_void_p]),
    "bb_copy_to_device": (None, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "bb_copy_to_host": (None, [C.c_void_p, C.c_void_p, C.c_size_t]),
}

# kernel kinds of the event timer (bb_device.h)
PROF_KINDS = [
    "u_star", "bc", "pp_setup", "pack", "spmv_p", "spmv_s", "update_xr",
    "update_p", "final", "resid64", "correct64", "spmv_p32", "spmv_s32",
    "update_xr32", "update_p32", "project", "update_pressure", "find_dt",
    "cages", "part_bc", "quadrature", "move_parts", "misc", "unet",
    "unet_conv0", "fspmv", "fvec", "sweep_a32", "sweep_b32",
]


def libs_present():
    return os.path.exists(HIP_LIB) and os.path.exists(HOST_LIB)


def load():
    """Load (once) and return (host, hip) ctypes handles."""
    global _libs
    if _libs is not None:
        return _libs
    if not libs_present():
        raise RuntimeError(
            "bluebottle_amd: the HIP library is not built (expected %s); run "
            "__graft_entry__.build()" % HIP_LIB)
    host = C.CDLL(HOST_LIB, mode=C.RTLD_GLOBAL)
    hip = C.CDLL(HIP_LIB, mode=C.RTLD_GLOBAL)
    for name, (res, args) in list(ENTRY_POINTS.items()) + list(EXTRA_HIP.items()):
        fn = getattr(hip, name)
        fn.restype = res
        fn.argtypes = args
    for name, (res, args) in HOST_FUNCS.items():
        fn = getattr(host, name)
        fn.restype = res
        fn.argtypes = args
    _libs = (host, hip)
    return _libs


def gvar(name, ctype):
    """The driver-side global `name` as a ctypes object."""
    host, _ = load()
    return ctype.in_dll(host, name)


def set_scalar(name, value):
    host, _ = load()
    if isinstance(value, int) and not isinstance(value, bool) and name in (
            "pp_max_iter", "lamb_max_iter", "out_plane", "stepnum",
            "init_cond", "nparts", "coeff_stride"):
        C.c_int.in_dll(host, name).value = value
    else:
        C.c_double.in_dll(host, name).value = float(value)


def get_scalar(name, kind="d"):
    host, _ = load()
    t = C.c_int if kind == "i" else C.c_double
    return t.in_dll(host, name).value


def array_view(name, count, dtype=np.float64):
    """numpy view of the host array the global pointer `name` holds."""
    host, _ = load()
    ptype = C.POINTER(C.c_int) if dtype == np.int32 else C.POINTER(C.c_double)
    ptr = ptype.in_dll(host, name)
    if not ptr:
        raise RuntimeError("global %s is NULL" % name)
    return np.ctypeslib.as_array(ptr, shape=(count,))


def dom():
    return gvar("Dom", abi.dom_struct)


def bc():
    return gvar("bc", abi.BC)


def parts(n):
    host, _ = load()
    ptr = C.POINTER(abi.part_struct).in_dll(host, "parts")
    return ptr, [ptr[i] for i in range(n)]


def exported_symbols(path=HIP_LIB):
    """Dynamic symbols a library defines (no GPU needed)."""
    lib = C.CDLL(HOST_LIB, mode=C.RTLD_GLOBAL) if path == HIP_LIB else None
    del lib
    hip = C.CDLL(path, mode=C.RTLD_GLOBAL)
    out = {}
    for name in list(ENTRY_POINTS) + list(EXTRA_HIP):
        try:
            getattr(hip, name)
            out[name] = True
        except AttributeError:
            out[name] = False
    return out
