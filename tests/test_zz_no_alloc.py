"""bb_lamb_err_noalloc, the variant of cuda_lamb_err that allocates nothing
(its per-particle scratch is a static device array), called through the C
ABI against cuda_lamb_err on the same coefficients: the same error value
and the same previous-coefficient arrays afterwards, bit for bit, on two
calls in a row, the second with fewer particles than the first."""
import ctypes as C

import numpy as np
import pytest

from bluebottle_amd import configs, runtime, sim

from helpers import COEFF_SET

pytestmark = pytest.mark.gpu

PREV = tuple(c + "0" for c in COEFF_SET)


def _device_ptrs(host, names):
    out = (C.c_void_p * 6)()
    for i, n in enumerate(names):
        p = C.POINTER(C.POINTER(C.c_double)).in_dll(host, "_" + n)[0]
        out[i] = C.cast(p, C.c_void_p).value
    return out


def _one_case(nparts, seed):
    case = configs.small_particle_case(n=24, nparts=nparts, radius_dx=2.5,
                                       seed=seed, lamb_max_iter=2)
    s = sim.Simulation(case)
    s.to_device()
    s.prepare()
    s.step()
    hip = s.hip
    hip.bb_lamb_err_noalloc.restype = C.c_double
    hip.bb_lamb_err_noalloc.argtypes = [C.c_int, C.c_int, C.c_double,
                                        C.c_void_p * 6, C.c_void_p * 6]
    # fresh coefficients, so that the previous ones differ from them
    hip.cuda_Lamb()
    s.pull()
    before = {k: np.array(s.array(k)) for k in PREV}
    want = hip.cuda_lamb_err()
    s.pull()
    want_prev = {k: np.array(s.array(k)) for k in PREV}
    # back to the same previous coefficients on the device
    for k, v in before.items():
        s.array(k)[:] = v
    s.push()
    stride = runtime.get_scalar("coeff_stride", "i")
    cut = runtime.get_scalar("lamb_cut")
    got = hip.bb_lamb_err_noalloc(nparts, stride, cut,
                                  _device_ptrs(s.host, COEFF_SET),
                                  _device_ptrs(s.host, PREV))
    s.pull()
    got_prev = {k: np.array(s.array(k)) for k in PREV}
    s.close()
    return want, got, want_prev, got_prev


def test_noalloc_lamb_err_matches_entry_point():
    for nparts, seed in ((6, 21), (2, 22)):
        print("no-alloc lamb_err: %d particles" % nparts, flush=True)
        want, got, want_prev, got_prev = _one_case(nparts, seed)
        print("  cuda_lamb_err %.17g, bb_lamb_err_noalloc %.17g" % (want, got),
              flush=True)
        assert want > 0.0
        assert got == want
        for k in PREV:
            assert np.array_equal(got_prev[k], want_prev[k]), k


def test_noalloc_lamb_err_rejects_oversize():
    _, hip = runtime.load()
    hip.bb_lamb_err_noalloc.restype = C.c_double
    hip.bb_lamb_err_noalloc.argtypes = [C.c_int, C.c_int, C.c_double,
                                        C.c_void_p * 6, C.c_void_p * 6]
    null = (C.c_void_p * 6)()
    assert hip.bb_lamb_err_noalloc(32769, 15, 1e-4, null, null) == -1.0
    assert hip.bb_lamb_err_noalloc(10, 22, 1e-4, null, null) == -1.0
