"""The learned preconditioner on the GPU (bb_unet.hip, bb_poisson.hip):

* the network's forward pass (BF16 MFMA convolutions, pooling, skip
  connections, the Jacobi term and the solid mask) against the PyTorch
  reference of tests/unet_ref.py with the same BF16 rounding points, on a
  periodic box and on a wall-bounded one (zero padding across the walls);
* the weight gradients of the training loss against PyTorch autograd
  through the same BF16-rounded forward pass (the GPU's backward runs in
  BF16 with FP32 accumulation);
* an untrained network (last layer zero) is the Jacobi preconditioner of
  the fluid system left when the solid unknowns (identity rows) are
  eliminated; Adam steps lower the loss on a fixed pair;
* the learned solver form reaches the same true FP64 residual as the
  others, matches the oracle's phi within the mixed form's tolerance, and
  a garbage network trips the fallback to Jacobi and still converges.
"""
import ctypes as C

import numpy as np
import pytest
import torch

from bluebottle_amd import abi, configs, runtime, sim
from oracle.pyoracle import OracleState

import unet_ref as ur
from helpers import compare_fields

pytestmark = pytest.mark.gpu

TOL_FWD = 1e-2      # BF16 network: one BF16 ulp flips propagate
# BF16 backward against FP32 autograd through the same BF16-rounded forward
# pass, per layer: |g - g_ref| / |g_ref|.  The coarse levels' gradients are
# an order of magnitude smaller and pass more BF16 roundings (measured
# 6-11 % there, 0.5-1 % at level 0); their direction is held to a cosine
# of 0.99 as well.
TOL_GRAD_L0 = 3e-2
TOL_GRAD_COARSE = 0.15
LEVEL0_LAYERS = (0, 1, 9, 10)


def _bind(hip):
    PF = C.POINTER(C.c_float)
    hip.bb_unet_num_params.restype = C.c_long
    hip.bb_unet_set_params.argtypes = [PF, C.c_long]
    hip.bb_unet_get_params.argtypes = [PF, C.c_long]
    hip.bb_unet_get_grads.argtypes = [PF, C.c_long]
    hip.bb_unet_init.argtypes = [C.c_ulong]
    hip.bb_unet_apply.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    hip.bb_unet_loss_grad.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_double]
    hip.bb_unet_loss_grad.restype = C.c_double
    hip.bb_learn_collect.argtypes = [C.c_int]
    hip.bb_learn_pairs.restype = C.c_int
    hip.bb_learn_train.argtypes = [C.c_int, C.c_double, C.c_ulong]
    hip.bb_learn_train.restype = C.c_double
    hip.bb_learn_enable.argtypes = [C.c_int]
    hip.bb_learn_set_fallback.argtypes = [C.c_int, C.c_double, C.c_int]
    L = C.POINTER(C.c_long)
    hip.bb_learn_stats.argtypes = [L, L, L, L, L, C.POINTER(C.c_double)]


class _Dev:
    def __init__(self, hip):
        self.hip, self.ptrs = hip, []

    def put(self, a):
        a = np.ascontiguousarray(a)
        d = self.hip.bb_device_alloc(a.nbytes)
        self.hip.bb_copy_to_device(d, a.ctypes.data, a.nbytes)
        self.ptrs.append(d)
        return d

    def get(self, d, like):
        out = np.empty_like(like)
        self.hip.bb_copy_to_host(out.ctypes.data, d, out.nbytes)
        return out

    def free(self):
        for d in self.ptrs:
            self.hip.bb_device_free(d)
        self.ptrs = []


def _flags(shape, per, seed):
    """Operator flags: every face active except the domain faces of
    non-periodic axes, and a solid ball of cells."""
    nz, ny, nx = shape
    f = np.full(shape, ur.FW | ur.FE | ur.FS | ur.FN | ur.FB | ur.FT, np.uint8)
    if not per[0]:
        f[:, :, 0] &= ~np.uint8(ur.FW)
        f[:, :, -1] &= ~np.uint8(ur.FE)
    if not per[1]:
        f[:, 0, :] &= ~np.uint8(ur.FS)
        f[:, -1, :] &= ~np.uint8(ur.FN)
    if not per[2]:
        f[0, :, :] &= ~np.uint8(ur.FB)
        f[-1, :, :] &= ~np.uint8(ur.FT)
    rng = np.random.default_rng(seed)
    c = rng.uniform(0.3, 0.7, 3) * np.array([nx, ny, nz])
    z, y, x = np.meshgrid(np.arange(nz) + 0.5, np.arange(ny) + 0.5,
                          np.arange(nx) + 0.5, indexing="ij")
    ball = (x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2 < (0.2 * ny) ** 2
    f[ball] |= ur.SOLID
    return f


def _setup(which):
    if which == "periodic":
        case = configs.periodic_case(16, 0.01)
        case.n = (16, 16, 8)
        case.hi = (1.0, 1.0, 0.5)
    else:
        case = configs.channel_case(n=(16, 12, 16), nparts=0)
    s = sim.Simulation(case)
    s.to_device()
    _bind(s.hip)
    d = s.dom
    per = (int(case.bc_p[0][0] == abi.PERIODIC),
           int(case.bc_p[2][0] == abi.PERIODIC),
           int(case.bc_p[4][0] == abi.PERIODIC))
    shape = (d.zn, d.yn, d.xn)
    s.hip.bb_unet_output_scale.restype = C.c_double
    h2 = np.float32(s.hip.bb_unet_output_scale())
    assert abs(h2 / (3.0 / (1 / d.dx ** 2 + 1 / d.dy ** 2 + 1 / d.dz ** 2)
                     / 64) - 1) < 1e-6
    dinv, diag = ur.dinv_tables(d.dx, d.dy, d.dz)
    dd = (np.float32(1 / d.dx ** 2), np.float32(1 / d.dy ** 2),
          np.float32(1 / d.dz ** 2))
    return s, per, shape, float(h2), dinv, diag, dd


def _weights(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    w = []
    for cin, cout in ur.LAYERS:
        w.append(rng.standard_normal(27 * cin * cout) * scale
                 * np.sqrt(2.0 / (27 * cin)))
    return np.concatenate(w).astype(np.float32)


@pytest.mark.parametrize("which", ["periodic", "channel"])
def test_forward_matches_torch(which):
    s, per, shape, h2, dinv, diag, dd = _setup(which)
    print("unet forward:", which, shape, flush=True)
    hip = s.hip
    w = _weights(1)
    hip.bb_unet_set_params(w.ctypes.data_as(C.POINTER(C.c_float)), w.size)
    rng = np.random.default_rng(2)
    r = rng.standard_normal(shape).astype(np.float32)
    fl = _flags(shape, per, 3)
    dev = _Dev(hip)
    try:
        rd, fd = dev.put(r), dev.put(fl)
        od = dev.put(np.zeros(shape, np.float32))
        hip.bb_unet_apply(rd, fd, od)
        got = dev.get(od, r)
    finally:
        dev.free()
    ref = ur.forward(torch.from_numpy(r), torch.from_numpy(w), fl, h2, dinv,
                     per, dd).numpy()
    rt = ur.eliminate(torch.from_numpy(r), fl, dd).numpy()
    solid = (fl & ur.SOLID) != 0
    assert np.array_equal(got[solid], r[solid])
    net = got - np.where(solid, r, dinv[fl & 127] * rt)
    net_ref = ref - np.where(solid, r, dinv[fl & 127] * rt)
    err = np.max(np.abs(net - net_ref)) / np.max(np.abs(net_ref))
    print("unet forward %s: rel err %.2e (scale %.2e)" % (
        which, err, np.max(np.abs(net_ref))), flush=True)
    s.close()
    assert err < TOL_FWD


def test_untrained_network_is_jacobi():
    s, per, shape, h2, dinv, diag, dd = _setup("periodic")
    hip = s.hip
    hip.bb_unet_init(7)
    rng = np.random.default_rng(4)
    r = rng.standard_normal(shape).astype(np.float32)
    fl = _flags(shape, per, 5)
    dev = _Dev(hip)
    try:
        rd, fd = dev.put(r), dev.put(fl)
        od = dev.put(np.zeros(shape, np.float32))
        hip.bb_unet_apply(rd, fd, od)
        got = dev.get(od, r)
    finally:
        dev.free()
    rt = ur.eliminate(torch.from_numpy(r), fl, dd).numpy()
    want = np.where((fl & ur.SOLID) != 0, r, dinv[fl & 127] * rt)
    s.close()
    # D^-1 rt, with rt formed in the kernel's operation order
    assert np.allclose(got, want, rtol=1e-6, atol=0)


@pytest.mark.parametrize("which", ["periodic", "channel"])
def test_gradients_match_torch(which):
    s, per, shape, h2, dinv, diag, dd = _setup(which)
    print("unet gradients:", which, flush=True)
    hip = s.hip
    w = _weights(11, scale=0.7)
    hip.bb_unet_set_params(w.ctypes.data_as(C.POINTER(C.c_float)), w.size)
    rng = np.random.default_rng(12)
    r = rng.standard_normal(shape).astype(np.float32)
    r /= np.linalg.norm(r)
    fl = _flags(shape, per, 13)
    # the target: the network with weights moved by 10 %, so that M(r) - e
    # is well above the BF16 rounding of M(r) (the loss weighs d with A^2,
    # which amplifies rounding noise)
    wp = (w * (1 + 0.1 * rng.standard_normal(w.size))).astype(np.float32)
    e = ur.forward(torch.from_numpy(r), torch.from_numpy(wp), fl, h2, dinv,
                   per, dd).numpy()
    dev = _Dev(hip)
    try:
        rd, ed, fd = dev.put(r), dev.put(e), dev.put(fl)
        loss_gpu = hip.bb_unet_loss_grad(rd, ed, fd, 0.0)
    finally:
        dev.free()
    g = np.zeros(w.size, np.float32)
    hip.bb_unet_get_grads(g.ctypes.data_as(C.POINTER(C.c_float)), g.size)
    wt = torch.from_numpy(w.copy()).requires_grad_(True)
    # the reference forward rounds to BF16 where the kernels do
    lt = ur.loss(torch.from_numpy(r), torch.from_numpy(e), wt, fl, h2, dinv,
                 diag, per, dd, bf16=True)
    lt.backward()
    gref = wt.grad.numpy()
    print("loss gpu %.6e torch %.6e" % (loss_gpu, lt.item()), flush=True)
    assert abs(loss_gpu / lt.item() - 1) < 5e-2
    o = 0
    bad = []
    for li, (cin, cout) in enumerate(ur.LAYERS):
        n = 27 * cin * cout
        a, b = g[o:o + n], gref[o:o + n]
        e_l = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        cos = float(np.dot(a, b) / (np.linalg.norm(a) * np.linalg.norm(b)))
        print("  layer %2d (%2d->%2d): |dW| %.3e rel err %.2e cos %.5f" % (
            li, cin, cout, np.linalg.norm(b), e_l, cos), flush=True)
        tol = TOL_GRAD_L0 if li in LEVEL0_LAYERS else TOL_GRAD_COARSE
        if not (e_l < tol and cos > 0.99):
            bad.append(li)
        o += n
    s.close()
    assert not bad, bad


def test_adam_lowers_the_loss():
    """Teacher and student: the target is the network itself with weights
    moved by 30 %; 40 Adam steps at 1e-3 from the unmoved weights must
    bring the loss below a fifth (PyTorch's Adam on the same problem goes
    from 1.46 to 0.08)."""
    s, per, shape, h2, dinv, diag, dd = _setup("periodic")
    hip = s.hip
    w0 = _weights(21, scale=0.7)
    hip.bb_unet_set_params(w0.ctypes.data_as(C.POINTER(C.c_float)), w0.size)
    rng = np.random.default_rng(22)
    r = rng.standard_normal(shape).astype(np.float32)
    r /= np.linalg.norm(r)
    fl = _flags(shape, per, 23)
    wt = (w0 * (1 + 0.3 * rng.standard_normal(w0.size))).astype(np.float32)
    e = ur.forward(torch.from_numpy(r), torch.from_numpy(wt), fl, h2, dinv,
                   per, dd).numpy()
    dev = _Dev(hip)
    try:
        rd, ed, fd = dev.put(r), dev.put(e), dev.put(fl)
        hip.bb_unet_init(1)      # fresh Adam moments ...
        hip.bb_unet_set_params(w0.ctypes.data_as(C.POINTER(C.c_float)),
                               w0.size)   # ... from the student's weights
        losses = [hip.bb_unet_loss_grad(rd, ed, fd, 1e-3) for _ in range(40)]
    finally:
        dev.free()
    s.close()
    print("losses: %.4e -> %.4e -> %.4e" % (losses[0], losses[9], losses[-1]),
          flush=True)
    assert losses[-1] < 0.2 * losses[0]


def _solve_case():
    return configs.small_particle_case(n=24, nparts=2, radius_dx=3.0,
                                       lamb_max_iter=2, pp_residual=1e-11)


@pytest.mark.parametrize("garbage", [False, True])
def test_learned_form_solves_like_the_oracle(garbage):
    """One pressure solve of the learned form from a cold start at relative
    residual 1e-11: converged on the true FP64 residual, phi within the
    mixed form's 1e-8 of the oracle's; a garbage network (large random
    weights) must trip the fallback and still converge."""
    case = _solve_case()
    s = sim.Simulation(case)
    s.to_device()
    s.prepare()
    _bind(s.hip)
    hip = s.hip
    s.step()
    hip.cuda_U_star_2()
    hip.cuda_dom_BC_star()
    s.pull()
    s.array("phi")[:] = 0.0
    s.push()
    if garbage:
        w = _weights(31, scale=30.0)
        hip.bb_unet_set_params(w.ctypes.data_as(C.POINTER(C.c_float)), w.size)
    else:
        hip.bb_unet_init(32)
    hip.bb_learn_enable(1)
    hip.bb_learn_set_fallback(5, 0.9, 0)   # every inner solve tries the network
    hip.bb_learn_reset_stats()
    s.set_solver(2)
    try:
        o = OracleState.from_sim(s)
        hip.cuda_PP_bicgstab(0)
        o.call("PP_bicgstab")
    finally:
        s.set_solver(0)
        hip.bb_learn_enable(0)
    s.pull()
    st = s.solve_stats()
    a, fb, fi, ji = C.c_long(), C.c_long(), C.c_long(), C.c_long()
    sk, ts = C.c_long(), C.c_double()
    hip.bb_learn_stats(C.byref(a), C.byref(fb), C.byref(fi), C.byref(ji),
                       C.byref(sk), C.byref(ts))
    print("learned solve (garbage=%s): outer %d inner %d resid %.2e; net "
          "applications %d, fallbacks %d, flexible its %d, Jacobi its %d" % (
              garbage, st.outer_steps, st.iterations, st.residual, a.value,
              fb.value, fi.value, ji.value), flush=True)
    assert st.converged and st.residual <= case.pp_residual
    assert a.value > 0
    if garbage:
        assert fb.value > 0
    e = compare_fields(s, o, ("phi",), 1e-8)
    print("phi rel err vs oracle %.2e" % e["phi"], flush=True)
    s.close()
