"""PyTorch (CPU, FP32) reference of the learned preconditioner's network,
for the GPU tests of bb_unet.hip.

`forward(r, W, flags, ...)` repeats the network with BF16 rounding at the
points where the HIP kernels store BF16 (convolution inputs and weights,
stored activations); `forward(..., bf16=False)` is the same function in
plain FP32, differentiable for the gradient check.  Weights are the
library's flat FP32 vector, layer by layer, each [cout][cin][27] with tap
(dz*3 + dy)*3 + dx, i.e. torch's conv3d weight layout.
"""
import numpy as np
import torch
import torch.nn.functional as F

LAYERS = [(1, 8), (8, 8), (8, 16), (16, 16), (16, 32), (32, 32), (32, 16),
          (16, 16), (16, 8), (8, 8), (8, 1)]
SLOPE = 0.1
FW, FE, FS, FN, FB, FT, SOLID = 1, 2, 4, 8, 16, 32, 64


def n_params():
    return sum(27 * a * b for a, b in LAYERS)


def split(w):
    out, o = [], 0
    for cin, cout in LAYERS:
        n = 27 * cin * cout
        out.append(w[o:o + n].reshape(cout, cin, 3, 3, 3))
        o += n
    return out


def _bf(x, on):
    return x.to(torch.bfloat16).to(torch.float32) if on else x


def _conv(x, W, per, bf):
    """x (C, nz, ny, nx); per = (perx, pery, perz)."""
    x = x[None]
    # pad z, y, x: circular on periodic axes, zeros elsewhere
    for axis, p in ((4, per[0]), (3, per[1]), (2, per[2])):
        if p:
            lo = x.narrow(axis, x.shape[axis] - 1, 1)
            hi = x.narrow(axis, 0, 1)
        else:
            shape = list(x.shape)
            shape[axis] = 1
            lo = hi = torch.zeros(shape, dtype=x.dtype)
        x = torch.cat([lo, x, hi], dim=axis)
    return F.conv3d(x, _bf(W, bf))[0]


def _act(x):
    return torch.where(x > 0, x, SLOPE * x)


def _pool(x):
    return F.avg_pool3d(x[None], 2)[0]


def _up(c):
    return c.repeat_interleave(2, 1).repeat_interleave(2, 2).repeat_interleave(2, 3)


def dinv_tables(dx, dy, dz):
    ddx, ddy, ddz = 1 / dx ** 2, 1 / dy ** 2, 1 / dz ** 2
    d = np.ones(128, np.float32)
    g = np.ones(128, np.float32)
    for f in range(128):
        if f & SOLID:
            continue
        fx = bool(f & FW) + bool(f & FE)
        fy = bool(f & FS) + bool(f & FN)
        fz = bool(f & FB) + bool(f & FT)
        dg = -(fx * ddx + fy * ddy + fz * ddz)
        g[f] = np.float32(dg)
        d[f] = np.float32(1.0 / dg) if dg != 0 else 1.0
    return d, g


def eliminate(r, flags, dd):
    """rt = r - A_fs r_s on fluid cells, 0 on solid cells (the solid rows
    of A are the identity)."""
    fl = torch.from_numpy(flags.astype(np.int64))
    solid = (fl & SOLID) != 0
    rs = torch.where(solid, r, torch.zeros_like(r))
    out = r.clone()
    for bit, axis, shift, d in ((FW, 2, 1, dd[0]), (FE, 2, -1, dd[0]),
                                (FS, 1, 1, dd[1]), (FN, 1, -1, dd[1]),
                                (FB, 0, 1, dd[2]), (FT, 0, -1, dd[2])):
        nb = torch.roll(rs, shifts=shift, dims=axis)
        out = out - torch.where((fl & bit) != 0, d * nb, torch.zeros_like(nb))
    return torch.where(solid, torch.zeros_like(out), out)


def forward(r, w, flags, h2, dinv, per, dd, bf16=True):
    """r (nz, ny, nx) float32 tensor; w flat tensor; flags uint8 array of the
    same shape; dd = (1/dx^2, 1/dy^2, 1/dz^2); returns M(r) (nz, ny, nx)."""
    Ws = split(w)
    b = bf16
    rt = eliminate(r, flags, dd)
    x = _bf(rt[None], b)
    a0 = _bf(_act(_conv(x, Ws[0], per, b)), b)
    x0 = _bf(_act(_conv(a0, Ws[1], per, b)), b)
    q0 = _bf(_pool(x0), b)
    a1 = _bf(_act(_conv(q0, Ws[2], per, b)), b)
    x1 = _bf(_act(_conv(a1, Ws[3], per, b)), b)
    q1 = _bf(_pool(x1), b)
    a2 = _bf(_act(_conv(q1, Ws[4], per, b)), b)
    x2 = _bf(_act(_conv(a2, Ws[5], per, b)), b)
    c2 = _bf(_conv(x2, Ws[6], per, b), b)
    u1 = _bf(_act(_up(c2) + x1), b)
    y1 = _bf(_act(_conv(u1, Ws[7], per, b)), b)
    c1 = _bf(_conv(y1, Ws[8], per, b), b)
    u0 = _bf(_act(_up(c1) + x0), b)
    y0 = _bf(_act(_conv(u0, Ws[9], per, b)), b)
    o = _conv(y0, Ws[10], per, b)[0]
    fl = torch.from_numpy(flags.astype(np.int64))
    solid = (fl & SOLID) != 0
    d = torch.from_numpy(dinv)[fl & 127]
    return torch.where(solid, r, d * rt + h2 * o)


def a_ff(y, flags, ddx, ddy, ddz, diag):
    """The Poisson operator between fluid cells (periodic index wraps; the
    face bits decide which neighbours couple); 0 on solid rows."""
    fl = torch.from_numpy(flags.astype(np.int64))
    fluid = (fl & SOLID) == 0
    ym = torch.where(fluid, y, torch.zeros_like(y))
    s = torch.from_numpy(diag)[fl & 127] * ym
    for bit, axis, shift, dd in ((FW, 2, 1, ddx), (FE, 2, -1, ddx),
                                 (FS, 1, 1, ddy), (FN, 1, -1, ddy),
                                 (FB, 0, 1, ddz), (FT, 0, -1, ddz)):
        nb = torch.roll(ym, shifts=shift, dims=axis)
        s = s + torch.where((fl & bit) != 0, dd * nb, torch.zeros_like(nb))
    return torch.where(fluid, s, torch.zeros_like(s))


def loss(r, e, w, flags, h2, dinv, diag, per, dd, bf16=False):
    out = forward(r, w, flags, h2, dinv, per, dd, bf16=bf16)
    fl = torch.from_numpy(flags.astype(np.int64))
    fluid = (fl & SOLID) == 0
    d = torch.where(fluid, out - e, torch.zeros_like(out))
    q = a_ff(d, flags, dd[0], dd[1], dd[2], diag)
    rt = eliminate(r, flags, dd)
    return (q * q).sum() / (rt * rt).sum()
