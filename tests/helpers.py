"""Shared helpers for the parity tests: run one entry point (or a whole
step) on the GPU through the C ABI and on the oracle from the same state,
and compare."""
import numpy as np

from bluebottle_amd import sim
from oracle.pyoracle import OracleState

REAL_FIELDS = ("u", "v", "w", "p", "phi", "u_star", "v_star", "w_star",
               "conv_u", "conv_v", "conv_w", "diff_u", "diff_v", "diff_w",
               "p0", "u0", "v0", "w0")
INT_FIELDS = ("phase", "phase_shell", "flag_u", "flag_v", "flag_w")


def start(case, prepare=True):
    """Simulation on the device (prepared) plus an oracle copy of its
    state."""
    s = sim.Simulation(case)
    s.to_device()
    if prepare:
        s.prepare()
    s.pull()
    o = OracleState.from_sim(s)
    return s, o


def rel_err(a, b):
    scale = max(float(np.max(np.abs(b))), 1e-300)
    return float(np.max(np.abs(a - b))) / scale


def compare_fields(s, o, names, tol):
    out = {}
    for n in names:
        e = rel_err(s.array(n), o.array(n))
        out[n] = e
        assert e <= tol, "%s: relative error %.3e above %.1e" % (n, e, tol)
    return out


def compare_ints(s, o, names=INT_FIELDS):
    for n in names:
        a, b = s.array(n), o.array(n)
        bad = np.flatnonzero(a != b)
        assert bad.size == 0, "%s differs at %d entries (first %s: %s vs %s)" % (
            n, bad.size, bad[:5], a[bad[:5]], b[bad[:5]])


PART_FIELDS = ("x", "y", "z", "u", "v", "w", "udot", "vdot", "wdot", "ox",
               "oy", "oz", "oxdot", "oydot", "ozdot", "Fx", "Fy", "Fz", "Lx",
               "Ly", "Lz", "iFx", "iFy", "iFz", "iLx", "iLy", "iLz", "axx",
               "axy", "axz", "ayx", "ayy", "ayz", "azx", "azy", "azz")


COEFF_SET = ("pnm_re", "pnm_im", "phinm_re", "phinm_im", "chinm_re",
             "chinm_im")


def compare_parts(s, o, tol):
    """Particle records: each real field relative to that field's largest
    magnitude over all particles; node flags and cages exactly."""
    ps, po = s.parts(), o.parts()
    worst = 0.0
    for f in PART_FIELDS:
        va = np.array([getattr(a, f) for a in ps])
        vb = np.array([getattr(b, f) for b in po])
        scale = max(float(np.max(np.abs(vb))) if vb.size else 0.0, 1e-300)
        e = float(np.max(np.abs(va - vb))) / scale if vb.size else 0.0
        worst = max(worst, e)
        assert e <= tol, "particle field %s: %r vs %r (rel %.3e)" % (
            f, va, vb, e)
    for a, b in zip(ps, po):
        assert list(a.nodes) == list(b.nodes)
        for f in ("cx", "cy", "cz", "is", "ibs", "ie", "ibe", "in", "js", "jbs",
                  "je", "jbe", "jn", "ks", "kbs", "ke", "kbe", "kn"):
            assert getattr(a.cage, f) == getattr(b.cage, f), "cage." + f
    return worst
