"""Write the golden fixtures (test infrastructure): for each small case,
the seeded initial state is set up on the host, the FP64 oracle runs the
driver's set-up (src/bluebottle.c:445-545) and one time step
(src/bluebottle.c:598-711), and the resulting fields, cell flags, particle
records and Lamb coefficients are stored as compressed numpy arrays.

    python tests/golden/make_fixtures.py

Nothing here touches a GPU.  tests/test_golden.py checks the oracle against
these files (bit-exact) and the HIP path against them (to the parity
tolerances)."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

REAL = ("u", "v", "w", "p", "phi")
INTS = ("phase", "phase_shell", "flag_u", "flag_v", "flag_w")
COEFFS = ("pnm_re", "pnm_im", "phinm_re", "phinm_im", "chinm_re", "chinm_im")
PART = ("x", "y", "z", "u", "v", "w", "ox", "oy", "oz", "Fx", "Fy", "Fz",
        "Lx", "Ly", "Lz")


def cases():
    from bluebottle_amd import configs
    return {
        # 16^3 Taylor-Green box with one moving sphere of radius 3 dx
        "tg16_sphere": configs.small_particle_case(n=16, nparts=1,
                                                   radius_dx=3.0,
                                                   lamb_max_iter=2),
        # 16 x 12 x 16 wall-bounded shear box with two spheres
        "channel16": configs.channel_case(n=(16, 12, 16), nparts=2),
    }


def oracle_step(case):
    """The oracle's state after set-up and one step of `case`."""
    from bluebottle_amd import sim
    from oracle.pyoracle import OracleState
    s = sim.Simulation(case)
    o = OracleState.from_sim(s)
    o.call("prepare")
    lamb_iters = o.call("step")
    return s, o, lamb_iters


def snapshot(o, lamb_iters):
    out = {"lamb_iters": np.array(lamb_iters), "dt": np.array(o.s.dt)}
    for f in REAL + INTS:
        out[f] = np.array(o.array(f))
    if o.nparts:
        for c in COEFFS:
            out[c] = np.array(o.array(c))
        parts = o.parts()
        for k in PART:
            out["part_" + k] = np.array([getattr(p, k) for p in parts])
    return out


def main():
    for name, case in cases().items():
        s, o, it = oracle_step(case)
        data = snapshot(o, it)
        s.close()
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **data)
        print("wrote", path, os.path.getsize(path), "bytes", flush=True)


if __name__ == "__main__":
    main()
