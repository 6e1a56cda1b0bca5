"""Diagnostics of the reference's own CPU case (64x32x32 dam break, water
column 0.25 x 0.5 x 0.25 m in the x = 0 corner, water/air 1000/1.2, no-slip
walls, Smagorinsky Cs = 0.1, dt = 0.2 h/(2 sqrt(g H)), 200 steps) computed by
the CPU oracle, whose routines match the flang-built reference bit for bit
(tests/test_oracle.py), every 10 steps: water volume, front position, kinetic
energy, max |u|.  The GPU test compares its own run against this file.

    python tests/golden/gen_config0.py      (about 2 minutes)
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")]
import oracle as O  # noqa: E402
from helpers import dambreak_diagnostics  # noqa: E402

NX, NY, NZ, STEPS, EVERY = 64, 32, 32, 200, 10
WIDTH, HEIGHT = 0.25, 0.25


def main():
    h = 1.0 / NX
    p = O.Params(nx=NX, ny=NY, nz=NZ, dx=h, dy=h, dz=h)
    s = O.State(p)
    s.ls[...] = O.dambreak_phi(p, WIDTH, HEIGHT)
    dt = 0.2 * h / (2 * np.sqrt(9.81 * HEIGHT))
    s.scal["dt"] = dt
    s.run("ls_ini")
    s.run("contorno", 1)
    s.run("contorno", 3)
    rec = []
    for it in range(1, STEPS + 1):
        s.scal["t"] = it * dt
        st = s._struct()  # one step with the step index of the driver loop
        st.t = it * dt
        O.lib().or_step(O.C.byref(st), st.dt)
        s._pull(st)
        if it % EVERY == 0:
            d = dambreak_diagnostics(p, s.u, s.v, s.w, s.ls, s.rho)
            d.update(step=it, vol_ins=s.scal["vol_ins"], cg_iters=s.diag.get("cg_iters"))
            rec.append(d)
            print(d, flush=True)
    out = {"nx": NX, "ny": NY, "nz": NZ, "width": WIDTH, "height": HEIGHT, "dt": dt, "steps": STEPS,
           "vol_ini": s.scal["vol_ini"], "records": rec}
    with open(os.path.join(HERE, "config0_diagnostics.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
