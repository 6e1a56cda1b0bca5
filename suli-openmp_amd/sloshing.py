"""Large sloshing tank (the survey's config 4) on one GPU.

Tank 2.0 x 1.0 x 1.0 m on an nx x nx/2 x nx/2 grid (default 512 x 256 x 256,
33.5 M cells), water at rest to a depth of 0.4 m (level set = 0.4 - z,
positive in water), air above; horizontal body force
gx(t) = 0.05 g sin(2 pi t / 1.2 s + phase), the phase drawn from the seed
with splitmix64; density ratio 1000 (rho_water 1000, rho_air from the
ratio); no-slip walls; Smagorinsky Cs = 0.1; mixed precision; dt = 0.2 h /
sqrt(g depth).  After the steps it solves the projection of the next stage
with each preconditioner (the learned ones from suli-openmp_amd/weights,
trained on dam breaks: unseen geometry and forcing) and reports ms per
solve, iterations and fallbacks.

    python3 sloshing.py --nx 512 --steps 20 --ratio 1000
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import pysuli as S  # noqa: E402
from train_unet import splitmix64  # noqa: E402


def run(nx, steps, ratio, seed=0, depth=0.4, precond="gmg"):
    ny = nz = nx // 2
    h = 2.0 / nx
    rho_w = 1000.0
    pc = {"jacobi": S.PRECOND_JACOBI, "gmg": S.PRECOND_GMG}[precond]
    s = S.Solver(nx, ny, nz, h, precision=S.PREC_MIXED, poisson=S.POISSON_FCG, precond=pc, tol_rel=1e-8,
                 rho_f1=rho_w / ratio, rho_f2=rho_w)
    out = {"grid": [nx, ny, nz], "h": h, "ratio": ratio, "seed": seed}
    try:
        z = (np.arange(1, nz + 1) - 0.5) * h
        phi = np.broadcast_to(depth - z, (nx, ny, nz)).copy(order="F")
        s.upload("ls", phi)
        vol0 = s.level_set_ini()
        s.contorno(1)
        s.contorno(3)
        phase = 2 * math.pi * next(splitmix64(seed))
        s.set_forcing(0.05 * 9.81, 1.2, phase)
        dt = 0.2 * h / math.sqrt(9.81 * depth)
        t0 = time.time()
        its = []
        for _ in range(steps):
            _, ps = s.step(dt)
            its += [p["iters"] for p in ps]
        s.sync()
        wall = time.time() - t0
        out.update(steps=steps, dt=dt, mcell_steps_per_s=nx * ny * nz * steps / wall / 1e6,
                   iters_per_solve=sum(its) / len(its), volume_drift=(s.volumes()[1] - vol0) / vol0)
        # the next projection with each preconditioner, warm-started from the current pressure
        s.level_set(dt)
        s.contorno(3)
        s.visco()
        s.convdiff()
        s.tempo(1, 0.5 * dt)
        s.contorno(2)
        p0 = s.download("prd1")
        cases = [("jacobi", S.PRECOND_JACOBI, None), ("gmg", S.PRECOND_GMG, None)]
        for m in S.shipped_learned():
            if nx % (1 << (m["entry"] + m["levels"])) == 0:
                cases.append((m["name"], S.PRECOND_LEARNED, m))
        out["poisson"] = {}
        for name, c, meta in cases:
            if meta is not None:
                S.use_learned(s, meta)
            s.upload("prd1", p0)
            s.set_solver(S.POISSON_FCG, c)
            st = s.graddin(0.5 * dt)
            out["poisson"][name] = {"ms": round(st["ms"], 2), "iters": st["iters"], "outer": st["outer"],
                                    "fallbacks": st["fallbacks"], "converged": st["converged"], "relres": st["res"]}
        return out
    finally:
        s.close()


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--nx", type=int, default=512)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--ratio", type=float, nargs="+", default=[1000.0])
    ap.add_argument("--seed", type=int, default=0)
    args = ap.parse_args()
    for r in args.ratio:
        print(json.dumps(run(args.nx, args.steps, r, args.seed)), flush=True)


if __name__ == "__main__":
    main()
