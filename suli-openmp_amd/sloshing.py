"""Large sloshing tank (the survey's config 4) on one GPU.

Tank 2.0 x 1.0 x 1.0 m on an nx x nx/2 x nx/2 grid (default 512 x 256 x 256,
33.5 M cells), water at rest to a depth of 0.4 m (level set = 0.4 - z,
positive in water), air above; horizontal body force
gx(t) = 0.05 g sin(2 pi t / 1.2 s + phase), the phase drawn from the seed
with splitmix64; density ratio 1000 (rho_water 1000, rho_air from the
ratio); no-slip walls; Smagorinsky Cs = 0.1; mixed precision; dt = 0.2 h /
sqrt(g depth).  Per window of --window steps it reports iterations, GPU ms
per solve and fallbacks of the run's own solves, the interface volume drift
and max |u|; at the first step of each window the first-stage projection
is solved from the current pressure with plain Jacobi-FCG, the default
learned preconditioner (trained on dam breaks: unseen geometry, forcing and
density ratio) and multigrid FCG, all to the 1e-8 FP64 residual.

    python3 sloshing.py --nx 512 --steps 2000 --ratio 100 500 1000
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


def _compare(s, dt, cases):
    """The next first-stage projection solved with each preconditioner from
    the current pressure (the extrapolated guess off for these solves; the
    run's guess history restarts after them)."""
    s.level_set(dt)
    s.contorno(3)
    s.visco()
    s.convdiff()
    s.tempo(1, 0.5 * dt)
    s.contorno(2)
    p0 = s.download("prd1")
    s.set_guess(0)
    out = {}
    for name, c, meta in cases:  # the run's own preconditioner last: its pressure continues the stage
        if meta is not None:
            S.use_learned(s, meta)
        s.upload("prd1", p0)
        s.set_solver(S.POISSON_FCG, c)
        st = s.graddin(0.5 * dt)
        out[name] = {"ms": round(st["ms"], 2), "iters": st["iters"], "outer": st["outer"],
                     "fallbacks": st["fallbacks"], "converged": st["converged"], "relres": st["res"]}
    s.set_guess(1)
    s.posdin(0.5 * dt)
    s.contorno(1)
    st2 = s.stage(2, dt)  # finish the step
    return out, st2


def run(nx, steps, ratio, seed=0, depth=0.4, precond="gmg", window=200, compare=True):
    ny = nz = nx // 2
    h = 2.0 / nx
    rho_w = 1000.0
    pcs = {"jacobi": S.PRECOND_JACOBI, "gmg": S.PRECOND_GMG, "learned": S.PRECOND_LEARNED}
    s = S.Solver(nx, ny, nz, h, precision=S.PREC_MIXED, poisson=S.POISSON_FCG, precond=pcs[precond], tol_rel=1e-8,
                 rho_f1=rho_w / ratio, rho_f2=rho_w)
    out = {"grid": [nx, ny, nz], "h": h, "ratio": ratio, "seed": seed, "precond": precond, "window": window}
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
        models = [m for m in S.shipped_learned() if nx % (1 << (m["entry"] + m["levels"])) == 0]
        default = next((m for m in models if m["default"]), None)
        if precond == "learned":
            S.use_learned(s, default)
        cases = [("jacobi", S.PRECOND_JACOBI, None)]
        if default is not None:
            cases.append(("learned", S.PRECOND_LEARNED, default))
        cases.append(("gmg", S.PRECOND_GMG, None))
        if precond != "gmg":  # the run's own preconditioner solves last
            cases = [c for c in cases if c[0] != precond] + [c for c in cases if c[0] == precond]
        windows, wall, nwall = [], 0.0, 0
        t = 0.0
        wstats = []
        for it in range(1, steps + 1):
            t += dt
            boundary = compare and (it % window == 1 or window == 1)
            if boundary:
                s.sync()
                s.set_time(t)
                cmp_, st2 = _compare(s, dt, cases)
                s.set_solver(S.POISSON_FCG, pcs[precond])
                if precond == "learned":
                    S.use_learned(s, default)
                ps = [dict(cmp_[precond], ms=cmp_[precond]["ms"], res=cmp_[precond]["relres"]), st2]
            else:
                t0 = time.time()
                _, ps = s.step(dt)
                wall += time.time() - t0
                nwall += 1
            wstats += ps
            if it % window == 0 or it == steps:
                u = s.download("u")
                rec = {"steps": f"{it - len(wstats) // 2 + 1}-{it}", "t": round(t, 4),
                       "iters_per_solve": round(sum(p["iters"] for p in wstats) / len(wstats), 2),
                       "ms_per_solve": round(sum(p["ms"] for p in wstats) / len(wstats), 2),
                       "fallbacks": sum(p["fallbacks"] for p in wstats),
                       "volume_drift": (s.volumes()[1] - vol0) / vol0, "max_u": float(np.abs(u).max())}
                if compare:
                    rec["compare_at_step"] = it - len(wstats) // 2 + 1
                    rec["compare"] = cmp_
                windows.append(rec)
                print(json.dumps(rec), file=sys.stderr, flush=True)
                wstats = []
        s.sync()
        out.update(steps=steps, dt=dt, t_end=t, mcell_steps_per_s=nx * ny * nz * nwall / max(wall, 1e-9) / 1e6,
                   volume_drift=(s.volumes()[1] - vol0) / vol0, windows=windows)
        return out
    finally:
        s.close()


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--nx", type=int, default=512)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--ratio", type=float, nargs="+", default=[1000.0])
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--precond", default="gmg", choices=["jacobi", "gmg", "learned"])
    ap.add_argument("--window", type=int, default=200)
    ap.add_argument("--no-compare", action="store_true")
    args = ap.parse_args()
    for r in args.ratio:
        print(json.dumps(run(args.nx, args.steps, r, args.seed, precond=args.precond, window=args.window,
                             compare=not args.no_compare)), flush=True)


if __name__ == "__main__":
    main()
