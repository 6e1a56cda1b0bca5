"""Per-field accuracy of the GPU time step at the benchmark size (not a
pytest module: run it on a GPU machine, it prints one JSON line).

    python tests/accuracy_256.py [--nx 256] [--steps 1 10 100] [--no-oracle]

Dam break 256x128x128 (h = 1/256, column of seed 0 as the benchmark,
dt = 0.2 h / (2 sqrt(g H))).

A. Mixed precision against FP64 on the GPU: the same run from the same
   initial state, once with FP64 fields (FCG-multigrid projection to a
   1e-12 FP64 relative residual) and once in the benchmark's mixed
   precision (FP32 fields, stencils and Krylov sweeps, projection to 1e-8),
   compared after each of --steps steps.
B. FP64 on the GPU against the CPU oracle (the C restatement of the
   reference, bit for bit with the flang-built reference on the golden
   vectors): from the FP64 state after 10 steps, one more step on both with
   the reference's own projection (its CG and stopping rule, Sum r^2 <=
   1e-4/N).

Errors per field: relative L2 (||a - b|| / ||b||) and relative Linf
(max |a - b| / max |b|) for u, v, w (also Linf over the largest speed of
the three: the column spans the tank's width, so v is round-off around
zero) and the pressure (both pressures
shifted to zero mean over the interior cells: the closed tank fixes it only
up to a constant), and for the level set over the interface band |phi| <= 3h
(L2 relative to the band's norm, Linf in units of h).  Interface volume
drift of each run.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("", "oracle", "suli-openmp_amd", "tests"):
    sys.path.insert(0, os.path.join(ROOT, sub))

FIELDS = ("u", "v", "w", "prd1", "ls")


def column_geometry(seed):
    from bench import column_geometry as cg  # the benchmark's seeded column
    return cg(seed)


def errors(a, b, h):
    out = {}
    vmax = max(float(np.abs(b[n]).max()) for n in ("u", "v", "w"))
    for n in ("u", "v", "w"):
        d = a[n] - b[n]
        out[n] = {"rel_l2": float(np.linalg.norm(d) / np.linalg.norm(b[n])),
                  "rel_linf": float(np.abs(d).max() / np.abs(b[n]).max()),
                  "linf_over_max_speed": float(np.abs(d).max() / vmax), "max_abs": float(np.abs(b[n]).max())}
    pa = a["prd1"][1:-1, 1:-1, 1:-1]
    pb = b["prd1"][1:-1, 1:-1, 1:-1]
    pa = pa - pa.mean()
    pb = pb - pb.mean()
    out["p"] = {"rel_l2": float(np.linalg.norm(pa - pb) / np.linalg.norm(pb)),
                "rel_linf": float(np.abs(pa - pb).max() / np.abs(pb).max())}
    band = np.abs(b["ls"]) <= 3 * h
    d = (a["ls"] - b["ls"])[band]
    out["phi_band"] = {"rel_l2": float(np.linalg.norm(d) / np.linalg.norm(b["ls"][band])),
                       "linf_over_h": float(np.abs(d).max() / h), "cells": int(band.sum())}
    return out


def gpu_run(nx, prec, steps, width, height, dt, tol):
    import pysuli as S
    ny = nz = nx // 2
    h = 1.0 / nx
    sv = S.Solver(nx, ny, nz, h, precision=prec, poisson=S.POISSON_FCG, precond=S.PRECOND_GMG, tol_rel=tol)
    snaps, worst = {}, 0.0
    try:
        vol0 = sv.init_dambreak(width, height)
        sv.contorno(1)
        for it in range(1, max(steps) + 1):
            _, ps = sv.step(dt)
            worst = max(worst, max(p["res"] for p in ps))
            if it in steps:
                snaps[it] = {n: sv.download(n) for n in FIELDS}
                snaps[it]["drift"] = (sv.volumes()[1] - vol0) / vol0
        return snaps, worst
    finally:
        sv.close()


def oracle_step_check(nx, width, height, dt, warm_steps=10):
    """FP64 GPU with the reference CG against the oracle, one step from the
    GPU state after warm_steps steps."""
    import oracle as O
    import pysuli as S
    ny = nz = nx // 2
    h = 1.0 / nx
    p = O.Params(nx=nx, ny=ny, nz=nz, dx=h, dy=h, dz=h)
    sv = S.Solver.from_oracle_params(p, precision=S.PREC_FP64, poisson=S.POISSON_REF)
    try:
        vol0 = sv.init_dambreak(width, height)
        sv.contorno(1)
        for _ in range(warm_steps):
            sv.step(dt)
        st = O.State(p)
        for n in ["u", "v", "w", "prd1", "rho", "ls_nu", "ls", "mod_ls", "kurv", "hs", "hsx", "hsy", "hsz",
                  "ddlsdx", "ddlsdy", "ddlsdz", "nut", "u0", "v0", "w0"]:
            st.arr(n)[...] = sv.download(n)
        st.scal["dt"] = dt
        st.scal["vol_ini"] = vol0
        st.run("contorno", 1)  # the friction planes ub, vb, wb, which the GPU keeps implicit
        _, ps = sv.step(dt)
        gpu = {n: sv.download(n) for n in FIELDS}
    finally:
        sv.close()
    t0 = time.time()
    st.run("level_set")
    st.run("contorno", 3)
    print(f"accuracy: oracle level set done ({time.time() - t0:.0f}s)", file=sys.stderr, flush=True)
    its = []
    for tt in (1, 2):
        st.run("stage", tt, it=warm_steps + 1)
        its.append(st.diag["cg_iters"])
        print(f"accuracy: oracle stage {tt} done, {its[-1]} CG iterations ({time.time() - t0:.0f}s)",
              file=sys.stderr, flush=True)
    cpu_s = time.time() - t0
    ref = {n: st.arr(n).copy() for n in FIELDS}
    return {"errors": errors(gpu, ref, h), "cg_iters_gpu": [q["iters"] for q in ps], "cg_iters_oracle": its,
            "oracle_step_s": round(cpu_s, 1), "from_step": warm_steps}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--nx", type=int, default=256)
    ap.add_argument("--steps", type=int, nargs="+", default=[1, 10, 100])
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-oracle", action="store_true")
    args = ap.parse_args()
    import pysuli as S
    nx = args.nx
    h = 1.0 / nx
    width, height = column_geometry(args.seed)
    dt = 0.2 * h / (2.0 * math.sqrt(9.81 * height))
    steps = sorted(set(args.steps))
    out = {"grid": [nx, nx // 2, nx // 2], "dt": dt, "column": [width, height], "steps": steps}
    print(f"accuracy: FP64 run to step {steps[-1]}", file=sys.stderr, flush=True)
    ref, w64 = gpu_run(nx, S.PREC_FP64, steps, width, height, dt, 1e-12)
    print(f"accuracy: mixed run to step {steps[-1]}", file=sys.stderr, flush=True)
    mix, w32 = gpu_run(nx, S.PREC_MIXED, steps, width, height, dt, 1e-8)
    out["mixed_vs_fp64"] = {str(it): dict(errors(mix[it], ref[it], h), drift_mixed=mix[it]["drift"],
                                          drift_fp64=ref[it]["drift"]) for it in steps}
    out["max_relres"] = {"fp64": w64, "mixed": w32}
    if not args.no_oracle:
        print("accuracy: FP64 GPU step against the oracle (CPU, a few minutes)", file=sys.stderr, flush=True)
        out["fp64_vs_oracle"] = oracle_step_check(nx, width, height, dt)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
