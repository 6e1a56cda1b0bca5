#!/usr/bin/env python3
"""Benchmark of the SuLi-IC time step on one MI355X.

Workload: 3-D dam break, water/air (rho 1000/1.2 kg/m^3, mu 1e-3/1.8e-5
Pa s, g = 9.81 in -z), 1.0 x 0.5 x 0.5 m tank on a 256x128x128 staggered
grid (h = 1/256), no-slip walls, Smagorinsky Cs = 0.1, WENO5 level set.
Water column width/height drawn from the seed (splitmix64): width in
[0.2, 0.4] m, height in [0.3, 0.45] m.  Fixed dt = 0.2 h / (2 sqrt(g H)).
Fields FP32 (stencils and Krylov sweeps), FP64 pressure, residuals, dot
products and volume integrals; Poisson solve = FCG with FP64 iterative
refinement to a 1e-8 FP64 relative residual, geometric-multigrid
preconditioner.  Synthetic data: the tank is initialised in place.

A "step" is one full time step of 1_principal.f90:74-95 (level set, two RK2
stages each with visco, convdiff+tempo, contorno, projection, posdin).
After --warmup untimed steps, --steps steps are timed between two stream
synchronisations.  value = nx*ny*nz*steps/seconds/1e6 (Mcell-steps/s).

The last warm-up step is timed kernel by kernel (HIP events around every
launch) to pick the dominant fine-grid kernel; during the timed steps only
every 7th launch of that kernel carries events, attached to its dispatches
(hipExtLaunchKernelGGL; an event pair costs a few microseconds of dispatch
time), so the other launches run undisturbed.  The JSON
line carries:
- roofline: algorithmic bytes per launch of the dominant kernel / its average
  duration over the timed steps, against 8 TB/s; traffic = HBM bytes per
  launch from rocprof FETCH_SIZE + WRITE_SIZE (profiles/hbm_traffic_*.json,
  made by --pmc-summary from counter runs of this script, scaled by a 256 MiB
  device copy measured the same way);
- poisson.comparison: ms per solve of one projection state with the plain
  (Jacobi) FCG, the multigrid FCG and each shipped learned preconditioner
  (suli-openmp_amd/weights; the default one as "learned"), all to a 1e-8
  FP64 relative residual;
- cpu_baseline: the flang-built reference harness (oracle/_ref) timed on one
  host core on a bounded sample of the same workload (one level_set plus
  one RK stage from the GPU's post-warmup state; per-step time = level_set
  + 2 x stage).
"""
import argparse
import json
import math
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "suli-openmp_amd"))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s

# Algorithmic bytes per cell of each timed kernel (each distinct array read
# once and written once per cell; FP32 fields = 4 B, FP64 = 8 B).  Kernels
# whose grid is a coarse multigrid level use that level's cells.
BYTES_PER_CELL = {
    "ls_advect_rk3": 6 * 4,        # phi, phi_n, u, v, w -> phi
    "reinit_sign": 2 * 4,          # phi -> S
    "reinit_rk3": 4 * 4,           # phi, phi_it, S -> phi
    "heaviside_normals": 11 * 4,   # phi -> H, Hx, Hy, Hz, rho, mu, |grad phi|, n (3)
    "curvature": 7 * 4,            # n (3) -> d (3), kappa
    "volume_correction": 3 * 4,
    "visco": 5 * 4,                # u, v, w, rho -> nu_t
    "convdiff_tempo": 16 * 4,      # u,v,w,nu_t,mu,rho,kappa,Hx,Hy,Hz (+u0,v0,w0) -> u*,v*,w* (+u0,v0,w0)
    "posdin": 8 + 7 * 4,           # p (FP64), rho, u, v, w -> u, v, w
    "est_div": 3 * 4,
    "fcg_spmv_dot": 8 * 4,         # z, p_old, r, ax, ay, az -> p, q
    "fcg_update": 6 * 4,           # p, q, r, d -> d, r
    "fcg_zq": 2 * 4,
    "fcg_residual64": 7 * 8,       # x, b, sqrt(D), ax, ay, az (FP64) -> r (FP64)
    "fcg_project": 8 + 8 + 4 + 4,  # r, sqrt(D) (FP64) -> r32, d
    "fcg_xupdate": 8 + 4 + 8,
    "mg_gs_l0": 2 + 2 + 6 + 2,     # x of the other colour, b of this one (FP32, half the cells each),
                                   # c_x,c_y,c_z (BF16; the diagonal is their sum) -> x of this colour
    "mg_gs_l0_in": 4 + 4 + 2 + 4 + 2,  # r, sqrt(D) (FP32), red dg (half) -> b (both colours), red x
    "mg_gs_l0_prolong": 12 + 1,    # a half sweep + the coarse x (1/8 of the cells)
    "mg_gs_l0_out": 2 + 6 + 2 + 3 * 4,  # black x, faces, red b, sqrt(D), q -> z
    "mg_restrict_l0": 4 + 4 + 3 * 2 + 1,  # b, x (FP32), BF16 faces -> coarse b (1/8)
    "mg_prolong_l0": 2 * 4,
    "mg_scale": 3 * 4,
}


def splitmix64(seed):
    x = seed & 0xFFFFFFFFFFFFFFFF
    while True:
        x = (x + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        z = z ^ (z >> 31)
        yield (z >> 11) / float(1 << 53)


def column_geometry(seed):
    r = splitmix64(seed)
    width = 0.2 + 0.2 * next(r)
    height = 0.3 + 0.15 * next(r)
    return width, height


def prof_table(L):
    import ctypes as C
    n = L.suli_prof_count()
    out = {}
    for q in range(n):
        name = C.create_string_buffer(128)
        cnt = C.c_int64()
        tot = C.c_double()
        L.suli_prof_get(q, name, 128, C.byref(cnt), C.byref(tot))
        out[name.value.decode()] = (int(cnt.value), float(tot.value))
    return out


def cpu_baseline(sv, nx, ny, nz, h, dt, vol_ini, budget_s):
    """Time the flang-built reference on a bounded sample of the workload."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    p = O.Params(nx=nx, ny=ny, nz=nz, dx=h, dy=h, dz=h)
    exe = os.path.join(O.REF_DIR, O.ref_tag(p), "harness")
    if not os.path.exists(exe):
        return None
    st = O.State(p)
    for n in ["u", "v", "w", "prd1", "ls", "nut"]:
        st.arr(n)[...] = sv.download(n)
    st.scal.update(dt=dt, vol_ini=vol_ini)
    tmp = tempfile.mkdtemp(prefix="suli_cpu_")
    fin = os.path.join(tmp, "state.bin")
    st.to_file(fin)
    del st
    t0 = time.time()
    def _big_stack():  # the reference keeps large automatic arrays on the stack
        import resource
        hard = resource.getrlimit(resource.RLIMIT_STACK)[1]
        resource.setrlimit(resource.RLIMIT_STACK, (hard, hard))

    logf = os.path.join(tmp, "harness.out")
    try:
        with open(logf, "w") as fo:
            proc = subprocess.Popen(["taskset", "-c", "0", exe, fin, "-", "bench"], stdout=fo,
                                    stderr=subprocess.STDOUT, text=True, env=dict(os.environ, OMP_NUM_THREADS="1"),
                                    preexec_fn=_big_stack)
            while True:
                try:
                    proc.wait(timeout=15)
                    break
                except subprocess.TimeoutExpired:
                    el = time.time() - t0
                    progress(f"CPU baseline (reference harness, 1 core) running {el:.0f}s")
                    if el > budget_s:
                        proc.kill()
                        proc.wait()
                        break
        out = err = open(logf).read()
    finally:
        for f in (fin, logf):
            if os.path.exists(f):
                os.remove(f)
        os.rmdir(tmp)
    wall = time.time() - t0
    import re
    vals = {k: float(v) for k, v in re.findall(r"(\w+)=\s*([-+0-9.Ee]+)", out)}
    info = {"unit": "Mcell-steps/s", "cores": 1, "kind": "reference", "cpu_model": _cpu_model(),
            "host_nproc": os.cpu_count(), "wall_s": round(wall, 1)}
    if "level_set_seconds" not in vals:
        info.update(value=None, sample=(f"reference level_set at {nx}x{ny}x{nz} did not finish within {budget_s:.0f}s "
                                        f"(rc {proc.returncode}: {(err or '').strip()[-200:]})"))
        return info
    ls_s = vals["level_set_seconds"]
    if "stage1_seconds" not in vals:
        # the stage (with its CG) ran past the budget: the step takes at least
        # level_set + 2 x (budget left for the stage), so the value is an upper bound
        stage_min = max(wall - ls_s - 5.0, 0.0)
        step_min = ls_s + 2 * stage_min
        info.update(value=nx * ny * nz / step_min / 1e6, bound="upper",
                    sample=(f"flang -O2 reference harness, 1 core, {nx}x{ny}x{nz} from the GPU post-warmup state: "
                            f"level_set {ls_s:.1f}s; RK stage 1 still running after {stage_min:.0f}s (stopped); "
                            f"step >= level_set + 2 x {stage_min:.0f}s = {step_min:.0f}s, so the value is an upper bound"))
        return info
    step_s = ls_s + 2 * vals["stage1_seconds"]
    info.update(value=nx * ny * nz / step_s / 1e6,
                sample=(f"flang -O2 reference harness, 1 core: one level_set ({ls_s:.1f}s) + one RK stage "
                        f"({vals['stage1_seconds']:.1f}s, of which graddin {vals['graddin_seconds']:.1f}s) at "
                        f"{nx}x{ny}x{nz} from the GPU post-warmup state; step = level_set + 2 stages = {step_s:.1f}s"))
    return info


def progress(msg):
    print(f"bench.py: {msg}", file=sys.stderr, flush=True)


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_256x128x128.json")


def poisson_comparison(sv, dt, repeats):
    """ms per solve of the same projection (the first RK stage of the next
    step, warm-started from the current pressure) with each preconditioner:
    Jacobi (plain), multigrid, and every shipped learned preconditioner (the
    default one as "learned", all of them under "learned_by_entry")."""
    import pysuli
    sv.level_set(dt)
    sv.contorno(3)
    sv.visco()
    sv.convdiff()
    sv.tempo(1, 0.5 * dt)
    sv.contorno(2)
    p0 = sv.download("prd1")
    out = {}
    cases = [("jacobi", pysuli.PRECOND_JACOBI, None), ("gmg", pysuli.PRECOND_GMG, None)]
    shipped = [m for m in pysuli.shipped_learned() if sv.nx % (1 << (m["entry"] + m["levels"] - 1)) == 0
               and sv.ny % (1 << (m["entry"] + m["levels"] - 1)) == 0]
    cases += [(m["name"], pysuli.PRECOND_LEARNED, m) for m in shipped]
    for name, pc, meta in cases:
        if meta is not None:
            pysuli.use_learned(sv, meta)
        sv.set_solver(pysuli.POISSON_FCG, pc)
        runs = []
        for _ in range(repeats):
            sv.upload("prd1", p0)
            runs.append(sv.graddin(0.5 * dt))
        ms = sorted(r["ms"] for r in runs)
        rec = {"ms_per_solve": round(ms[len(ms) // 2], 3), "ms_min": round(ms[0], 3),
               "iters": runs[-1]["iters"], "outer": runs[-1]["outer"], "fallbacks": runs[-1]["fallbacks"],
               "converged": runs[-1]["converged"], "relres": runs[-1]["res"]}
        if meta is None:
            out[name] = rec
            continue
        rec.update(entry=meta["entry"], levels=meta["levels"], weights=f"suli-openmp_amd/weights/{name}.npy",
                   plain_over=round(out["jacobi"]["ms_per_solve"] / rec["ms_per_solve"], 3))
        out.setdefault("learned_by_entry", {})[name] = rec
        if meta["default"] or "learned" not in out:
            out["learned"] = rec
    if "learned" in out:
        out["plain_over_learned"] = out["learned"]["plain_over"]
        out["gmg_over_learned"] = round(out["gmg"]["ms_per_solve"] / out["learned"]["ms_per_solve"], 3)
    out["plain_over_gmg"] = round(out["jacobi"]["ms_per_solve"] / out["gmg"]["ms_per_solve"], 3)
    out["note"] = ("one first-stage projection of the 256x128x128 state after the timed steps, warm-started "
                   f"from the current pressure, median of {repeats} solves, FP64 relative residual <= 1e-8; "
                   "plain = Jacobi-FCG (the reference's symmetric scaling); learned = the default shipped "
                   "network; entry = the multigrid level a network solves (0 = the whole grid); the entry >= 1 "
                   "networks were trained on the column geometries of seeds 1-6, not on this one")
    return out


def hbm_traffic(kernel):
    """HBM bytes per launch of a kernel from the committed rocprof counter
    summary (FETCH_SIZE + WRITE_SIZE, scaled by a copy kernel), or None."""
    try:
        t = json.load(open(TRAFFIC_FILE))
    except (OSError, ValueError):
        return None
    k = t.get("kernels", {}).get(kernel)
    return None if k is None else k.get("bytes_per_launch")


# timing scope -> kernel of that scope (the finest-grid dispatches of it)
PMC_KERNELS = {"mg_gs_l0": "k_mg0_gs", "mg_gs_l0_in": "k_mg0_in", "mg_gs_l0_prolong": "k_mg0_prolong",
               "mg_gs_l0_out": "k_mg0_out", "mg_restrict_l0": "k_mg0_restrict", "mg_prolong_l0": "k_mg_prolong",
               "mg_scale": "k_scale", "fcg_spmv_dot": "k_fcg1", "fcg_update": "k_fcg2", "fcg_zq": "k_zq",
               "fcg_residual64": "k_res64", "fcg_project": "k_r32", "fcg_xupdate": "k_xupd",
               "convdiff_tempo": "k_momentum", "ls_advect_rk3": "k_ls_adv", "reinit_rk3": "k_reinit_stage",
               "reinit_sign": "k_reinit_sign", "heaviside_normals": "k_heavi", "curvature": "k_curv",
               "visco": "k_visco", "posdin": "k_posdin", "est_div": "k_divmax"}


def _kernel_ident(name):
    """'void k_momentum<float>(Grid, ...)' -> 'k_momentum'."""
    import re
    n = name[5:] if name.startswith("void ") else name
    return re.split(r"[<(]", n, maxsplit=1)[0].strip()


def _pmc_rows(d):
    """{dispatch: (kernel, grid, {counter: value})} from a rocprofv3 --pmc csv directory."""
    import csv
    import glob
    out = {}
    for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
        for r in csv.DictReader(open(f)):
            k = out.setdefault(r["Dispatch_Id"], [r["Kernel_Name"], int(r["Grid_Size"]), {}])
            k[2][r["Counter_Name"]] = k[2].get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
    return out


def pmc_summary(fetch_dir, write_dir, cal_fetch_dir, cal_write_dir, ncell):
    """HBM bytes per launch of each timed kernel: rocprof FETCH_SIZE and
    WRITE_SIZE (KiB per dispatch) scaled by what they read for a 256 MiB
    device copy, written to the traffic file read by the benchmark."""
    cal_bytes = (64 << 20) * 4

    def factor(d, counter):
        vals = [c[counter] for _, _, c in _pmc_rows(d).values() if c.get(counter, 0) * 1024 > 0.5 * cal_bytes]
        return cal_bytes / (1024 * sum(vals) / len(vals)), len(vals)

    ff, nf = factor(cal_fetch_dir, "FETCH_SIZE")
    fw, nw = factor(cal_write_dir, "WRITE_SIZE")
    fetch, write = _pmc_rows(fetch_dir), _pmc_rows(write_dir)
    res = {"factor_fetch": ff, "factor_write": fw, "calibration": f"256 MiB torch copy, {nf}/{nw} dispatches",
           "cells": ncell, "kernels": {}}
    for scope, kern in PMC_KERNELS.items():
        def per_launch(rows, counter):
            ks = [(g, c.get(counter)) for name, g, c in rows.values() if _kernel_ident(name) == kern]
            ks = [(g, v) for g, v in ks if v is not None]
            if not ks:
                return None, 0
            gmax = max(g for g, _ in ks)
            v = [x for g, x in ks if g == gmax]
            # dispatches enqueued past the end of a converged pass return at
            # once and move almost nothing: only those that ran are averaged
            v = [x for x in v if x > 0.25 * max(v)]
            return sum(v) / len(v), len(v)
        fk, n1 = per_launch(fetch, "FETCH_SIZE")
        wk, n2 = per_launch(write, "WRITE_SIZE")
        if fk is None or wk is None:
            continue
        b = (fk * ff + wk * fw) * 1024
        res["kernels"][scope] = {"kernel": kern, "fetch_kib": round(fk, 1), "write_kib": round(wk, 1),
                                 "bytes_per_launch": round(b), "bytes_per_cell": round(b / ncell, 2),
                                 "algorithmic_bytes_per_cell": BYTES_PER_CELL.get(scope), "dispatches": [n1, n2]}
    with open(TRAFFIC_FILE, "w") as f:
        json.dump(res, f, indent=1)
    print(json.dumps(res, indent=1))


def pmc_calibrate():
    """A plain device copy of known size, for scaling the rocprof counters."""
    import torch
    n = 64 << 20  # 64 Mi floats = 256 MiB each way
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_each_way": n * 4, "launches": 5}))


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nx", type=int, default=256)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--precond", default="gmg", choices=["jacobi", "gmg", "learned"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-budget", type=float, default=180.0)
    ap.add_argument("--no-poisson", action="store_true", help="skip the preconditioner comparison")
    ap.add_argument("--poisson-repeats", type=int, default=3)
    ap.add_argument("--pmc-calibrate", action="store_true",
                    help="only run a 256 MiB device copy (scale factor of the rocprof byte counters)")
    ap.add_argument("--pmc-summary", nargs=4, metavar=("FETCH_DIR", "WRITE_DIR", "CAL_FETCH_DIR", "CAL_WRITE_DIR"),
                    help="summarise rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs into " + TRAFFIC_FILE)
    args = ap.parse_args()
    if args.pmc_calibrate:
        pmc_calibrate()
        return
    if args.pmc_summary:
        pmc_summary(*args.pmc_summary, ncell=args.nx * (args.nx // 2) ** 2)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs the time step on one GPU; --gpus {args.gpus} is not supported "
              f"(the multi-GPU slab decomposition is a design, not an implementation)", file=sys.stderr)
        sys.exit(2)
    import pysuli
    nx = args.nx
    ny, nz = nx // 2, nx // 2
    h = 1.0 / nx
    width, height = column_geometry(args.seed)
    g = 9.81
    dt = 0.2 * h / (2.0 * math.sqrt(g * height))
    pc = {"jacobi": pysuli.PRECOND_JACOBI, "gmg": pysuli.PRECOND_GMG, "learned": pysuli.PRECOND_LEARNED}[args.precond]
    sv = pysuli.Solver(nx, ny, nz, h, precision=pysuli.PREC_MIXED, poisson=pysuli.POISSON_FCG, precond=pc,
                       tol_rel=1e-8)
    L = pysuli.lib()
    if args.precond == "learned":  # the default shipped network
        shipped = pysuli.shipped_learned()
        if not shipped:
            sys.exit("bench.py: --precond learned needs the weights in suli-openmp_amd/weights")
        pysuli.use_learned(sv, shipped[0])
    vol_ini = sv.init_dambreak(width, height)
    sv.contorno(1)
    progress(f"{nx}x{ny}x{nz} initialised; {args.warmup} warm-up steps")
    for _ in range(max(args.warmup - 1, 0)):
        sv.step(dt)
    # one untimed step with every kernel timed: the per-kernel table and the
    # choice of the dominant kernel (events around every launch slow a step,
    # so the timed steps below record events around that kernel only)
    L.suli_prof_enable(1)
    sv.step(dt)
    sv.sync()
    prof = prof_table(L)
    known = {k: v for k, v in prof.items() if k in BYTES_PER_CELL}
    dname = max(known.items(), key=lambda kv: kv[1][1])[0]
    L.suli_prof_enable(1)
    # events on every 7th launch of it: 7 is prime to the 8 finest-grid sweeps
    # of a V-cycle, so the sample covers every position in the cycle
    L.suli_prof_select(dname.encode(), 7)
    stats = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ls, ps = sv.step(dt)
        stats.append((ls, ps))
    sv.sync()
    t1 = time.perf_counter()
    progress(f"timed {args.steps} steps in {t1 - t0:.3f}s")
    dcount, dms = prof_table(L).get(dname, (0, 0.0))
    L.suli_prof_enable(0)
    secs = t1 - t0
    ncell = nx * ny * nz
    value = ncell * args.steps / secs / 1e6
    solves = [p for _, ps in stats for p in ps]
    pois_ms = sum(p["ms"] for p in solves) / len(solves)
    pois_it = sum(p["iters"] for p in solves) / len(solves)
    # roofline of the dominant kernel over the timed steps
    bpc = BYTES_PER_CELL.get(dname)
    roof = None
    if bpc is not None and dcount > 0:
        avg_s = dms / dcount / 1e3
        achieved = bpc * ncell / avg_s / 1e9
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": hbm_traffic(dname), "kernel": dname,
                "bytes_per_launch": bpc * ncell, "avg_launch_us": round(avg_s * 1e6, 2), "launches_timed": dcount,
                "sampling": "HIP events on every 7th launch over the timed steps",
                "note": "largest fine-grid kernel by total time; coarse multigrid levels (mg_gs, mg_restrict, "
                        "mg_prolong on levels >= 1) are latency-bound and reported under kernels"}
    vol_now = sv.volumes()[1]
    maxdiv = sv.est_div()
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(sv, nx, ny, nz, h, dt, vol_ini, args.cpu_budget)
    if not args.no_poisson:
        progress("preconditioner comparison")
    comparison = None if args.no_poisson else poisson_comparison(sv, dt, args.poisson_repeats)
    sv.close()
    step_ms = sum(t for _, t in prof.values())
    kernels = {k: {"launches": c, "ms_total": round(t, 3), "share_of_timed_kernels": round(t / step_ms, 4)}
               for k, (c, t) in sorted(prof.items(), key=lambda kv: -kv[1][1])}
    line = {
        "metric": "Mcell-steps/s at 256x128x128; Poisson ms/solve to 1e-8, learned vs plain",
        "value": round(value, 3),
        "unit": "Mcell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(secs / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 (fp64 pressure, residuals, dots, integrals)",
        "data": "synthetic (seeded dam-break tank initialised in place)",
        "config": {"workload": f"dam break {nx}x{ny}x{nz}, h=1/{nx}, column {width:.4f} x {height:.4f} m (seed {args.seed})",
                   "grid": [nx, ny, nz], "dt": dt, "poisson": f"FCG+FP64 refinement to 1e-8, {args.precond}",
                   "parallelism": "single GPU"},
        "poisson": {"ms_per_solve": round(pois_ms, 3), "iters_per_solve": round(pois_it, 2),
                    "outer_per_solve": round(sum(p["outer"] for p in solves) / len(solves), 2),
                    "fallbacks": sum(p["fallbacks"] for p in solves),
                    "all_converged": all(p["converged"] for p in solves),
                    "max_relres": max(p["res"] for p in solves), "precond": args.precond,
                    "comparison": comparison},
        "diagnostics": {"volume_drift": (vol_now - vol_ini) / vol_ini, "max_div": maxdiv},
        "roofline": roof,
        "cpu_baseline": cpu,
        "kernels": {"note": "one untimed step after the warm-up, every launch timed with HIP events", **kernels},
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
