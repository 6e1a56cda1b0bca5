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
stages each with visco, convdiff+tempo, contorno, projection, posdin).  The
run has --total steps (default 1000, 0.2 s of the collapse): --advance
untimed steps (default 500: the column is falling and the surge runs along
the floor), --warmup steps of which the last is timed kernel by kernel,
then --steps timed steps between two stream synchronisations, then the rest
of the run untimed.  value = nx*ny*nz*steps/seconds/1e6 (Mcell-steps/s).

During the timed steps only every 7th launch of the dominant kernel carries
events, attached to its dispatches (hipExtLaunchKernelGGL; an event pair
costs a few microseconds of dispatch time), so the other launches run
undisturbed.  The JSON line carries:
- roofline: algorithmic bytes per launch of the dominant kernel / its average
  duration over the timed steps, against 8 TB/s; traffic = HBM bytes per
  launch from rocprof FETCH_SIZE + WRITE_SIZE (profiles/hbm_traffic_*.json,
  made by --pmc-summary from counter runs of this script, scaled by a 256 MiB
  device copy measured the same way);
- poisson: GPU ms per solve of the timed steps, and per 100-step window of
  the whole run; comparison = the projection benchmark set (both RK stages
  of steps 1, 10, 100, 500, 1000, each warm- and cold-started) solved with
  the plain (Jacobi) FCG, the multigrid FCG and the default learned
  preconditioner, and every shipped network on the step-500 state, all to a
  1e-8 FP64 relative residual;
- diagnostics: interface volume drift after the timed steps and at the end
  of the run, max |div u|;
- cpu_baseline: the flang-built reference harness (-O3, oracle/_ref) timed on
  one host core on a bounded sample of the same workload (one level_set plus
  one RK stage from the GPU's state after the timed steps; per-step time =
  level_set + 2 x stage), with the reference's own CG stopping rule and with
  the relative 1e-8 rule, with the reference's CG iteration counts.
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
    "reinit_rk3": 4 * 4,           # phi, phi_it, S -> phi (the first stage: phi -> S, phi)
    "heaviside_normals": 11 * 4,   # phi -> H, Hx, Hy, Hz, rho, mu, |grad phi|, n (3)
    "curvature": 7 * 4,            # n (3) -> d (3), kappa
    "volume_correction": 3 * 4,
    "visco": 5 * 4,                # u, v, w, rho -> nu_t
    "convdiff_tempo": 16 * 4,      # u,v,w,nu_t,mu,rho,kappa,Hx,Hy,Hz (+u0,v0,w0) -> u*,v*,w* (+u0,v0,w0)
    "posdin": 8 + 7 * 4,           # p (FP64), rho, u, v, w -> u, v, w
    "est_div": 3 * 4,
    "fcg_spmv_dot": 8 * 4,         # z, p_old, r, ax, ay, az -> p, q
    "fcg_update": 6 * 4 + 9,       # p, q, r, d -> d, r; and the next V-cycle's finest input:
                                   # s -> b (FP16, either colour), red cells D (FP32) -> x (FP16; half the cells each)
    "fcg_zq": 2 * 4,
    "fcg_residual64": 7 * 8,       # x, b, sqrt(D), ax, ay, az (FP64) -> r (FP64)
    "fcg_project": 8 + 8 + 4 + 4 + 9,  # r, sqrt(D) (FP64) -> r32, d; and the first V-cycle's input as fcg_update
    "fcg_xupdate": 8 + 4 + 8,
    "mg_gs_l0": 1 + 1 + 6 + 1,     # x of the other colour, b of this one (FP16, half the cells each),
                                   # c_x,c_y,c_z (BF16; the diagonal is their sum) -> x of this colour
    "mg_gs_l0_in": 4 + 4 + 2 + 4 + 2,  # r, sqrt(D), red dg (half) -> b (both colours), red x (FP32: unstaged input)
    "mg_gs_l0_prolong": 9 + 0.5,   # a half sweep + the coarse x (FP32, 1/8 of the cells)
    "mg_gs_l0_out": 1 + 6 + 1 + 3 * 4,  # black x, faces, red b (FP16), sqrt(D), q -> z
    "mg_restrict_l0": 2 + 2 + 3 * 2 + 1,  # b, x (FP16), BF16 faces -> coarse b, x (FP32, 1/8)
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


def _run_harness(exe, fin, budget_s, label):
    """Run a reference harness 'bench' on one pinned core; its output, or
    what it printed before the budget ran out, and the wall time."""
    def _big_stack():  # the reference keeps large automatic arrays on the stack
        import resource
        hard = resource.getrlimit(resource.RLIMIT_STACK)[1]
        resource.setrlimit(resource.RLIMIT_STACK, (hard, hard))

    logf = fin + f".{label}.out"
    t0 = time.time()
    killed = False
    with open(logf, "w") as fo:
        proc = subprocess.Popen(["taskset", "-c", "0", exe, fin, "-", "bench"], stdout=fo, stderr=subprocess.STDOUT,
                                text=True, env=dict(os.environ, OMP_NUM_THREADS="1"), preexec_fn=_big_stack,
                                cwd=os.path.dirname(fin))
        while True:
            try:
                proc.wait(timeout=15)
                break
            except subprocess.TimeoutExpired:
                el = time.time() - t0
                progress(f"CPU baseline ({label}, reference harness, 1 core) running {el:.0f}s")
                if el > budget_s:
                    proc.kill()
                    proc.wait()
                    killed = True
                    break
    out = open(logf).read()
    os.remove(logf)
    return out, time.time() - t0, killed, proc.returncode


def cpu_baseline(sv, nx, ny, nz, h, dt, vol_ini, budget_s, step_label):
    """Time the flang-built reference (-O3, as its Makefile:7) on a bounded
    sample of the workload: one level_set plus one RK stage from the GPU's
    state after the timed steps, once with the reference's own stopping rule
    (sum r_s^2 <= 1e-4/N, 9_grad.f90:175) and once with the relative rule
    the GPU solves to (||r_s|| <= 1e-8 ||b_s||).  Per step = level_set + 2
    stages (the second stage's solve is taken as long as the first's)."""
    import re
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    p = O.Params(nx=nx, ny=ny, nz=nz, dx=h, dy=h, dz=h)
    exes = {v: os.path.join(O.REF_DIR, O.ref_tag(p) + "_" + v, "harness") for v in ("bench", "bench_rel")}
    if not os.path.exists(exes["bench"]):
        return None
    st = O.State(p)
    for n in ["u", "v", "w", "prd1", "ls", "nut"]:
        st.arr(n)[...] = sv.download(n)
    st.scal.update(dt=dt, vol_ini=vol_ini)
    tmp = tempfile.mkdtemp(prefix="suli_cpu_")
    fin = os.path.join(tmp, "state.bin")
    st.to_file(fin)
    del st
    runs = {}
    try:
        for v, share in (("bench", 0.4), ("bench_rel", 0.6)):
            if not os.path.exists(exes[v]):
                continue
            out, wall, killed, rc = _run_harness(exes[v], fin, share * budget_s, v)
            vals = {k: float(x) for k, x in re.findall(r"(\w+)=\s*([-+0-9.Ee]+)", out)}
            its = [int(x) for x in re.findall(r"REF_CG_ITERS\s+(\d+)", out)]
            runs[v] = dict(vals=vals, cg_iters=its[0] if its else None, wall=wall, killed=killed, rc=rc,
                           tail=out.strip()[-200:])
    finally:
        if os.path.exists(fin):
            os.remove(fin)
        os.rmdir(tmp)
    info = {"unit": "Mcell-steps/s", "cores": 1, "kind": "reference", "cpu_model": _cpu_model(),
            "host_nproc": os.cpu_count(), "build": "flang -O3 -fdefault-real-8 -fdefault-double-8 (Makefile:7)"}

    def per_step(r):
        vals = r["vals"]
        if "level_set_seconds" not in vals:
            return None, None
        ls_s = vals["level_set_seconds"]
        if "stage1_seconds" in vals:
            return ls_s + 2 * vals["stage1_seconds"], False
        stage_min = max(r["wall"] - ls_s - 5.0, 0.0)  # still running when stopped: a lower bound
        return ls_s + 2 * stage_min, True

    a = runs.get("bench")
    step_s, bound = per_step(a)
    if step_s is None:
        info.update(value=None, sample=(f"reference level_set at {nx}x{ny}x{nz} did not finish within "
                                        f"{0.4 * budget_s:.0f}s (rc {a['rc']}: {a['tail']})"))
        return info
    v = a["vals"]
    info.update(value=nx * ny * nz / step_s / 1e6,
                sample=(f"flang -O3 reference harness, 1 core (taskset), {nx}x{ny}x{nz} from the GPU state after "
                        f"{step_label}: one level_set ({v['level_set_seconds']:.1f}s) + one RK stage "
                        + (f"({v['stage1_seconds']:.1f}s, of which graddin {v['graddin_seconds']:.1f}s in "
                           f"{a['cg_iters']} CG iterations to the reference's sum r^2 <= 1e-4/N)"
                           if not bound else "(stopped at the budget)")
                        + f"; step = level_set + 2 stages = {step_s:.1f}s"),
                reference_cg_iters=a["cg_iters"], level_set_s=v.get("level_set_seconds"),
                stage_s=v.get("stage1_seconds"), graddin_s=v.get("graddin_seconds"))
    if bound:
        info["bound"] = "upper"
    r = runs.get("bench_rel")
    if r is not None:
        srel, brel = per_step(r)
        vr = r["vals"]
        info["relative_1e-8"] = {
            "value": None if srel is None else nx * ny * nz / srel / 1e6,
            "bound": "upper" if brel else None,
            "cg_iters": r["cg_iters"], "stage_s": vr.get("stage1_seconds"), "graddin_s": vr.get("graddin_seconds"),
            "graddin_ms_per_solve": None if "graddin_seconds" not in vr else round(1e3 * vr["graddin_seconds"], 1),
            "note": "the same sample with the reference's CG stopped at ||r_s|| <= 1e-8 ||b_s|| (the GPU's rule; "
                    "build variant bench_rel of oracle/ref/build_ref.sh)"}
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


SNAP_FIELDS = ("u", "v", "w", "rho", "prd1")


def step_capturing(sv, dt, t_new):
    """One time step as suli_step runs it (1_principal.f90:74-95), call by
    call, keeping each RK stage's projection input (u*, rho, the pressure it
    starts from) before its solve.  Returns the solves' stats and the two
    snapshots."""
    sv.set_time(t_new)
    sv.level_set(dt)
    sv.contorno(3)
    stats, snaps = [], []
    for tt in (1, 2):
        dts = 0.5 * dt if tt == 1 else dt
        sv.visco()
        sv.convdiff()
        sv.tempo(tt, dts)
        sv.contorno(2)
        snaps.append({"stage": tt, "dt": dts, **{n: sv.download(n) for n in SNAP_FIELDS}})
        stats.append(sv.graddin(dts))
        sv.posdin(dts)
        sv.contorno(1)
    return stats, snaps


def _solve_record(runs):
    ms = sorted(r["ms"] for r in runs)
    r = runs[-1]
    return {"ms_per_solve": round(ms[len(ms) // 2], 3), "ms_min": round(ms[0], 3), "iters": r["iters"],
            "outer": r["outer"], "fallbacks": r["fallbacks"], "converged": r["converged"], "relres": r["res"]}


def _load_snapshot(sv, snap):
    for n in SNAP_FIELDS:
        sv.upload(n, snap[n])


def _solve_snapshot(sv, snap, cold, repeats):
    """Solves of a loaded snapshot (the solve writes only the pressure)."""
    runs = []
    for _ in range(repeats):
        sv.upload("prd1", 0.0 * snap["prd1"] if cold else snap["prd1"])
        runs.append(sv.graddin(snap["dt"]))
    return _solve_record(runs)


def poisson_comparison(sv, snaps, repeats):
    """ms per solve of the projection benchmark set: the (u*, rho, p) of both
    RK stages captured at steps 1, 10, 100, 500 and 1000 of the run, each
    solved warm (from the pressure the reference would start from,
    9_grad.f90:121) and cold (from p = 0), to a 1e-8 FP64 relative residual,
    with plain Jacobi-FCG, multigrid FCG and the default learned
    preconditioner; every shipped learned preconditioner on the step-500
    stage-1 warm state.  The solver's own extrapolated guess is off here
    (suli_set_guess(0)); ms = GPU time between stream events around each
    solve, median of `repeats`."""
    import pysuli
    sv.set_guess(0)
    shipped = [m for m in pysuli.shipped_learned() if sv.nx % (1 << (m["entry"] + m["levels"] - 1)) == 0
               and sv.ny % (1 << (m["entry"] + m["levels"] - 1)) == 0]
    default = next((m for m in shipped if m["default"]), shipped[0] if shipped else None)
    cases = [("jacobi", pysuli.PRECOND_JACOBI, None), ("gmg", pysuli.PRECOND_GMG, None)]
    if default is not None:
        cases.append(("learned", pysuli.PRECOND_LEARNED, default))
    table = []
    totals = {name: {"ms": 0.0, "iters": 0, "fallbacks": 0, "solves": 0, "all_converged": True}
              for name, _, _ in cases}
    for snap in snaps:
        _load_snapshot(sv, snap)
        for cold in (False, True):
            row = {"step": snap["step"], "stage": snap["stage"], "start": "cold" if cold else "warm"}
            for name, pc, meta in cases:
                if meta is not None:
                    pysuli.use_learned(sv, meta)
                sv.set_solver(pysuli.POISSON_FCG, pc)
                rec = _solve_snapshot(sv, snap, cold, repeats)
                row[name] = {k: rec[k] for k in ("ms_per_solve", "iters", "outer", "fallbacks", "converged")}
                t = totals[name]
                t["ms"] += rec["ms_per_solve"]
                t["iters"] += rec["iters"]
                t["fallbacks"] += rec["fallbacks"]
                t["solves"] += 1
                t["all_converged"] &= bool(rec["converged"])
            table.append(row)
            progress(f"Poisson set: step {row['step']} stage {row['stage']} {row['start']}: "
                     + ", ".join(f"{n} {row[n]['ms_per_solve']:.2f} ms/{row[n]['iters']} it" for n, _, _ in cases))
    out = {}
    for name, t in totals.items():
        n = max(t["solves"], 1)
        out[name] = {"ms_per_solve": round(t["ms"] / n, 3), "iters_per_solve": round(t["iters"] / n, 2),
                     "fallbacks": t["fallbacks"], "solves": t["solves"], "all_converged": t["all_converged"]}
    if "learned" in out:
        out["learned"].update(weights=f"suli-openmp_amd/weights/{default['name']}.npy", entry=default["entry"],
                              levels=default["levels"])
        out["plain_over_learned"] = round(out["jacobi"]["ms_per_solve"] / out["learned"]["ms_per_solve"], 3)
        out["gmg_over_learned"] = round(out["gmg"]["ms_per_solve"] / out["learned"]["ms_per_solve"], 3)
    out["plain_over_gmg"] = round(out["jacobi"]["ms_per_solve"] / out["gmg"]["ms_per_solve"], 3)
    # every shipped network on one state
    ref = next((s for s in snaps if s["step"] == 500 and s["stage"] == 1), snaps[len(snaps) // 2])
    _load_snapshot(sv, ref)
    sv.set_solver(pysuli.POISSON_FCG, pysuli.PRECOND_JACOBI)
    plain = _solve_snapshot(sv, ref, False, repeats)
    by_entry = {}
    for m in shipped:
        pysuli.use_learned(sv, m)
        sv.set_solver(pysuli.POISSON_FCG, pysuli.PRECOND_LEARNED)
        rec = _solve_snapshot(sv, ref, False, repeats)
        rec.update(entry=m["entry"], levels=m["levels"], weights=f"suli-openmp_amd/weights/{m['name']}.npy",
                   plain_over=round(plain["ms_per_solve"] / rec["ms_per_solve"], 3))
        by_entry[m["name"]] = rec
    out["learned_by_entry"] = {"state": f"step {ref['step']} stage {ref['stage']} warm", "jacobi": plain, **by_entry}
    out["set"] = table
    out["note"] = (f"projection benchmark set: steps {sorted({s['step'] for s in snaps})} x both RK stages, warm "
                   "(from the current pressure) and cold (p = 0); ms_per_solve = mean over the set of the median "
                   f"of {repeats} solves (GPU time), FP64 relative residual <= 1e-8; plain = Jacobi-FCG (the "
                   "reference's symmetric scaling); learned = the default shipped network; entry = the multigrid "
                   "level a network solves (0 = the whole grid); the entry >= 1 networks were trained on the "
                   "column geometries of seeds 1-6, not on this one")
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
PMC_KERNELS = {"mg_gs_l0": "k_mg0_gs2", "mg_gs_l0_in": "k_mg0_in", "mg_gs_l0_prolong": "k_mg0_prolong",
               "mg_gs_l0_out": "k_mg0_out2", "mg_restrict_l0": "k_mg0_restrict", "mg_prolong_l0": "k_mg_prolong",
               "mg_scale": "k_scale", "fcg_spmv_dot": "k_fcg1", "fcg_update": "k_fcg2", "fcg_zq": "k_zq",
               "fcg_residual64": "k_res64", "fcg_project": "k_r32", "fcg_xupdate": "k_xupd",
               "convdiff_tempo": ("k_momentum_tile", "k_momentum"), "ls_advect_rk3": "k_ls_adv", "reinit_rk3": "k_reinit_stage",
               "reinit_sign": "k_reinit_sign", "heaviside_normals": "k_heavi", "curvature": "k_curv",
               "visco": "k_visco", "posdin": "k_posdin", "est_div": "k_divmax"}


def _kernel_ident(name):
    """'void k_momentum<float>(Grid, ...)' -> 'k_momentum'; a name the profiler
    left mangled ('_Z9k_mg0_gs2IDF16_E...', the _Float16 instantiations) ->
    'k_mg0_gs2'."""
    import re
    m = re.match(r"_Z(\d+)", name)
    if m:
        return name[m.end():m.end() + int(m.group(1))]
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
    for scope, kerns in PMC_KERNELS.items():
        def per_launch(rows, counter, kern):
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
        # a scope of several kernels (the momentum's inner rows and boundary
        # rows): the bytes of one launch of each, summed
        kerns = kerns if isinstance(kerns, tuple) else (kerns,)
        fk = wk = 0.0
        n1 = n2 = 0
        for kern in kerns:
            f1, c1 = per_launch(fetch, "FETCH_SIZE", kern)
            w1, c2 = per_launch(write, "WRITE_SIZE", kern)
            if f1 is None or w1 is None:
                fk = None
                break
            fk, wk, n1, n2 = fk + f1, wk + w1, n1 + c1, n2 + c2
        if fk is None:
            continue
        kern = "+".join(kerns)
        b = (fk * ff + wk * fw) * 1024
        res["kernels"][scope] = {"kernel": kern, "fetch_kib": round(fk, 1), "write_kib": round(wk, 1),
                                 "bytes_per_launch": round(b), "bytes_per_cell": round(b / ncell, 2),
                                 "algorithmic_bytes_per_cell": BYTES_PER_CELL.get(scope), "dispatches": [n1, n2]}
    with open(TRAFFIC_FILE, "w") as f:
        json.dump(res, f, indent=1)
    print(json.dumps(res, indent=1))


def stream_probe():
    """Practical streaming rates of this GPU for the roofline fractions:
    torch's device copy, sum and fill over working sets of 16 MB to 1 GiB,
    each launched 50 times back to back (sets up to about 256 MB stay in the
    Infinity Cache between launches; 1 GiB ones stream from HBM)."""
    import torch
    out = []
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    for mb in (16, 38, 64, 150, 256, 1024):
        n = (mb << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        half_a, half_b = a[: n // 2], torch.empty(n // 2, dtype=torch.float32, device="cuda")
        for kind in ("copy", "sum", "fill"):
            def run():
                if kind == "copy":
                    half_b.copy_(half_a)  # mb moved: half read, half written
                elif kind == "sum":
                    a.sum()
                else:
                    a.fill_(1.0)
            for _ in range(3):
                run()
            ev0.record()
            for _ in range(50):
                run()
            ev1.record()
            torch.cuda.synchronize()
            us = ev0.elapsed_time(ev1) * 1e3 / 50
            rec = {"kind": kind, "mb_moved": mb, "us_per_launch": round(us, 2), "tb_per_s": round((n * 4) / us / 1e6, 2)}
            out.append(rec)
            print(json.dumps(rec), flush=True)
        del a, half_a, half_b
        torch.cuda.empty_cache()
    return out


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


def _windows(per_step, width=100):
    """Solve statistics per window of `width` steps: iterations and GPU ms
    per solve, fallbacks, largest final residual."""
    out = []
    for w0 in range(1, max(per_step) + 1, width):
        ss = [p for it in range(w0, w0 + width) for p in per_step.get(it, [])]
        if not ss:
            continue
        out.append({"steps": f"{w0}-{min(w0 + width - 1, max(per_step))}", "solves": len(ss),
                    "iters_per_solve": round(sum(p["iters"] for p in ss) / len(ss), 2),
                    "ms_per_solve": round(sum(p["ms"] for p in ss) / len(ss), 3),
                    "outer_per_solve": round(sum(p["outer"] for p in ss) / len(ss), 2),
                    "fallbacks": sum(p["fallbacks"] for p in ss),
                    "max_relres": max(p["res"] for p in ss)})
    return out


CONFIG0 = dict(nx=64, ny=32, nz=32, width=0.25, height=0.25, steps=200)


def config0(cpu=True):
    """The reference's own CPU-runnable case (64x32x32 dam break, column
    0.25 x 0.25 m, dt = 0.2 h / (2 sqrt(g H)), 200 steps): wall time of the
    200 steps on the GPU in FP64 with the reference's CG, in mixed precision
    with GMG-FCG to 1e-10 (the case's Poisson tolerance), and of the
    flang-built reference (-O3) on one pinned core of the host, with its CG
    iteration counts.  A parity case, not the benchmark: printed as its own
    JSON line."""
    import re

    import pysuli
    c = CONFIG0
    nx, ny, nz = c["nx"], c["ny"], c["nz"]
    h = 1.0 / nx
    dt = 0.2 * h / (2.0 * math.sqrt(9.81 * c["height"]))
    out = {"workload": f"dam break {nx}x{ny}x{nz}, column {c['width']} x {c['height']} m, {c['steps']} steps",
           "dt": dt, "gpu": {}}
    for label, kw in (("fp64_reference_cg", dict(precision=pysuli.PREC_FP64, poisson=pysuli.POISSON_REF)),
                      ("mixed_gmg_fcg_1e-10", dict(precision=pysuli.PREC_MIXED, poisson=pysuli.POISSON_FCG,
                                                   precond=pysuli.PRECOND_GMG, tol_rel=1e-10))):
        sv = pysuli.Solver(nx, ny, nz, h, **kw)
        try:
            sv.init_dambreak(c["width"], c["height"])
            sv.contorno(1)
            sv.contorno(3)
            sv.step(dt)  # first-touch
            sv.sync()
            iters = 0
            t0 = time.time()
            for _ in range(c["steps"] - 1):
                _, ps = sv.step(dt)
                iters += sum(q["iters"] for q in ps)
            sv.sync()
            wall = time.time() - t0
            out["gpu"][label] = {"seconds_per_step": wall / (c["steps"] - 1),
                                 "mcell_steps_per_s": nx * ny * nz * (c["steps"] - 1) / wall / 1e6,
                                 "iters_per_solve": iters / (2 * (c["steps"] - 1))}
        finally:
            sv.close()
        progress(f"config 0 on the GPU ({label}): {out['gpu'][label]['seconds_per_step'] * 1e3:.2f} ms/step")
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        p = O.Params(nx=nx, ny=ny, nz=nz, dx=h, dy=h, dz=h)
        exe = os.path.join(O.REF_DIR, O.ref_tag(p) + "_bench", "harness")
        if os.path.exists(exe):
            st = O.State(p)
            st.ls[...] = O.dambreak_phi(p, c["width"], c["height"])
            st.scal["dt"] = dt
            st.run("ls_ini")
            st.run("contorno", 1)
            st.run("contorno", 3)
            tmp = tempfile.mkdtemp(prefix="suli_c0_")
            fin = os.path.join(tmp, "state.bin")
            st.to_file(fin)

            def _big_stack():  # the reference keeps large automatic arrays on the stack
                import resource
                hard = resource.getrlimit(resource.RLIMIT_STACK)[1]
                resource.setrlimit(resource.RLIMIT_STACK, (hard, hard))

            progress("config 0 on the CPU: the reference's 200 steps on one core")
            t0 = time.time()
            r = subprocess.run(["taskset", "-c", "0", exe, fin, "-", "step", str(c["steps"])], capture_output=True,
                               text=True, env=dict(os.environ, OMP_NUM_THREADS="1"), preexec_fn=_big_stack, cwd=tmp,
                               timeout=1800)
            os.remove(fin)
            os.rmdir(tmp)
            vals = {k: float(x) for k, x in re.findall(r"(\w+)=\s*([-+0-9.Ee]+)", r.stdout)}
            its = [int(x) for x in re.findall(r"REF_CG_ITERS\s+(\d+)", r.stdout)]
            out["cpu_reference"] = {"rc": r.returncode, "wall_s": time.time() - t0, "cores": 1,
                                    "cpu_model": _cpu_model(),
                                    "seconds": vals.get("seconds"), "graddin_seconds": vals.get("graddin_seconds"),
                                    "level_set_seconds": vals.get("level_set_seconds"),
                                    "iters_per_solve": sum(its) / len(its) if its else None,
                                    "build": "flang -O3 -fdefault-real-8 -fdefault-double-8 (Makefile:7)"}
            if vals.get("seconds"):
                s = vals["seconds"] / c["steps"]
                out["cpu_reference"].update(seconds_per_step=s, mcell_steps_per_s=nx * ny * nz / s / 1e6)
    print(json.dumps({"config0": out}), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--advance", type=int, default=500,
                    help="untimed steps before the warm-up: the timed window starts in the collapse of the column")
    ap.add_argument("--total", type=int, default=1000,
                    help="steps of the whole run (the steps after the timed window run untimed)")
    ap.add_argument("--snap-steps", default="1,10,100,500,1000",
                    help="steps whose two projections join the Poisson benchmark set")
    ap.add_argument("--nx", type=int, default=256)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--precond", default="gmg", choices=["jacobi", "gmg", "learned"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-budget", type=float, default=240.0)
    ap.add_argument("--no-poisson", action="store_true", help="skip the preconditioner comparison")
    ap.add_argument("--poisson-repeats", type=int, default=3)
    ap.add_argument("--config0", action="store_true",
                    help="time the reference's own 64x32x32 case (GPU FP64 and mixed, and the CPU reference) instead")
    ap.add_argument("--stream-probe", action="store_true",
                    help="only time device copy/sum/fill over 16 MB..1 GiB (practical streaming rates)")
    ap.add_argument("--pmc-calibrate", action="store_true",
                    help="only run a 256 MiB device copy (scale factor of the rocprof byte counters)")
    ap.add_argument("--pmc-summary", nargs=4, metavar=("FETCH_DIR", "WRITE_DIR", "CAL_FETCH_DIR", "CAL_WRITE_DIR"),
                    help="summarise rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs into " + TRAFFIC_FILE)
    args = ap.parse_args()
    if args.stream_probe:
        stream_probe()
        return
    if args.pmc_calibrate:
        pmc_calibrate()
        return
    if args.pmc_summary:
        pmc_summary(*args.pmc_summary, ncell=args.nx * (args.nx // 2) ** 2)
        return
    if args.config0:
        config0(cpu=not args.no_cpu)
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
    ncell = nx * ny * nz
    prof_step = args.advance + args.warmup           # the step timed kernel by kernel
    first, last = prof_step + 1, prof_step + args.steps  # the timed window
    total = max(args.total, last)
    snap_steps = sorted({int(s) for s in args.snap_steps.split(",") if s.strip()
                         if 1 <= int(s) <= total and not prof_step <= int(s) <= last})
    progress(f"{nx}x{ny}x{nz} initialised; steps 1-{prof_step - 1} untimed (snapshots at {snap_steps}), "
             f"step {prof_step} profiled, steps {first}-{last} timed, run to step {total}")
    per_step = {}
    snaps = []
    t = 0.0
    run_wall, run_steps = 0.0, 0

    def untimed(it_from, it_to):
        nonlocal t, run_wall, run_steps
        seg0, nseg = time.perf_counter(), 0
        for it in range(it_from, it_to + 1):
            t = t + dt
            if it in snap_steps:
                sv.sync()
                run_wall += time.perf_counter() - seg0
                run_steps += nseg
                ps, ss = step_capturing(sv, dt, t)
                for s in ss:
                    s["step"] = it
                snaps.extend(ss)
                seg0, nseg = time.perf_counter(), 0
            else:
                _, ps = sv.step(dt)
                nseg += 1
            per_step[it] = ps
            if it % 100 == 0:
                progress(f"step {it}: {ps[0]['iters']}/{ps[1]['iters']} FCG iterations")
        sv.sync()
        run_wall += time.perf_counter() - seg0
        run_steps += nseg

    untimed(1, prof_step - 1)
    # one untimed step with every kernel timed: the per-kernel table and the
    # choice of the dominant kernel (events around every launch slow a step,
    # so the timed steps below record events around that kernel only)
    L.suli_prof_enable(1)
    t = t + dt
    _, per_step[prof_step] = sv.step(dt)
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
    for it in range(first, last + 1):
        ls, ps = sv.step(dt)
        stats.append((ls, ps))
        per_step[it] = ps
    sv.sync()
    t1 = time.perf_counter()
    for _ in range(first, last + 1):
        t = t + dt
    progress(f"timed steps {first}-{last} in {t1 - t0:.3f}s")
    dcount, dms = prof_table(L).get(dname, (0, 0.0))
    L.suli_prof_enable(0)
    secs = t1 - t0
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
    vol_window = sv.volumes()[1]
    div_window = sv.est_div()
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(sv, nx, ny, nz, h, dt, vol_ini, args.cpu_budget, f"step {last} (t = {t:.4f} s)")
    untimed(last + 1, total)
    vol_end = sv.volumes()[1]
    div_end = sv.est_div()
    if not args.no_poisson:
        progress(f"Poisson benchmark set: {len(snaps)} projections x warm/cold")
    comparison = None if args.no_poisson or not snaps else poisson_comparison(sv, snaps, args.poisson_repeats)
    sv.close()
    step_ms = sum(tk for _, tk in prof.values())
    kernels = {}
    for k, (c, tk) in sorted(prof.items(), key=lambda kv: -kv[1][1]):
        kernels[k] = {"launches": c, "ms_total": round(tk, 3), "share_of_timed_kernels": round(tk / step_ms, 4)}
        if k in BYTES_PER_CELL and tk > 0:
            kernels[k]["algorithmic_gbs"] = round(BYTES_PER_CELL[k] * ncell * c / (tk * 1e-3) / 1e9, 1)
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
        "config": {"workload": (f"dam break {nx}x{ny}x{nz}, h=1/{nx}, column {width:.4f} x {height:.4f} m "
                                f"(seed {args.seed}); timed steps {first}-{last} of a {total}-step run "
                                f"(t = {(first - 1) * dt:.4f}-{last * dt:.4f} s of {total * dt:.4f} s)"),
                   "grid": [nx, ny, nz], "dt": dt, "poisson": f"FCG+FP64 refinement to 1e-8, {args.precond}",
                   "timed_steps": [first, last], "run_steps": total, "parallelism": "single GPU"},
        "poisson": {"ms_per_solve": round(pois_ms, 3), "iters_per_solve": round(pois_it, 2),
                    "outer_per_solve": round(sum(p["outer"] for p in solves) / len(solves), 2),
                    "fallbacks": sum(p["fallbacks"] for p in solves),
                    "all_converged": all(p["converged"] for p in solves),
                    "max_relres": max(p["res"] for p in solves), "precond": args.precond,
                    "ms_note": "GPU time of each solve of the timed steps (stream events at its start and after "
                               "its last kernel)",
                    "run_windows": _windows(per_step),
                    "comparison": comparison},
        "diagnostics": {"volume_drift_after_timed": (vol_window - vol_ini) / vol_ini,
                        "max_div_after_timed": div_window,
                        "volume_drift": (vol_end - vol_ini) / vol_ini, "volume_drift_step": total,
                        "max_div": div_end,
                        "untimed_steps_ms_per_step": round(1e3 * run_wall / max(run_steps, 1), 3),
                        "untimed_steps": run_steps},
        "roofline": roof,
        "cpu_baseline": cpu,
        "kernels": {"note": f"step {prof_step}, every launch timed with HIP events; algorithmic_gbs from the "
                            "algorithmic bytes per cell of the fine-grid kernels", **kernels},
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
