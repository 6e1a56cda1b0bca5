"""Generate the golden vectors of the oracle tests from the flang-built
reference (oracle/ref/build_ref.sh + harness.f90), in this container only.

For each case the script builds a deterministic input state (dam-break level
set, seeded smooth random velocities and pressure), runs one reference
subroutine through the harness, and stores the input state and every array
the call changed in tests/golden/<case>.npz (float64, Fortran order).

    python tests/golden/gen_golden.py            # RK2 cases (g8_*, per_*)
    python tests/golden/gen_golden.py schemes    # Euler, RK3, AB2 (g8tt*_*)
    python tests/golden/gen_golden.py variants   # advection forms, directional LES
    python tests/golden/gen_golden.py obstacle   # bottom obstacle columns (g8obst_*)
    python tests/golden/gen_golden.py obstacle_neg  # with columns below the bed, -1 and 0 (g8obstneg_*)
    python tests/golden/gen_golden.py geometry   # obstaculo's columns per obst_t (obstaculo_*)
    python tests/golden/gen_golden.py est        # est's divergence and its maximum (est/est_*)
    python tests/golden/gen_golden.py periodic   # graddin on the x-periodic grid (per_graddin)
    python tests/golden/gen_golden.py weng       # the manufactured solution of 13_mms.f90 (weng/*)
    python tests/golden/gen_golden.py waves      # the wave inflow, Stokes I, II, V (waves/*)

The scheme cases use reference builds with t_tempo = 0, 2, 3
(0_modules.f90:24) and start from nonzero kept forces fu0, fu1.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle as O  # noqa: E402

CASES = {
    # name: (params, state builder, command, arg)
    "g8": O.Params(),
    "per": O.Params(nx=12, ny=8, nz=6, dx=1 / 12, dy=1 / 12, dz=1 / 12, ccx0=0, ccxf=0, cczf=1),
}
# time schemes other than RK2: t_tempo -> stages (ntt, 2_iniciais.f90:85-102)
SCHEMES = {0: 1, 2: 3, 3: 1}
COMMANDS = [("ls_ini", None), ("heaviside", None), ("mod_ls1", None), ("reinic", None), ("level_set", None),
            ("visco", None), ("convdiff", None), ("tempo", 1), ("tempo", 2), ("contorno", 1),
            ("contorno", 2), ("graddin", None), ("posdin", None), ("stage", 1), ("step", 2)]


def base_state(p, seed):
    s = O.random_state(p, seed=seed, perturb=0.4)
    s.run("visco")
    s.run("contorno", 1)
    s.run("convdiff")
    s.u0[...] = 0.9 * s.u[1:-1, 1:-1, 1:-1]
    s.v0[...] = 0.9 * s.v[1:-1, 1:-1, 1:-1]
    s.w0[...] = 0.9 * s.w[1:-1, 1:-1, 1:-1]
    return s


def _save(path, p, s_in, s_out):
    data = {"__scal_in": np.array([s_in.scal[k] for k in O.SCALARS]),
            "__scal_out": np.array([s_out.scal[k] for k in O.SCALARS]),
            "__params": np.array([p.nx, p.ny, p.nz, p.ccx0, p.ccxf, p.ccy0, p.ccyf, p.ccz0, p.cczf]),
            "__h": np.array([p.dx, p.dy, p.dz]), "__tt": np.array([p.t_tempo]),
            "__model": np.array([p.der, p.adv_type, p.m_turb, p.obst_t])}
    for name, _ in O.ARRAYS:
        a, b = s_in.arr(name), s_out.arr(name)
        data["in_" + name] = a
        if not np.array_equal(a, b):
            data["out_" + name] = b
    np.savez_compressed(path, **data)
    return sorted(k[4:] for k in data if k.startswith("out_"))


def schemes():
    import dataclasses
    for tt_scheme, ntt in SCHEMES.items():
        p = dataclasses.replace(O.Params(), t_tempo=tt_scheme)
        O.ref_build(p)
        cmds = [("tempo", t) for t in range(1, ntt + 1)] + [("stage", t) for t in range(1, ntt + 1)] + [("step", 2)]
        for q, (cmd, arg) in enumerate(cmds):
            s_in = base_state(p, seed=300 + 10 * tt_scheme + q)
            rng = np.random.default_rng(1000 + q)
            for n in ["fu0", "fv0", "fw0", "fu1", "fv1", "fw1"]:
                s_in.arr(n)[...] = 0.1 * rng.standard_normal(s_in.arr(n).shape)
            s_out = s_in.copy()
            O.ref_run(s_out, cmd, arg)
            tag = f"{cmd}{arg}"
            changed = _save(os.path.join(HERE, f"g8tt{tt_scheme}_{tag}.npz"), p, s_in, s_out)
            print(f"g8tt{tt_scheme}", tag, "changed:", changed)


# advection (der, adv_type of 0_modules.f90:26-27) and turbulence-model
# variants: case name -> Params overrides
VARIANTS = {"g8adv12": dict(der=2), "g8adv13": dict(der=3), "g8adv21": dict(adv_type=2),
            "g8adv31": dict(adv_type=3), "g8les2": dict(m_turb=2)}


def variants():
    import dataclasses
    for case, kw in VARIANTS.items():
        p = dataclasses.replace(O.Params(), **kw)
        O.ref_build(p)
        cmds = [("visco", None), ("stage", 1)] if "les" in case else [("convdiff", None), ("stage", 1)]
        for q, (cmd, arg) in enumerate(cmds):
            s_in = base_state(p, seed=500 + q)
            if cmd == "convdiff":
                for n in ["Fu", "Fv", "Fw"]:
                    s_in.arr(n)[...] = 0.0
            s_out = s_in.copy()
            O.ref_run(s_out, cmd, arg)
            tag = cmd + ("" if arg is None else str(arg))
            changed = _save(os.path.join(HERE, f"{case}_{tag}.npz"), p, s_in, s_out)
            print(case, tag, "changed:", changed)


def obstacle_masks(p, seed=0, neg=False):
    """Columns of a bottom obstacle (ku, kv, kw of 0_modules.f90:76-79): a
    seeded bump one to three cells high in the middle of the tank, clear of
    the side walls (contorno(3) reads ls(i, j-1, k) beside a lower column,
    outside the array for an obstacle on the first or last row), with ku,
    kv <= nz and kw <= nz + 1 as obstaculo clamps them."""
    rng = np.random.default_rng(seed)
    def bump(nxe, nye, top):
        m = np.zeros((nxe, nye))
        m[nxe // 3: 2 * nxe // 3 + 1, 2: nye - 2] = rng.integers(1, top + 1, size=(2 * nxe // 3 + 1 - nxe // 3, nye - 4))
        return m
    ms = [bump(p.nx + 3, p.ny + 2, 2), bump(p.nx + 2, p.ny + 3, 2), bump(p.nx + 2, p.ny + 2, 3)]
    if neg:
        # obstaculo's columns below the bottom (3_contorno.f90:741-760 gives
        # -1 where the bed lies under z = 0): ku = kv = -1 and kw = 0 on the
        # row after the bump, ku = kv = kw = -1 on the row before it
        for q, m in enumerate(ms):
            n0, n1 = m.shape
            m[2 * n0 // 3 + 1, 2: n1 - 2] = 0 if q == 2 else -1
            m[n0 // 3 - 1, 2: n1 - 2] = -1
    return tuple(ms)


def obstacle(neg=False):
    import dataclasses
    p = dataclasses.replace(O.Params(), obst_t=1)
    O.ref_build(p)
    cmds = [("contorno", 1), ("contorno", 2), ("contorno", 3), ("stage", 1), ("step", 2)]
    case = "g8obstneg" if neg else "g8obst"
    for q, (cmd, arg) in enumerate(cmds):
        s_in = base_state(p, seed=(750 if neg else 700) + q)
        s_in.ku[...], s_in.kv[...], s_in.kw[...] = obstacle_masks(p, seed=q, neg=neg)
        s_out = s_in.copy()
        O.ref_run(s_out, cmd, arg)
        tag = f"{cmd}{arg}"
        changed = _save(os.path.join(HERE, f"{case}_{tag}.npz"), p, s_in, s_out)
        print(case, tag, "changed:", changed)


# grid of the obstacle-geometry vectors: 3 m x 0.6 m x 0.6 m
GEOM = dict(nx=60, ny=6, nz=12, dx=0.05, dy=0.1, dz=0.05)


def geometry():
    import dataclasses
    for t in (1, 2, 3, 4, 5, 6, 7, 10):
        p = dataclasses.replace(O.Params(**GEOM), obst_t=t)
        O.ref_build(p)
        s = O.State(p)
        O.ref_run(s, "obstaculo")
        np.savez_compressed(os.path.join(HERE, "obstaculo", f"obstaculo_{t}.npz"), ku=s.ku, kv=s.kv, kw=s.kw,
                            grid=np.array([p.nx, p.ny, p.nz]), h=np.array([p.dx, p.dy, p.dz]))
        print("obst_t", t, "ku range", s.ku.min(), s.ku.max(), "kw range", s.kw.min(), s.kw.max())


def est():
    """est (4_plot.f90:500-525) on three states per grid: seeded velocities,
    the same after a reference projection (graddin + posdin, divergence at
    the solver's residual level), and one face velocity raised at a wall
    cell.  Stored: the velocities, the divergence field and its maximum."""
    os.makedirs(os.path.join(HERE, "est"), exist_ok=True)
    for case, p in CASES.items():
        O.ref_build(p)
        for q in range(3):
            s = base_state(p, seed=900 + q)
            if q == 1:
                s.run("contorno", 2)
                O.ref_run(s, "graddin")
                O.ref_run(s, "posdin")
            if q == 2:
                s.u[1, 1, 1] += 7.0
            out = s.copy()
            O.ref_run(out, "est")
            np.savez_compressed(os.path.join(HERE, "est", f"est_{case}_{q}.npz"), u=s.u, v=s.v, w=s.w,
                                div=out.kurv, divmax=np.array([out.scal["vol_ins"]]),
                                grid=np.array([p.nx, p.ny, p.nz]), h=np.array([p.dx, p.dy, p.dz]))
            print("est", case, q, "max |div|", out.scal["vol_ins"])


def periodic():
    """graddin on the x-periodic grid: the reference's operator is not
    symmetric across the periodic faces (the face density is copied at both
    ends, 8_derivadas.f90:318-323), its CG does not reach the stopping test
    and runs the full 10000 iterations (9_grad.f90:175); the fixture pins the
    oracle's recurrence over all of them.  Same input as main() would build."""
    p = CASES["per"]
    O.ref_build(p)
    q = COMMANDS.index(("graddin", None))
    s_in = base_state(p, seed=100 + q)
    s_in.run("contorno", 2)
    s_out = s_in.copy()
    O.ref_run(s_out, "graddin")
    changed = _save(os.path.join(HERE, "per_graddin.npz"), p, s_in, s_out)
    print("per graddin changed:", changed)


# the manufactured-solution cases: (mms_t, a, coef); grid 12x10x8 of
# 0.5 x 0.6 x 0.3, every wall prescribed (code 3), no turbulence model, no
# gravity (the mode has no projection to balance it)
WENG_CASES = {"weng1": (1, 0.0, 0.0), "weng2": (2, 0.0, 1.0), "weng1a": (1, 0.5, 0.0), "weng2a": (2, 0.5, 1.0)}
WENG_DT, WENG_STEPS, WENG_T_SRC = 0.01, 5, 0.7


def weng_params(mms_t, a, coef):
    return O.Params(nx=12, ny=10, nz=8, dx=0.5, dy=0.6, dz=0.3, ccx0=3, ccxf=3, ccy0=3, ccyf=3, ccz0=3, cczf=3,
                    m_turb=0, gx=0.0, gz=0.0, mms_t=mms_t, mms_a=a, mms_h0=2.0, mms_coef=coef)


def weng():
    """The manufactured-solution mode of reference builds with mms_t = 1, 2
    (13_mms.f90): the run of 1_principal.f90:39-108 from t = 0 for WENG_STEPS
    steps (harness mms_run: final fields, the errors of mms() at t = 0 and
    after every step) and, at t = WENG_T_SRC (the steady planes at t = 0),
    mms_bc's planes and the sources of termo_fonte1/termo_fonte2 (harness
    mms_src).  a = 0 is the reference's own setting; with a = 0.5 the
    reference's hand-derived sources differ from the exact ones, so those
    cases store the initial state (mms_run with 0 steps), the planes and the
    reference's sources only."""
    os.makedirs(os.path.join(HERE, "weng"), exist_ok=True)
    for case, (mms_t, a, coef) in WENG_CASES.items():
        p = weng_params(mms_t, a, coef)
        O.ref_build(p)
        s = O.State(p)
        s.scal["dt"] = WENG_DT
        nsteps = WENG_STEPS if a == 0.0 else 0
        out, ex = O.ref_run_mms(s, "mms_run", nsteps)
        errs = np.array([[float(x) for x in l.split()[2:]] for l in out.splitlines() if l.startswith("MMS_ERR")])
        src = O.State(p)
        src.scal["dt"] = WENG_DT
        src.scal["t"] = 0.0 if mms_t == 1 else WENG_T_SRC
        _, exs = O.ref_run_mms(src, "mms_src")
        data = {"mms": np.array([mms_t, a, 2.0, coef]), "dt": np.array([WENG_DT]), "nsteps": np.array([nsteps]),
                "t_src": np.array([src.scal["t"]]), "errors": errs,
                "grid": np.array([p.nx, p.ny, p.nz]), "h": np.array([p.dx, p.dy, p.dz])}
        for n in ("u", "v", "w", "ls", "prd1"):
            data["out_" + n] = s.arr(n)
        for n in O.PLANE_NAMES:
            data["plane_" + n] = exs[n]
        for n in ("tf_u", "tf_v", "tf_w"):
            data["src_" + n] = exs[n]
        np.savez_compressed(os.path.join(HERE, "weng", f"{case}.npz"), **data)
        print(case, "errors after the last step:", errs[-1])


WAVE_STEPS = 3


def wave_params(wave_t):
    """A 4 x 1 x 1.2 m channel at 0.25 x 0.25 x 0.1 m cells, still water 1 m
    deep, the wave entering through the prescribed wall x = 0 (code 3),
    outflow at x = 4 (code 4), free-slip sides and lid, no-slip bed."""
    return O.Params(nx=16, ny=4, nz=12, dx=0.25, dy=0.25, dz=0.1, ccx0=3, ccxf=4, ccy0=1, ccyf=1, ccz0=2, cczf=1,
                    m_turb=1, gx=0.0, gz=9.81, wave_t=wave_t)


def wave_still_water(p):
    s = O.State(p)
    z = (np.arange(1, p.nz + 1) - 0.5) * p.dz
    s.ls[...] = 1.0 - z[None, None, :]
    s.scal["dt"] = 0.01
    s.run("ls_ini")
    s.run("contorno", 1)
    s.run("contorno", 3)
    return s


def waves():
    """The wave inflow of reference builds with wave_t = 1, 2, 5
    (3_contorno.f90:952-1205): waves_coef + one boundary_waves at t = 0.7 on
    a still-water level set with seeded noise (harness waves: coefficients,
    level set, planes), and WAVE_STEPS steps of the time loop from still
    water (harness wave_run)."""
    os.makedirs(os.path.join(HERE, "waves"), exist_ok=True)
    for wt in (1, 2, 5):
        p = wave_params(wt)
        O.ref_build(p)
        s = O.State(p)
        z = (np.arange(1, p.nz + 1) - 0.5) * p.dz
        s.ls[...] = 1.0 - z[None, None, :] + 1e-3 * np.random.default_rng(wt).standard_normal(s.ls.shape)
        s.scal["dt"] = 0.01
        s.scal["t"] = 0.7
        ls_in = s.ls.copy()
        out, ex = O.ref_run_mms(s, "waves")
        coef = np.array([float(x) for x in out.split("WAVE_COEF")[1].split()[:14]])
        data = {"coef": coef, "ls_in": ls_in, "ls_out": s.ls.copy(), "t": np.array([0.7]), "steps": np.array([WAVE_STEPS])}
        for n in ("bxx0", "bxx1", "bxy0", "bxz0"):
            data["plane_" + n] = ex[n]
        r = wave_still_water(p)
        _, exr = O.ref_run_mms(r, "wave_run", WAVE_STEPS)
        for n in ("u", "v", "w", "prd1", "ls"):
            data["run_" + n] = r.arr(n)
        for n in ("bxx0", "bxx1", "bxz0"):
            data["run_plane_" + n] = exr[n]
        np.savez_compressed(os.path.join(HERE, "waves", f"wave{wt}.npz"), **data)
        print("wave", wt, "l_w", coef[0], "max u after the run", np.abs(r.u).max())


def main():
    for case, p in CASES.items():
        O.ref_build(p)
        for q, (cmd, arg) in enumerate(COMMANDS):
            if case == "per" and cmd in ("graddin", "stage", "step"):
                continue  # periodic-x CG does not converge in the reference (asymmetric wall faces)
            s_in = base_state(p, seed=100 + q)
            rng = np.random.default_rng(q)
            if cmd in ("ls_ini", "heaviside", "mod_ls1"):
                for n in ["hs", "hsx", "hsy", "hsz", "mod_ls", "kurv", "ddlsdx", "ddlsdy", "ddlsdz", "rho", "ls_nu"]:
                    s_in.arr(n)[...] = 0.0
                s_in.rho[...] = 1.0
                s_in.ls[...] += 0.01 * p.dx * rng.standard_normal(s_in.ls.shape)
            if cmd == "convdiff":
                for n in ["Fu", "Fv", "Fw"]:
                    s_in.arr(n)[...] = 0.0
            if cmd == "contorno":
                for n in ["u", "v", "w", "prd1"]:  # scramble every ghost plane
                    a = s_in.arr(n)
                    g = rng.standard_normal(a.shape)
                    a[0], a[-1] = g[0], g[-1]
                    a[:, 0], a[:, -1] = g[:, 0], g[:, -1]
                    a[:, :, 0], a[:, :, -1] = g[:, :, 0], g[:, :, -1]
            if cmd in ("graddin",):
                s_in.run("contorno", 2)
            s_out = s_in.copy()
            O.ref_run(s_out, cmd, arg)
            data = {"__scal_in": np.array([s_in.scal[k] for k in O.SCALARS]),
                    "__scal_out": np.array([s_out.scal[k] for k in O.SCALARS]),
                    "__params": np.array([p.nx, p.ny, p.nz, p.ccx0, p.ccxf, p.ccy0, p.ccyf, p.ccz0, p.cczf]),
                    "__h": np.array([p.dx, p.dy, p.dz])}
            for name, _ in O.ARRAYS:
                a, b = s_in.arr(name), s_out.arr(name)
                data["in_" + name] = a
                if not np.array_equal(a, b):
                    data["out_" + name] = b
            tag = cmd + ("" if arg is None else str(arg))
            np.savez_compressed(os.path.join(HERE, f"{case}_{tag}.npz"), **data)
            print(case, tag, "changed:", sorted(k[4:] for k in data if k.startswith("out_")))


if __name__ == "__main__":
    if sys.argv[1:] == ["schemes"]:
        schemes()
    elif sys.argv[1:] == ["variants"]:
        variants()
    elif sys.argv[1:] == ["obstacle"]:
        obstacle()
    elif sys.argv[1:] == ["obstacle_neg"]:
        obstacle(neg=True)
    elif sys.argv[1:] == ["geometry"]:
        geometry()
    elif sys.argv[1:] == ["est"]:
        est()
    elif sys.argv[1:] == ["periodic"]:
        periodic()
    elif sys.argv[1:] == ["weng"]:
        weng()
    elif sys.argv[1:] == ["waves"]:
        waves()
    else:
        main()
