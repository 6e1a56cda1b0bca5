"""GPU parity: every hot-path routine of the HIP library against the CPU
oracle (itself bit-exact with the flang-built reference) on the same inputs.

FP64 path: per-routine tolerance 1e-12 relative to the field's max norm
(the kernels evaluate the reference's expressions in its order without FMA;
the remaining differences come from the GPU's sin/cos/sqrt rounding and the
order of the FP64 reductions).  Mixed path (FP32 stencils, three steps):
level set 3e-6, pressure 1e-5, velocity 1e-4 of the largest speed.
"""
import numpy as np
import pytest

import oracle as O
from helpers import CELL, VEL, consistent_state, gpu_from_state, rel_err

pytestmark = pytest.mark.gpu

GRIDS = {
    "g8": O.Params(),
    "g16": O.Params(nx=16, ny=12, nz=10, dx=1 / 16, dy=1 / 16, dz=1 / 16),
    "per": O.Params(nx=12, ny=8, nz=8, dx=1 / 12, dy=1 / 12, dz=1 / 12, ccx0=0, ccxf=0, cczf=1),
}


def _cmp(sv, st, names, tol, label):
    worst = {}
    for n in names:
        e = rel_err(sv.download(n), st.arr(n))
        worst[n] = e
    bad = {k: v for k, v in worst.items() if not v <= tol}
    assert not bad, f"{label}: {bad} (tol {tol})"
    return worst


@pytest.fixture(scope="module", params=sorted(GRIDS))
def grid(request):
    return GRIDS[request.param]


def test_level_set_ini(grid):
    st = O.random_state(grid, seed=2)
    sv = gpu_from_state(st)
    vol = sv.level_set_ini()
    st.run("ls_ini")
    _cmp(sv, st, CELL, 1e-12, "ls_ini")
    assert abs(vol - st.scal["vol_ini"]) <= 1e-12 * st.scal["vol_ini"]
    sv.close()


def test_level_set(grid):
    st = consistent_state(grid, seed=3, perturb=0.4)
    sv = gpu_from_state(st)
    sv.level_set_ini()
    st.run("ls_ini")
    info = sv.level_set(1e-2)
    st.scal["dt"] = 1e-2
    st.run("level_set")
    assert info["reinit_iters"] == st.diag["reinit_iters"]
    assert info["mass_corr_iters"] == st.diag["mass_corr_iters"]
    _cmp(sv, st, CELL, 1e-11, "level_set")
    sv.close()


def test_visco(grid):
    st = consistent_state(grid, seed=4)
    sv = gpu_from_state(st)
    sv.visco()
    st.run("visco")
    _cmp(sv, st, ["nut", "xnut", "ynut", "znut"], 1e-12, "visco")
    sv.close()


def test_convdiff_forces(grid):
    st = consistent_state(grid, seed=5)
    sv = gpu_from_state(st)
    sv.convdiff_forces()
    st.run("convdiff")
    _cmp(sv, st, ["Fu", "Fv", "Fw"], 1e-12, "convdiff")
    sv.close()


@pytest.mark.parametrize("tt", [1, 2])
def test_tempo(grid, tt):
    st = consistent_state(grid, seed=6)
    st.u0[...] = st.u[1:-1, 1:-1, 1:-1] * 0.9  # stage-2 input
    st.v0[...] = st.v[1:-1, 1:-1, 1:-1] * 0.9
    st.w0[...] = st.w[1:-1, 1:-1, 1:-1] * 0.9
    sv = gpu_from_state(st)
    sv.convdiff()
    sv.tempo(tt, 1e-3)
    st.run("convdiff")
    st.scal["dt"] = 1e-3
    st.run("tempo", tt)
    _cmp(sv, st, VEL + ["u0", "v0", "w0"], 1e-12, f"tempo{tt}")
    sv.close()


@pytest.mark.parametrize("nlock", [1, 2])
def test_contorno(grid, nlock):
    st = O.random_state(grid, seed=7)
    sv = gpu_from_state(st)
    sv.contorno(nlock)
    st.run("contorno", nlock)
    _cmp(sv, st, VEL + ["prd1"], 0.0, f"contorno{nlock}")
    sv.close()


def test_graddin(grid):
    st = consistent_state(grid, seed=8)
    st.run("contorno", 2)
    sv = gpu_from_state(st)
    st.scal["dt"] = 1e-3
    if grid.ccx0 == 0:
        # The reference's periodic-x operator is not symmetric (the face
        # density is copied at both ends, 8_derivadas.f90:318-323), so its CG
        # never meets the stopping test and runs 10000 iterations, whose
        # iterate is round-off dominated.  Both solvers follow the same
        # recurrence for a fixed 20 iterations instead.
        import pysuli
        sv.set_solver(pysuli.POISSON_REF, pysuli.PRECOND_JACOBI, 0.0, 20)
        st.cg_max_it = 20
        info = sv.graddin(1e-3)
        st.run("graddin")
        assert info["iters"] == st.diag["cg_iters"] == 20, (info, st.diag)
        errs = _cmp(sv, st, ["prd1"], 1e-10, "graddin, periodic x, 20 iterations")
        print("periodic graddin, 20 iterations: relative error", errs)
        sv.close()
        return
    info = sv.graddin(1e-3)
    st.run("graddin")
    assert abs(info["iters"] - st.diag["cg_iters"]) <= 1, (info, st.diag)
    tol = 1e-10 if info["iters"] == st.diag["cg_iters"] else 1e-4
    _cmp(sv, st, ["prd1"], tol, "graddin")
    sv.close()


def test_posdin(grid):
    st = consistent_state(grid, seed=9)
    sv = gpu_from_state(st)
    sv.posdin(1e-3)
    st.scal["dt"] = 1e-3
    st.run("posdin")
    _cmp(sv, st, VEL, 1e-13, "posdin")
    sv.close()


def test_est_div(grid):
    st = consistent_state(grid, seed=10)
    sv = gpu_from_state(st)
    assert abs(sv.est_div() - st.est_div()) <= 1e-13 * st.est_div()
    sv.close()


def _vel_err(sv, st):
    """velocity error normalised by the largest speed of the three components"""
    vmax = max(float(np.max(np.abs(st.arr(n)))) for n in VEL)
    return max(float(np.max(np.abs(sv.download(n) - st.arr(n)))) for n in VEL) / vmax


def test_step_fp64():
    """Three dam-break steps, FP64: level-set fields to 1e-8; pressure,
    velocity and eddy viscosity to 2e-5 when every Poisson solve stops after
    the same number of CG iterations as the reference's (1e-4 if one stops an
    iteration apart at the absolute threshold sum r^2 <= 1e-4/N)."""
    p = O.Params(nx=16, ny=8, nz=8, dx=1 / 16, dy=1 / 16, dz=1 / 16)
    st = O.State(p)
    st.ls[...] = O.dambreak_phi(p, 0.25, 0.25)
    st.scal["dt"] = 1e-3
    st.run("ls_ini")
    st.run("contorno", 1)
    sv = gpu_from_state(st)
    same = True
    for _ in range(3):
        ls, ps = sv.step(1e-3)
        st.run("level_set")
        st.run("contorno", 3)
        for tt in (1, 2):
            st.run("stage", tt)
            same &= ps[tt - 1]["iters"] == st.diag["cg_iters"]
            print("solve", tt, "iterations gpu", ps[tt - 1]["iters"], "oracle", st.diag["cg_iters"])
    # The reference CG's iterate at its loose absolute stopping point is
    # sensitive to the order of its FP64 sums: two CPU CGs that differ only in
    # summation order stop after the same iterations with pressures 6e-7
    # apart (tests/test_oracle.py::test_cg_summation_order_sensitivity).
    tol_flow = 2e-5 if same else 1e-4
    ev = _vel_err(sv, st)
    errs = _cmp(sv, st, ["ls", "hs", "rho", "kurv"], 1e-8, "3 steps (level set)")
    errs.update(_cmp(sv, st, ["prd1", "nut"], tol_flow, "3 steps (flow)"))
    print("3-step FP64 relative errors:", errs, "velocity", ev)
    assert ev <= tol_flow, ev
    sv.close()


def test_mixed_precision_step():
    p = O.Params(nx=16, ny=8, nz=8, dx=1 / 16, dy=1 / 16, dz=1 / 16)
    st = O.State(p)
    st.ls[...] = O.dambreak_phi(p, 0.25, 0.25)
    st.scal["dt"] = 1e-3
    st.run("ls_ini")
    st.run("contorno", 1)
    import pysuli
    sv = gpu_from_state(st, precision=pysuli.PREC_MIXED)
    for _ in range(3):
        sv.step(1e-3)
        st.run("step", 1)
    # measured: level set 2.5e-7, pressure 1.1e-6, velocity 8.2e-6 (of the
    # largest speed); the bars are about 10x those (the pressure and the
    # velocity differ from the oracle mostly by where the reference CG stops)
    errs = _cmp(sv, st, ["ls", "hs", "rho"], 3e-6, "mixed 3 steps (level set)")
    errs.update(_cmp(sv, st, ["prd1"], 1e-5, "mixed 3 steps (pressure)"))
    errs["velocity"] = _vel_err(sv, st)
    print("mixed-precision relative errors after 3 steps:", errs)
    assert errs["velocity"] <= 1e-4, errs
    sv.close()


def test_sloshing_forcing_matches_constant_gx(grid):
    """gx(t) = amp*sin(2 pi t/period + phase) at a time where the sine is 1
    must give the forces of the reference with gx = amp (6_convdiff.f90:155)."""
    import dataclasses
    amp, period = 0.05 * 9.81, 1.2
    st = consistent_state(dataclasses.replace(grid, gx=amp), seed=5)
    sv = gpu_from_state(st, gx=0.0)
    sv.set_forcing(amp, period, 0.0)
    sv.set_time(0.25 * period)
    sv.convdiff_forces()
    st.run("convdiff")
    _cmp(sv, st, ["Fu", "Fv", "Fw"], 1e-12, "forcing")
    sv.set_time(0.75 * period)  # sine -1: the forces move by -2 amp in x only
    sv.convdiff_forces()
    Fu = sv.download("Fu")
    assert np.allclose(Fu, st.arr("Fu") - 2 * amp, rtol=0, atol=1e-12 * max(1.0, np.abs(Fu).max()))
    assert np.array_equal(sv.download("Fw"), st.arr("Fw"))
    sv.close()
