"""GPU: the wave inflow (3_contorno.f90:952-1205) through the C ABI against
the reference's fixtures (tests/golden/waves) and the oracle, which matches
them bit for bit (tests/test_waves.py)."""
import os
import sys

import numpy as np
import pytest

import oracle as O
import pysuli as S
from helpers import gpu_from_state, rel_err

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
from gen_golden import WAVE_STEPS, wave_params, wave_still_water  # noqa: E402

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden", "waves")


def _fixture(wt):
    return np.load(os.path.join(GOLD, f"wave{wt}.npz"))


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("wt", [1, 2, 5])
def test_gpu_boundary_waves(wt, prec):
    """suli_waves_coef's coefficients and one suli_boundary_waves at t = 0.7:
    the planes bxx0, bxx1, bxy0, bxz0 and the level set (single-precision
    rounding, first column from the wave) against the reference."""
    z = _fixture(wt)
    p = wave_params(wt)
    fp64 = prec == "fp64"
    print(f"waves wave_t {wt} {prec}", flush=True)
    sv = S.Solver.from_oracle_params(p, precision=S.PREC_FP64 if fp64 else S.PREC_MIXED)
    try:
        coef = sv.waves_coef(wt)
        assert np.allclose(coef, z["coef"], rtol=1e-13, atol=0.0)
        sv.upload("ls", z["ls_in"])
        sv.set_time(float(z["t"][0]))
        sv.boundary_waves()
        tol = 1e-13 if fp64 else 2e-7
        for n in ("bxx0", "bxx1", "bxy0", "bxz0"):
            ref = z["plane_" + n]
            d = np.abs(sv.download(n) - ref).max()
            print(f"  {n}: max diff {d:.2e} (max {np.abs(ref).max():.2e})")
            assert d <= tol * max(np.abs(ref).max(), 1e-300), n
        ls = sv.download("ls")
        if fp64:
            assert np.array_equal(ls, z["ls_out"])
        else:
            assert np.abs(ls - z["ls_out"]).max() <= 1e-7
    finally:
        sv.close()


@pytest.mark.parametrize("wt", [1, 2, 5])
def test_gpu_wave_loop_with_shared_pressure(wt):
    """The time loop with the wave entering at x = 0, routine by routine
    against the oracle (FP64): level_set, contorno(3), visco, convdiff, tempo,
    boundary_waves, contorno(2), then the projection, posdin and contorno(1)
    with the GPU's pressure handed to the oracle.  The pressure itself is not
    compared: with a prescribed inflow and a zero-gradient outflow the net
    flux through the walls is not zero, the reference's Neumann projection
    is an inconsistent singular system, and its CG's answer is set by
    round-off (tests/test_waves.py: a CG with another summation order runs
    10000 iterations and ends far away).  The GPU solves it with the FCG,
    which removes the inconsistent (null-space) part of the residual."""
    p = wave_params(wt)
    st = wave_still_water(p)
    dt = float(st.scal["dt"])
    sv = gpu_from_state(st, precision=S.PREC_FP64, poisson=S.POISSON_FCG, precond=S.PRECOND_GMG, tol_rel=1e-12)
    F = ("u", "v", "w", "prd1", "ls", "rho", "hs", "kurv", "nut")

    def cmp(label, names=F, tol=1e-10):
        # velocities against the largest speed (v is round-off around zero
        # in this spanwise-uniform channel), the other fields against their own maxima
        vmax = max(max(np.abs(st.arr(n)).max() for n in ("u", "v", "w")), 1e-300)  # still water at the start
        errs = {n: (np.abs(sv.download(n) - st.arr(n)).max() / vmax if n in ("u", "v", "w")
                    else rel_err(sv.download(n), st.arr(n))) for n in names}
        bad = {n: e for n, e in errs.items() if e > tol}
        assert not bad, (label, bad)

    try:
        sv.waves_coef(wt)
        st.run("waves_coef")
        for it in range(1, WAVE_STEPS + 1):
            t = it * dt
            sv.set_time(t)
            st.scal["t"] = t
            sv.level_set(dt)
            st.run("level_set")
            sv.contorno(3)
            st.run("contorno", 3)
            cmp(f"step {it} level set")
            for tt in (1, 2):
                d = 0.5 * dt if tt == 1 else dt
                st.scal["dt"] = d
                sv.visco()
                st.run("visco")
                sv.convdiff()
                st.run("convdiff")
                sv.tempo(tt, d)
                st.run("tempo", tt)
                sv.boundary_waves()
                st.run("boundary_waves")
                sv.contorno(2)
                st.run("contorno", 2)
                cmp(f"step {it} stage {tt} predictor")
                for n in ("bxx0", "bxx1", "bxz0"):
                    assert rel_err(sv.download(n), st.planes[n]) <= 1e-13, n
                ps = sv.graddin(d)
                assert ps["converged"], ps
                st.prd1[...] = sv.download("prd1")
                sv.posdin(d)
                st.run("posdin")
                sv.contorno(1)
                st.run("contorno", 1)
                cmp(f"step {it} stage {tt} corrector")
                st.scal["dt"] = dt
        print(f"wave loop {wt}: max |u| {np.abs(st.u).max():.3e}, {WAVE_STEPS} steps matched")
    finally:
        sv.close()


def test_gpu_wave_run_mixed_and_checkpoint(tmp_path):
    """Mixed precision (FP32 fields, FCG-multigrid projection): the Stokes V
    channel runs with the wave entering (planes equal to the FP64 oracle's
    to single precision, finite fields), and a checkpoint carries the wave:
    the run resumed in a fresh context continues bit for bit."""
    p = wave_params(5)
    st = wave_still_water(p)
    dt = float(st.scal["dt"])
    kw = dict(precision=S.PREC_MIXED, poisson=S.POISSON_FCG, precond=S.PRECOND_GMG, tol_rel=1e-8)
    sv = gpu_from_state(st, **kw)
    path = str(tmp_path / "wave.ckpt")
    st.run("waves_coef")
    try:
        sv.waves_coef(5)
        for it in range(1, WAVE_STEPS + 1):
            _, ps = sv.step(dt)
            assert all(q["converged"] for q in ps), ps
        st.scal["t"] = WAVE_STEPS * dt
        st.run("boundary_waves")
        for n in ("bxx0", "bxx1", "bxz0"):
            assert rel_err(sv.download(n), st.planes[n]) <= 1e-6, n
        for n in ("u", "v", "w", "ls"):
            assert np.all(np.isfinite(sv.download(n))), n
        sv.checkpoint_save(path)
        sv.step(dt)
        want = {n: sv.download(n) for n in ("u", "v", "w", "ls", "bxx0")}
    finally:
        sv.close()
    sv = S.Solver.from_oracle_params(p, **kw)
    try:
        sv.checkpoint_load(path)
        sv.step(dt)
        for n, a in want.items():
            assert np.array_equal(sv.download(n), a), n
    finally:
        sv.close()
