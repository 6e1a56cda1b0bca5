"""GPU parity of the time schemes of tempo (11_diversos.f90:10-228) other
than RK2: explicit Euler, RK3 and AB2 (t_tempo 0, 2, 3 of 0_modules.f90:24),
against the CPU oracle, which reproduces the flang-built reference of each
scheme bit for bit (tests/golden/g8tt*_*.npz).

FP64 tolerances as for RK2: 1e-12 relative for one tempo call (the forces
differ from the oracle's by the GPU's sqrt rounding only); whole steps to
the reference CG's stopping point, 2e-5 (1e-4 if a solve stops one CG
iteration apart).  Mixed precision: 4e-4 (10x the largest measured, 3.5e-5).
"""
import dataclasses

import numpy as np
import pytest

import oracle as O
import pysuli
from helpers import VEL, consistent_state, gpu_from_state, rel_err

pytestmark = pytest.mark.gpu

KEPT = ["fu0", "fv0", "fw0", "fu1", "fv1", "fw1"]
NTT = {0: 1, 1: 2, 2: 3, 3: 1}
G16 = O.Params(nx=16, ny=12, nz=10, dx=1 / 16, dy=1 / 16, dz=1 / 16)
PER = O.Params(nx=12, ny=8, nz=8, dx=1 / 12, dy=1 / 12, dz=1 / 12, ccx0=0, ccxf=0, cczf=1)


def _state(p, scheme, seed):
    st = consistent_state(dataclasses.replace(p, t_tempo=scheme), seed=seed)
    rng = np.random.default_rng(seed)
    st.u0[...] = st.u[1:-1, 1:-1, 1:-1] * 0.9
    st.v0[...] = st.v[1:-1, 1:-1, 1:-1] * 0.9
    st.w0[...] = st.w[1:-1, 1:-1, 1:-1] * 0.9
    for n in KEPT:
        st.arr(n)[...] = 0.1 * rng.standard_normal(st.arr(n).shape)
    return st


def _cmp(sv, st, names, tol, label):
    errs = {n: rel_err(sv.download(n), st.arr(n)) for n in names}
    bad = {k: v for k, v in errs.items() if not v <= tol}
    assert not bad, f"{label}: {bad} (tol {tol})"
    return errs


@pytest.mark.parametrize("p", [G16, PER], ids=["g16", "per"])
@pytest.mark.parametrize("scheme,tt", [(0, 1), (2, 1), (2, 2), (2, 3), (3, 1)])
def test_tempo_scheme(p, scheme, tt):
    print(f"tempo scheme {scheme} stage {tt} grid {p.nx}x{p.ny}x{p.nz}", flush=True)
    st = _state(p, scheme, seed=40 + 3 * scheme + tt)
    sv = gpu_from_state(st)
    try:
        sv.convdiff()
        sv.tempo(tt, 1e-3)
        st.run("convdiff")
        st.scal["dt"] = 1e-3
        st.run("tempo", tt)
        _cmp(sv, st, VEL + ["u0", "v0", "w0"] + KEPT, 1e-12, f"scheme {scheme} tempo {tt}")
    finally:
        sv.close()


@pytest.mark.parametrize("p", [G16, PER], ids=["g16", "per"])
def test_ab2_second_step(p):
    """The AB2 update proper (it > 1): the first tempo after selecting the
    scheme is Euler and keeps F in fu0, the second uses it."""
    print(f"AB2 two updates grid {p.nx}x{p.ny}x{p.nz}", flush=True)
    st = _state(p, 3, seed=51)
    sv = gpu_from_state(st)
    try:
        st.scal["dt"] = 1e-3
        for it in (1, 2):
            if it == 2:  # as between the driver's stages: ghosts and friction planes follow u
                sv.contorno(1)
                st.run("contorno", 1)
            sv.convdiff()
            sv.tempo(1, 1e-3)
            st.run("convdiff")
            st.run("tempo", 1, it=it)
            _cmp(sv, st, VEL + KEPT[:3], 1e-12, f"AB2 update {it}")
    finally:
        sv.close()


@pytest.mark.parametrize("scheme", [0, 2, 3])
def test_step_scheme_fp64(scheme):
    """Three dam-break steps of the scheme through suli_step, FP64, with the
    reference's CG: one Poisson statistics entry per stage."""
    print(f"3 steps FP64 scheme {scheme}", flush=True)
    p = O.Params(nx=16, ny=8, nz=8, dx=1 / 16, dy=1 / 16, dz=1 / 16, t_tempo=scheme)
    st = O.State(p)
    st.ls[...] = O.dambreak_phi(p, 0.25, 0.25)
    st.scal["dt"] = 1e-3
    st.run("ls_ini")
    st.run("contorno", 1)
    sv = gpu_from_state(st)
    try:
        assert sv.time_stages() == NTT[scheme]
        same = True
        for it in range(1, 4):
            ls, ps = sv.step(1e-3)
            assert len(ps) == NTT[scheme]
            st.run("level_set")
            st.run("contorno", 3)
            for tt in range(1, NTT[scheme] + 1):
                st.run("stage", tt, it=it)
                same &= ps[tt - 1]["iters"] == st.diag["cg_iters"]
        tol = 2e-5 if same else 1e-4
        vmax = max(float(np.max(np.abs(st.arr(n)))) for n in VEL)
        ev = max(float(np.max(np.abs(sv.download(n) - st.arr(n)))) for n in VEL) / vmax
        errs = _cmp(sv, st, ["ls", "hs", "rho"], 1e-8, f"scheme {scheme} level set")
        errs.update(_cmp(sv, st, ["prd1"], tol, f"scheme {scheme} pressure"))
        print(f"scheme {scheme}: velocity {ev:.2e}", errs)
        assert ev <= tol, ev
        if scheme in (2, 3):
            _cmp(sv, st, KEPT[:3], tol, f"scheme {scheme} kept forces")
    finally:
        sv.close()


@pytest.mark.parametrize("scheme", [0, 2, 3])
def test_step_scheme_mixed_fcg(scheme):
    """Mixed precision with the FCG projection (multigrid, extrapolated
    initial guess) under each scheme stays within 1e-3 of the FP64 oracle
    over three steps, and each solve reaches its 1e-8 target."""
    print(f"3 steps mixed FCG scheme {scheme}", flush=True)
    p = O.Params(nx=16, ny=8, nz=8, dx=1 / 16, dy=1 / 16, dz=1 / 16, t_tempo=scheme)
    st = O.State(p)
    st.ls[...] = O.dambreak_phi(p, 0.25, 0.25)
    st.scal["dt"] = 1e-3
    st.run("ls_ini")
    st.run("contorno", 1)
    sv = gpu_from_state(st, precision=pysuli.PREC_MIXED, poisson=pysuli.POISSON_FCG,
                        precond=pysuli.PRECOND_GMG, tol_rel=1e-8)
    try:
        for _ in range(3):
            _, ps = sv.step(1e-3)
            assert all(q["converged"] for q in ps), ps
        st.run("step", 3)
        vmax = max(float(np.max(np.abs(st.arr(n)))) for n in VEL)
        ev = max(float(np.max(np.abs(sv.download(n) - st.arr(n)))) for n in VEL) / vmax
        print(f"mixed-precision velocity error {ev:.2e}", flush=True)
        assert ev <= 4e-4, ev  # measured 1e-5 to 3.5e-5: about 10x
        _cmp(sv, st, ["ls", "rho"], 1e-5, f"mixed scheme {scheme}")
    finally:
        sv.close()


def test_checkpoint_ab2_resumes(tmp_path):
    """A checkpoint keeps the scheme and the AB2 history (fu0 and whether
    the Euler start is done): a run resumed from it continues bit for bit,
    in a context that had not selected the scheme."""
    print("AB2 checkpoint resume 32x16x16", flush=True)
    p = O.Params(nx=32, ny=16, nz=16, dx=1 / 32, dy=1 / 32, dz=1 / 32)
    path = str(tmp_path / "ab2.ckpt")
    names = ["u", "v", "w", "prd1", "ls", "fu0", "fv0", "fw0"]

    def tank():
        return pysuli.Solver.from_oracle_params(p, poisson=pysuli.POISSON_FCG, precond=pysuli.PRECOND_GMG,
                                                tol_rel=1e-8)
    sv = tank()
    try:
        sv.set_time_scheme(pysuli.TIME_AB2)
        sv.init_dambreak(0.3, 0.3)
        sv.contorno(1)
        for _ in range(2):
            sv.step(1e-3)
        sv.checkpoint_save(path)
        for _ in range(2):
            sv.step(1e-3)
        want = {n: sv.download(n) for n in names}
    finally:
        sv.close()
    sv = tank()
    try:
        sv.checkpoint_load(path)
        assert sv.time_stages() == 1
        for _ in range(2):
            sv.step(1e-3)
        for n in names:
            assert np.array_equal(sv.download(n), want[n]), n
    finally:
        sv.close()


def test_scheme_api_errors():
    p = O.Params()
    sv = pysuli.Solver.from_oracle_params(p)
    try:
        with pytest.raises(pysuli.SuliError, match="t_tempo"):
            sv.set_time_scheme(4)
        sv.set_time_scheme(pysuli.TIME_EULER)
        sv.convdiff()
        with pytest.raises(pysuli.SuliError, match="outside 1..1"):
            sv.tempo(2, 1e-3)
    finally:
        sv.close()
