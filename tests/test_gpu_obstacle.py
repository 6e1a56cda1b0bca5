"""GPU parity of the bottom obstacle (ku, kv, kw columns, obst_t != 0 in
0_modules.f90:29) against the CPU oracle, itself bit-exact with a reference
build with obst_t = 1 (tests/golden/g8obst_*.npz): contorno(1), contorno(2)
zeroing the columns and moving the friction planes (3_contorno.f90:266-291),
contorno(3) filling the level set inside the obstacle (:445-490), the
momentum friction at the plane above each column, whole steps, the field
output's kaux and the restart file's columns.

Tolerances as elsewhere: boundary conditions exact, routines 1e-12
relative (FP64), steps 2e-5 (1e-4 if a reference CG solve stops one
iteration apart), mixed precision 2e-4 (10x the measured 2e-5).
"""
import dataclasses

import numpy as np
import pytest

import oracle as O
import pysuli
from helpers import CELL, VEL, gpu_from_state, obstacle_state, rel_err

pytestmark = pytest.mark.gpu

G16 = dataclasses.replace(O.Params(nx=16, ny=12, nz=10, dx=1 / 16, dy=1 / 16, dz=1 / 16), obst_t=1)
PER = dataclasses.replace(O.Params(nx=12, ny=8, nz=8, dx=1 / 12, dy=1 / 12, dz=1 / 12, ccx0=0, ccxf=0, cczf=1),
                          obst_t=1)


def _cmp(sv, st, names, tol, label):
    errs = {n: rel_err(sv.download(n), st.arr(n)) for n in names}
    bad = {k: v for k, v in errs.items() if not v <= tol}
    assert not bad, f"{label}: {bad} (tol {tol})"
    return errs


@pytest.mark.parametrize("p", [G16, PER], ids=["g16", "per"])
@pytest.mark.parametrize("nlock", [1, 2, 3])
def test_contorno_obstacle(p, nlock):
    print(f"contorno({nlock}) with obstacle grid {p.nx}x{p.ny}x{p.nz}", flush=True)
    st = obstacle_state(p, seed=31 + nlock)
    if nlock == 3:  # a level set that is not already smooth inside the obstacle
        st.ls[...] += 0.05 * np.random.default_rng(nlock).standard_normal(st.ls.shape)
    sv = gpu_from_state(st)
    try:
        sv.contorno(nlock)
        st.run("contorno", nlock)
        if nlock == 3:
            _cmp(sv, st, ["ls"], 0.0, "contorno(3) level set")
            _cmp(sv, st, ["hs", "hsx", "hsy", "hsz", "rho", "ls_nu"], 1e-12, "contorno(3) heaviside")
        else:
            _cmp(sv, st, VEL, 0.0, f"contorno({nlock})")
    finally:
        sv.close()


@pytest.mark.parametrize("p", [G16, PER], ids=["g16", "per"])
def test_convdiff_friction_above_obstacle(p):
    print(f"convdiff with obstacle grid {p.nx}x{p.ny}x{p.nz}", flush=True)
    st = obstacle_state(p, seed=35)
    sv = gpu_from_state(st)
    try:
        sv.convdiff_forces()
        st.run("convdiff")
        _cmp(sv, st, ["Fu", "Fv", "Fw"], 1e-12, "convdiff with obstacle")
    finally:
        sv.close()


def _dambreak_over_obstacle():
    p = dataclasses.replace(O.Params(nx=16, ny=8, nz=8, dx=1 / 16, dy=1 / 16, dz=1 / 16), obst_t=1)
    st = O.State(p)
    st.ls[...] = O.dambreak_phi(p, 0.25, 0.25)
    st.kw[8:12, 2:8] = 2  # a ridge across the tank, downstream of the column
    st.ku[8:13, 2:8] = 1
    st.kv[8:12, 2:9] = 1
    st.scal["dt"] = 1e-3
    st.run("ls_ini")
    st.run("contorno", 1)
    return p, st


def test_steps_obstacle_fp64():
    print("2 steps FP64 dam break over an obstacle", flush=True)
    p, st = _dambreak_over_obstacle()
    sv = gpu_from_state(st)
    try:
        same = True
        for it in range(1, 3):
            _, ps = sv.step(1e-3)
            st.run("level_set")
            st.run("contorno", 3)
            for tt in (1, 2):
                st.run("stage", tt, it=it)
                same &= ps[tt - 1]["iters"] == st.diag["cg_iters"]
        tol = 2e-5 if same else 1e-4
        vmax = max(float(np.max(np.abs(st.arr(n)))) for n in VEL)
        ev = max(float(np.max(np.abs(sv.download(n) - st.arr(n)))) for n in VEL) / vmax
        errs = _cmp(sv, st, ["ls", "hs", "rho"], 1e-8, "obstacle level set")
        errs.update(_cmp(sv, st, ["prd1"], tol, "obstacle pressure"))
        print(f"obstacle: velocity {ev:.2e}", errs)
        assert ev <= tol, ev
    finally:
        sv.close()


def test_steps_obstacle_mixed():
    print("3 steps mixed FCG dam break over an obstacle", flush=True)
    p, st = _dambreak_over_obstacle()
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
        assert ev <= 2e-4, ev  # measured 2.0e-5: about 10x
    finally:
        sv.close()


def test_kaux_and_restart_carry_the_obstacle(tmp_path):
    print("kaux and restart columns 16x12x10", flush=True)
    st = obstacle_state(G16, seed=37)
    nx, ny, nz = G16.nx, G16.ny, G16.nz
    sv = gpu_from_state(st)
    try:
        rec = sv.frame_record(t_plot=0)
        n1, n = (nx + 1) * (ny + 1) * (nz + 1), nx * ny * nz
        kaux = rec[6 * n1 + n: 6 * n1 + 2 * n].reshape((nx, ny, nz), order="F")
        kw = st.kw[1:nx + 1, 1:ny + 1].astype(int)
        want = (np.arange(1, nz + 1)[None, None, :] <= kw[:, :, None] - 1).astype(np.float32)
        assert np.array_equal(kaux, want) and want.sum() > 0
        sv.restart_write(str(tmp_path / "a"))
    finally:
        sv.close()
    # the columns follow u, v, w in the file (11_diversos.f90:323); a context
    # that reads the file takes them over and writes them back
    nu, nv, nw = (nx + 3) * (ny + 2) * (nz + 2), (nx + 2) * (ny + 3) * (nz + 2), (nx + 2) * (ny + 2) * (nz + 3)
    sizes = [(nx + 3) * (ny + 2), (nx + 2) * (ny + 3), (nx + 2) * (ny + 2)]

    def columns(path):
        tok = np.array(open(path).read().split(), dtype=float)
        out, o = [], nu + nv + nw
        for n in sizes:
            out.append(tok[o:o + n])
            o += n
        return out
    for got, name in zip(columns(tmp_path / "a"), ("ku", "kv", "kw")):
        assert np.array_equal(got, st.arr(name).ravel(order="F")), name
    other = pysuli.Solver.from_oracle_params(G16)
    try:
        other.restart_read(str(tmp_path / "a"))
        other.restart_write(str(tmp_path / "b"))
        for a, b in zip(columns(tmp_path / "a"), columns(tmp_path / "b")):
            assert np.array_equal(a, b)
    finally:
        other.close()


def test_obstacle_api_errors():
    sv = pysuli.Solver.from_oracle_params(G16)
    try:
        nx, ny, nz = G16.nx, G16.ny, G16.nz
        ku, kv, kw = np.zeros((nx + 3, ny + 2)), np.zeros((nx + 2, ny + 3)), np.zeros((nx + 2, ny + 2))
        kw[3, 3] = nz + 2
        with pytest.raises(pysuli.SuliError, match="kw"):
            sv.set_obstacle(ku, kv, kw)
        kw[3, 3] = nz + 1
        sv.set_obstacle(ku, kv, kw)
        sv.set_obstacle(None, None, None)
    finally:
        sv.close()
