"""The fast forms of the stencil kernels (suli_set_fast_stencils(1)) against
their plain forms, and in FP64 against the oracle:

- the LDS-tiled momentum predictor (k_momentum_tile: the inner cells of
  convdiff + tempo, 6_convdiff.f90:147-186, 7_advectivo.f90:33-113,
  11_diversos.f90:10-228) against the kernel that reads every stencil value
  from memory;
- the reinitialisation's sign function (14_level_set.f90:400-435) computed
  inside its first RK stage against its own pass;
- the finest multigrid level's half sweeps and output sweep with two cells
  of a colour per thread (k_mg0_gs2, k_mg0_out2) against one (k_mg0_gs,
  k_mg0_out), in the projection.

The grids split the inner box into ragged tiles along every direction
(64-cell rows, 4-row tiles, 16-plane chunks), so tile edges, the x ends
(mirrored halo columns), the halo loads at the frame's end and the last
partial chunk are all exercised.  Both kernels evaluate the same
expressions in the same order: the FP64 and FP32 results must agree bit
for bit."""
import dataclasses

import numpy as np
import pytest

import oracle as O
from helpers import CELL, consistent_state, gpu_from_state, obstacle_state, rel_err

pytestmark = pytest.mark.gpu

# (the third splits into 8 tile rows, which turns the XCD-band block order on)
SHAPES = [(150, 14, 40), (70, 9, 21), (70, 34, 20)]


def _params(shape, **kw):
    nx, ny, nz = shape
    return O.Params(nx=nx, ny=ny, nz=nz, dx=1 / nx, dy=1 / nx, dz=1 / nx, **kw)


def _run(st, prec, tiles):
    """Forces, then the RK2 first-stage update, with the tile kernel on or off."""
    sv = gpu_from_state(st, precision=prec)
    sv.set_fast_stencils(tiles)
    sv.convdiff_forces()
    out = {n: sv.download(n) for n in ("Fu", "Fv", "Fw")}
    sv.convdiff()
    sv.tempo(1, 1e-3)
    out.update({n: sv.download(n) for n in ("u", "v", "w", "u0", "v0", "w0")})
    sv.close()
    return out


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_tiles_match_direct_loads(shape, prec):
    import pysuli
    pc = pysuli.PREC_FP64 if prec == "fp64" else pysuli.PREC_MIXED
    st = consistent_state(_params(shape), seed=11)
    a, b = _run(st, pc, 0), _run(st, pc, 1)
    diff = {n: float(np.max(np.abs(a[n] - b[n]))) for n in a}
    print(f"{shape} {prec}: max |tiles - direct| per field", diff)
    assert all(v == 0.0 for v in diff.values()), diff


def test_tiles_with_obstacle_match_direct_loads():
    import pysuli
    st = obstacle_state(_params((70, 9, 21), obst_t=1), seed=12)
    a, b = _run(st, pysuli.PREC_FP64, 0), _run(st, pysuli.PREC_FP64, 1)
    for n in a:
        assert np.array_equal(a[n], b[n]), n


@pytest.mark.parametrize("shape", SHAPES)
def test_tiles_match_oracle(shape):
    """FP64 forces of the tiled kernel against the oracle's convdiff (the
    parity bar of the other stencils: 1e-12 of the field's largest value)."""
    import pysuli
    st = consistent_state(_params(shape), seed=13)
    sv = gpu_from_state(st, precision=pysuli.PREC_FP64)
    sv.set_fast_stencils(1)
    sv.convdiff_forces()
    st.run("convdiff")
    errs = {n: rel_err(sv.download(n), st.arr(n)) for n in ("Fu", "Fv", "Fw")}
    sv.close()
    print(f"{shape}: tiled forces vs oracle", errs)
    assert all(e <= 1e-12 for e in errs.values()), errs



@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("case", ["rk3", "ab2", "euler", "per_xy", "sources"])
def test_tiles_other_updates_match_direct_loads(prec, case):
    """The tile kernel's run-time stage updates (Euler, the three RK3 stages,
    AB2 after its Euler start), periodic x and y walls, and momentum sources
    tf_u, tf_v, tf_w, each against the direct-load kernel."""
    import pysuli
    pc = pysuli.PREC_FP64 if prec == "fp64" else pysuli.PREC_MIXED
    p = _params((70, 9, 21))
    scheme = {"rk3": 2, "ab2": 3, "euler": 0}.get(case, 1)
    p = dataclasses.replace(p, t_tempo=scheme)
    if case == "per_xy":
        p = dataclasses.replace(p, ccx0=0, ccxf=0, ccy0=0, ccyf=0)
    st = consistent_state(p, seed=15)
    rng = np.random.default_rng(15)
    for n in ("fu0", "fv0", "fw0", "fu1", "fv1", "fw1"):
        st.arr(n)[...] = 0.1 * rng.standard_normal(st.arr(n).shape)
    names = ["u", "v", "w", "u0", "v0", "w0", "fu0", "fv0", "fw0", "fu1", "fv1", "fw1"]
    out = {}
    for fast in (0, 1):
        sv = gpu_from_state(st, precision=pc)
        sv.set_fast_stencils(fast)
        if case == "sources":
            for n in ("tf_u", "tf_v", "tf_w"):
                sv.upload(n, 0.3 * np.random.default_rng(16).standard_normal(sv.shape(n)))
        for tt in range(1, sv.time_stages() + 1):
            sv.convdiff()
            sv.tempo(tt, 1e-3)
            sv.contorno(1)
        if case == "ab2":  # the AB2 update proper after its Euler start
            sv.convdiff()
            sv.tempo(1, 1e-3)
        out[fast] = {n: sv.download(n) for n in names}
        sv.close()
    diff = {n: float(np.max(np.abs(out[0][n] - out[1][n]))) for n in names}
    print(f"{case} {prec}: max |tiles - direct| per field", diff)
    assert all(v == 0.0 for v in diff.values()), diff


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("per", [False, True])
def test_level_set_fast_matches_plain(prec, per):
    """One whole level_set (advection, reinitialisation, volume, heaviside,
    curvature) in both forms, with walls or periodic x; then the FP64 fast
    form against the oracle at the level-set parity bar."""
    import pysuli
    pc = pysuli.PREC_FP64 if prec == "fp64" else pysuli.PREC_MIXED
    p = _params((70, 9, 21))
    if per:
        p = dataclasses.replace(p, ccx0=0, ccxf=0)
    st = consistent_state(p, seed=14, perturb=0.5)
    out = {}
    for fast in (0, 1):
        sv = gpu_from_state(st, precision=pc)
        sv.set_fast_stencils(fast)
        sv.level_set_ini()
        stats = sv.level_set(2e-3)
        out[fast] = {n: sv.download(n) for n in CELL}
        out[fast]["vol"] = np.array(sv.volumes())
        sv.close()
    diff = {n: float(np.max(np.abs(out[0][n] - out[1][n]))) for n in out[0]}
    print(f"{prec} per={per}: max |fast - plain| per field", diff, "stats", stats)
    assert all(v == 0.0 for v in diff.values()), diff
    if prec == "fp64":
        st.run("ls_ini")
        st.scal["dt"] = 2e-3
        st.run("level_set")
        errs = {n: rel_err(out[1][n], st.arr(n)) for n in ("ls", "hs", "rho", "kurv")}
        print("fast FP64 level set vs oracle", errs)
        assert all(e <= 1e-10 for e in errs.values()), errs


@pytest.mark.parametrize("grid", [(128, 64, 64), (70, 34, 20), (96, 48, 40)])
@pytest.mark.parametrize("form", [1, 2])
@pytest.mark.parametrize("fuse_max", [65536, 4096])
def test_multigrid_pair_sweeps_match_single(grid, form, fuse_max):
    """The finest level's kernels with two cells per thread (fast: the half
    sweeps and the output sweep) against one (plain):
    the preconditioner output z = M r bit for bit (an application outside
    the solve: FP32 colour arrays, no z.q); projections with the colour
    arrays in FP32 (form 1) and FP16 (form 2): the same iterations and the
    same pressure up to the solve's rounding (the z.q sums of the two forms
    add their cells in different orders; the constant of the Neumann problem
    is left out).  nx = 70 gives 35 pairs per row, so the last thread's
    second pair is the ghost pair and the output sweep takes its row end
    cell by cell.  With at most 4096 cells in a fused level, levels 1 (all
    three grids) and 2 (128x64x64) run one launch per half sweep, their first
    red half sweep computed by the restriction that fills them (fast) or by
    its own launch (plain)."""
    import pysuli
    nx, ny, nz = grid
    sv = pysuli.Solver(nx, ny, nz, 1.0 / nx, precision=pysuli.PREC_MIXED, poisson=pysuli.POISSON_FCG,
                       precond=pysuli.PRECOND_GMG, tol_rel=1e-8)
    try:
        sv.set_mg(fuse_max_cells=fuse_max, fuse_fine=form)
        sv.init_dambreak(0.3, 0.4)
        sv.contorno(1)
        sv.contorno(3)
        sv.step(2e-4)
        sv.set_guess(0)
        p0 = sv.download("prd1")
        r = np.random.default_rng(3).standard_normal((nz, ny, nx))
        out = {}
        for fast in (0, 1):
            sv.set_fast_stencils(fast)
            sv.upload("prd1", p0)
            st = sv.graddin(2e-4)
            out[fast] = (st, sv.download("prd1"), sv.precond_apply(r))
        pa, pb = (out[f][1][1:-1, 1:-1, 1:-1] for f in (0, 1))
        dp = np.abs((pb - pb.mean()) - (pa - pa.mean())).max() / np.abs(pa - pa.mean()).max()
        print(f"{grid} form {form} fuse_max {fuse_max}: iterations {out[0][0]['iters']} / {out[1][0]['iters']}, "
              f"max |dp| / max |p| = {dp:.2e} (mean removed)")
        assert np.array_equal(out[1][2], out[0][2])
        assert out[0][0]["converged"] and out[0][0]["iters"] > 0
        assert out[1][0]["iters"] == out[0][0]["iters"]
        assert dp <= 1e-6, dp
    finally:
        sv.close()
