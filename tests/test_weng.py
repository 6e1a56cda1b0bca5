"""The manufactured solution of Weng (2009) (13_mms.f90) and the
prescribed-velocity walls (wall code 3) in the oracle, against the
flang-built reference's fixtures (tests/golden/weng, gen_golden.py weng) and
against the mathematics the fixtures leave open."""
import glob
import os

import numpy as np
import pytest

import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden", "weng")


def _case(name):
    z = np.load(os.path.join(GOLD, name + ".npz"))
    mms_t, a, h0, coef = z["mms"]
    nx, ny, nz = (int(x) for x in z["grid"])
    dx, dy, dz = z["h"]
    p = O.Params(nx=nx, ny=ny, nz=nz, dx=dx, dy=dy, dz=dz, ccx0=3, ccxf=3, ccy0=3, ccyf=3, ccz0=3, cczf=3,
                 m_turb=0, gx=0.0, gz=0.0, mms_t=int(mms_t), mms_a=float(a), mms_h0=float(h0), mms_coef=float(coef))
    return p, z


def test_fixtures_present():
    assert sorted(os.path.basename(f) for f in glob.glob(os.path.join(GOLD, "*.npz"))) == [
        "weng1.npz", "weng1a.npz", "weng2.npz", "weng2a.npz"]


@pytest.mark.parametrize("name", ["weng1", "weng2"])
def test_run_matches_reference(name):
    """mms_run of the reference (its own setting a = 0): five steps without
    projection over prescribed walls; fields and the errors of mms() per step."""
    p, z = _case(name)
    st, errs = O.weng_run(p, float(z["dt"][0]), int(z["nsteps"][0]))
    for n in ("u", "v", "w", "ls", "prd1"):
        ref = z["out_" + n]
        scale = max(np.abs(ref).max(), 1e-300)
        assert np.abs(st.arr(n) - ref).max() <= 1e-13 * max(scale, 1.0), n
    ref_err = z["errors"][:, 1:]
    got = np.array(errs)
    assert got.shape == ref_err.shape
    fin = np.isfinite(ref_err)
    assert np.array_equal(fin, np.isfinite(got))  # the steady case has no exact pressure (NaN)
    assert np.allclose(got[fin], ref_err[fin], rtol=1e-10, atol=1e-300)
    # the errors grow from zero: the discrete operator is not exact
    assert np.all(got[1:, 0] > got[:-1, 0])


@pytest.mark.parametrize("name", ["weng1", "weng2", "weng1a", "weng2a"])
def test_planes_match_reference(name):
    """mms_bc's 24 planes (steady at t = 0, unsteady at t = 0.7), including the
    rows the reference's loops leave at zero and bzy0/bzyf of the unsteady
    case, which take the formula of w."""
    p, z = _case(name)
    planes = O.weng_planes(p, float(z["t_src"][0]))
    for n in O.PLANE_NAMES:
        assert planes[n].shape == z["plane_" + n].shape, n
        assert np.abs(planes[n] - z["plane_" + n]).max() <= 1e-14, n
    if p.mms_t == 2 and p.mms_a != 0.0:
        assert np.abs(z["plane_bzy0"]).max() > 0.0


@pytest.mark.parametrize("name", ["weng1a", "weng2a"])
def test_initial_state_matches_reference(name):
    """mms_i + level_set_ini's tail + contorno(1) + contorno(3) at t = 0 with
    a = 0.5 (the depth varies over the tank): level set, velocities with
    their prescribed ghosts, pressure."""
    p, z = _case(name)
    st, errs = O.weng_run(p, float(z["dt"][0]), 0)
    for n in ("u", "v", "w", "ls", "prd1"):
        assert np.abs(st.arr(n) - z["out_" + n]).max() <= 1e-14, n
    assert np.abs(st.ls).max() > 0.0


@pytest.mark.parametrize("name", ["weng1", "weng2"])
def test_sources_equal_reference_at_a0(name):
    """For a = 0 (the reference's setting) the exact sources equal the
    reference's hand-derived termo_fonte1 / termo_fonte2."""
    p, z = _case(name)
    src = O.weng_sources(p, float(z["t_src"][0]))
    for q, n in enumerate(("tf_u", "tf_v", "tf_w")):
        ref = z["src_" + n]
        assert np.abs(src[q] - ref).max() <= 1e-13 * max(np.abs(ref).max(), 1.0), n


def test_reference_sources_deviate_for_a_nonzero():
    """With a = 0.5 the reference's expressions are not the derivatives of
    its own fields: steady, tf_u's viscous part (tf_v, tf_w agree); unsteady,
    every component (the time derivative leaves out that h depends on t).
    The exact sources are pinned by finite differences in the next test."""
    p, z = _case("weng1a")
    src = O.weng_sources(p, 0.0)
    d = [np.abs(src[q] - z["src_" + n]).max() for q, n in enumerate(("tf_u", "tf_v", "tf_w"))]
    assert d[0] > 1e-7 and d[1] < 1e-12 and d[2] < 1e-12
    p, z = _case("weng2a")
    src = O.weng_sources(p, float(z["t_src"][0]))
    d = [np.abs(src[q] - z["src_" + n]).max() for q, n in enumerate(("tf_u", "tf_v", "tf_w"))]
    assert min(d) > 1e-3


@pytest.mark.parametrize("mms_t", [1, 2])
def test_sources_against_finite_differences(mms_t):
    """The sympy-derived sources against central differences of weng_vel
    (fourth order, step 1e-3): catches a dropped term, a wrong sign or a
    transposed derivative."""
    p = O.Params(nx=6, ny=5, nz=4, dx=0.5, dy=0.6, dz=0.3, mms_t=mms_t, mms_a=0.5, mms_h0=2.0)
    t = 0.7
    src = O.weng_sources(p, t)
    e = 1e-3
    nu = 0.00100798 / 998.0
    offs = [(1.0, 0.5, 0.5), (0.5, 1.0, 0.5), (0.5, 0.5, 1.0)]
    rngs = [((1, 7), (1, 5), (1, 4)), ((1, 6), (1, 6), (1, 4)), ((1, 6), (1, 5), (1, 5))]

    def f(comp, X, Y, Z, T):
        return O.weng_vel(p, "uvw"[comp], X, Y, Z, T)

    def d1(comp, P, ax):
        def at(s):
            Q = list(P)
            Q[ax] = Q[ax] + s
            return f(comp, *Q)
        return (8 * (at(e) - at(-e)) - (at(2 * e) - at(-2 * e))) / (12 * e)

    def d2(comp, P, a1, a2):
        if a1 == a2:
            def at(s):
                Q = list(P)
                Q[a1] = Q[a1] + s
                return f(comp, *Q)
            return (-at(2 * e) + 16 * at(e) - 30 * at(0) + 16 * at(-e) - at(-2 * e)) / (12 * e * e)

        def at(s1, s2):
            Q = list(P)
            Q[a1] = Q[a1] + s1
            Q[a2] = Q[a2] + s2
            return f(comp, *Q)
        return (at(e, e) - at(e, -e) - at(-e, e) + at(-e, -e)) / (4 * e * e)

    for c in range(3):
        X, Y, Z = O._grid(p, offs[c], rngs[c])
        P = [X, Y, Z, t + 0 * X]
        V = [f(k, *P) for k in range(3)]
        adv = sum(V[k] * d1(c, P, k) for k in range(3))
        visc = sum(d2(c, P, k, k) for k in range(3)) + sum(d2(k, P, c, k) for k in range(3))
        fd = adv - nu * visc + (d1(c, P, 3) if mms_t == 2 else 0.0)
        assert np.abs(fd - src[c]).max() <= 1e-7 * max(np.abs(src[c]).max(), 1.0), c


def test_wall_code3_contorno_reads_planes():
    """contorno(1) and contorno(2) with every wall prescribed copy the planes
    into the ghosts (and, nlock = 2, the wall-normal faces) in the order y,
    x, z walls (3_contorno.f90:58-443): with random planes, each ghost plane
    holds its plane except where a later wall overwrites an edge."""
    p = O.Params(nx=6, ny=5, nz=4, ccx0=3, ccxf=3, ccy0=3, ccyf=3, ccz0=3, cczf=3)
    rng = np.random.default_rng(5)
    st = O.State(p)
    st.planes = {n: rng.standard_normal(O.plane_shape(p, n)) for n in O.PLANE_NAMES}
    st.run("contorno", 2)
    P = st.planes
    nx, ny, nz = p.nx, p.ny, p.nz
    # z walls are written last in the ghost phase; the wall-normal faces of
    # nlock = 2 (u at i = 1, nx+1; v at j = 1, ny+1) come after them
    xi = [i for i in range(nx + 3) if i not in (1, nx + 1)]
    yj = [j for j in range(ny + 3) if j not in (1, ny + 1)]
    assert np.array_equal(st.u[xi, :, 0], P["bzx0"][xi]) and np.array_equal(st.u[xi, :, nz + 1], P["bzxf"][xi])
    assert np.array_equal(st.v[:, yj, 0], P["bzy0"][:, yj]) and np.array_equal(st.v[:, yj, nz + 1], P["bzyf"][:, yj])
    assert np.array_equal(st.w[:, :, 0], P["bzz0"]) and np.array_equal(st.w[:, :, nz + 2], P["bzzf1"])
    assert np.array_equal(st.u[1, :, 0], P["bxx1"][:, 0])
    # x walls: interior of the plane (z rows 1..nz are not overwritten)
    assert np.array_equal(st.u[0, :, 1:nz + 1], P["bxx0"][:, 1:nz + 1])
    assert np.array_equal(st.u[nx + 2, :, 1:nz + 1], P["bxxf1"][:, 1:nz + 1])
    assert np.array_equal(st.v[0, yj, 1:nz + 1], P["bxy0"][yj, 1:nz + 1])
    # y walls, away from the x and z ghosts
    assert np.array_equal(st.u[2:nx + 1, 0, 1:nz + 1], P["byx0"][2:nx + 1, 1:nz + 1])
    assert np.array_equal(st.v[1:nx + 1, ny + 2, 1:nz + 1], P["byyf1"][1:nx + 1, 1:nz + 1])
    # nlock = 2: the wall-normal faces
    assert np.array_equal(st.u[1, 1:ny + 1, 1:nz + 1], P["bxx1"][1:ny + 1, 1:nz + 1])
    assert np.array_equal(st.u[nx + 1, 1:ny + 1, 1:nz + 1], P["bxxf"][1:ny + 1, 1:nz + 1])
    assert np.array_equal(st.v[1:nx + 1, 1, 1:nz + 1], P["byy1"][1:nx + 1, 1:nz + 1])
    assert np.array_equal(st.w[:, :, 1], P["bzz1"]) and np.array_equal(st.w[:, :, nz + 1], P["bzzf"])
