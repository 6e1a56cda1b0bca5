"""The CPU oracle, pinned three ways:

1. golden vectors written by the flang-built reference harness
   (tests/golden/*.npz, generated by tests/golden/gen_golden.py) must be
   reproduced bit for bit, routine by routine;
2. where the reference build is available here (oracle/_ref), a fresh
   seeded state is run through both and must agree bit for bit;
3. closed forms that a plausible slip (wrong ghost formula, swapped
   one-sided derivative, dropped term) would break.
"""
import glob
import os

import numpy as np
import pytest

import oracle as O

GOLD = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz")))


def _params_from(d):
    nx, ny, nz, ccx0, ccxf, ccy0, ccyf, ccz0, cczf = [int(x) for x in d["__params"]]
    dx, dy, dz = [float(x) for x in d["__h"]]
    return O.Params(nx=nx, ny=ny, nz=nz, dx=dx, dy=dy, dz=dz, ccx0=ccx0, ccxf=ccxf, ccy0=ccy0, ccyf=ccyf,
                    ccz0=ccz0, cczf=cczf, t_tempo=int(d["__tt"][0]) if "__tt" in d else 1,
                    **(dict(zip(("der", "adv_type", "m_turb", "obst_t"), [int(x) for x in d["__model"]])) if "__model" in d
                       else {}))


@pytest.mark.parametrize("path", GOLD, ids=[os.path.basename(p)[:-4] for p in GOLD])
def test_golden_bitexact(path):
    d = np.load(path)
    p = _params_from(d)
    s = O.State(p)
    for name, _ in O.ARRAYS:  # fixtures written before the RK3/AB2 arrays existed hold none: zero
        s.arr(name)[...] = d["in_" + name] if ("in_" + name) in d else 0.0
    s.scal = dict(zip(O.SCALARS, [float(x) for x in d["__scal_in"]]))
    tag = os.path.basename(path)[:-4].split("_", 1)[1]
    cmd, arg = tag, 0
    for base in ("tempo", "contorno", "stage", "step"):
        if tag.startswith(base) and tag[len(base):].isdigit():
            cmd, arg = base, int(tag[len(base):])
    s.run(cmd, arg)
    n_changed = 0
    for name, _ in O.ARRAYS:
        got = s.arr(name)
        want = d["out_" + name] if ("out_" + name) in d else (d["in_" + name] if ("in_" + name) in d else
                                                               np.zeros_like(got))
        n_changed += ("out_" + name) in d
        if name == "Fw" and p.adv_type == 2:
            # rotacional reads dn(:,:,nz+2), which the reference never sets
            # (7_advectivo.f90:419 assigns dp there): the top face plane of Fw
            # is unpinned, every other value must match
            got, want = got[:, :, :-1], want[:, :, :-1]
        assert np.array_equal(got, want), f"{name}: max diff {np.max(np.abs(got - want)):.3e}"
    assert n_changed > 0
    for q, k in enumerate(O.SCALARS):
        assert s.scal[k] == float(d["__scal_out"][q]), k


@pytest.mark.skipif(not O.ref_available(O.Params()), reason="flang reference harness not built here")
@pytest.mark.parametrize("cmd,arg", [("level_set", None), ("stage", 1), ("step", 2)])
def test_live_reference(cmd, arg):
    p = O.Params()
    s = O.random_state(p, seed=77, perturb=0.3)
    s.run("visco")
    s.run("contorno", 1)
    a, b = s.copy(), s.copy()
    O.ref_run(a, cmd, arg)
    b.run(cmd, arg or 0)
    assert np.array_equal(a.buf, b.buf)


@pytest.mark.parametrize("scheme", [0, 2, 3])
def test_live_reference_schemes(scheme):
    """Three steps of Euler, RK3 and AB2 (AB2 through its it > 1 branch)
    against the reference built with that t_tempo, from nonzero kept forces."""
    import dataclasses
    p = dataclasses.replace(O.Params(), t_tempo=scheme)
    if not O.ref_available(p):
        pytest.skip("flang reference harness of this scheme not built here")
    s = O.random_state(p, seed=78 + scheme, perturb=0.3)
    s.run("visco")
    s.run("contorno", 1)
    rng = np.random.default_rng(scheme)
    for n in ["fu0", "fv0", "fw0", "fu1", "fv1", "fw1"]:
        s.arr(n)[...] = 0.1 * rng.standard_normal(s.arr(n).shape)
    a, b = s.copy(), s.copy()
    O.ref_run(a, "step", 3)
    b.run("step", 3)
    assert np.array_equal(a.buf, b.buf)


# ---- closed forms -------------------------------------------------------------
@pytest.mark.parametrize("ihs", [1, 2])
def test_weno_cubic_exact(ihs):
    """Every WENO5 candidate differentiates a cubic exactly and the weights
    sum to one, so both one-sided derivatives equal the exact derivative
    wherever the stencil stays inside the pencil (14_level_set.f90:572-608)."""
    n, h = 24, 0.1
    x = (np.arange(1, n + 1) - 0.5) * h
    phi = 0.3 * x**3 - x**2 + 2 * x - 1
    dphi = 0.9 * x**2 - 2 * x + 2
    dp, dn = O.weno1(phi, h, ihs)
    inner = slice(3, n - 3)
    assert np.allclose(dp[inner], dphi[inner], rtol=0, atol=1e-11)
    assert np.allclose(dn[inner], dphi[inner], rtol=0, atol=1e-11)


def test_weno_fifth_order():
    errs = []
    for n in (40, 80, 160):
        h = 1.0 / (n - 1)
        x = np.arange(n) * h
        dp, dn = O.weno1(np.sin(2 * np.pi * x), h, 0)
        exact = 2 * np.pi * np.cos(2 * np.pi * x)
        errs.append(np.max(np.abs(dp - exact)))
    order = np.log2(errs[0] / errs[1]), np.log2(errs[1] / errs[2])
    assert min(order) > 4.5, order


def test_weno_ghost_formulas():
    """Ghost extrapolation (14_level_set.f90:541-569): for ihs=1 the low side
    uses (12,-9,2)/5, exact for linear data, and the high side (18,-9,2)/11,
    which is not; ihs=2 uses the 1/11 formula on both sides, mirror-symmetric.
    A swapped or symmetric-by-mistake formula breaks one of these."""
    phi = np.arange(1, 11, dtype=float)
    dp, dn = O.weno1(phi, 1.0, 1)
    assert np.allclose(dn, 1.0, rtol=0, atol=1e-14)  # low side exact for linear data
    assert np.allclose(dp[:-1], 1.0, rtol=0, atol=1e-14) and dp[-1] < 0.9  # high side is not
    dp2, dn2 = O.weno1(phi, 1.0, 2)
    assert dn2[0] == pytest.approx(dp2[-1], abs=1e-14) and dp2[-1] == pytest.approx(dp[-1], abs=1e-14)
    assert np.allclose(dn2[1:], 1.0, rtol=0, atol=1e-14) and np.allclose(dp2[:-1], 1.0, rtol=0, atol=1e-14)
    dp0, dn0 = O.weno1(np.sin(2 * np.pi * np.arange(10) / 9.0), 1.0, 0)
    # periodic with period n-1: points 1 and n coincide
    assert dp0[0] == pytest.approx(dp0[-1], abs=1e-12) and dn0[0] == pytest.approx(dn0[-1], abs=1e-12)


def test_heaviside_planar():
    """For a planar level set phi = z0 - z, H, grad H and rho follow the
    smoothed Heaviside in closed form (14_level_set.f90:659-698)."""
    p = O.Params(nx=8, ny=8, nz=24, dx=1 / 24, dy=1 / 24, dz=1 / 24)
    s = O.State(p)
    z = (np.arange(1, p.nz + 1) - 0.5) * p.dz
    s.ls[...] = (0.5 - z)[None, None, :]
    s.run("heaviside")
    eps = p.alpha1 * p.dx1
    ph = 0.5 - z
    H = np.where(ph < -eps, 0.0, np.where(ph > eps, 1.0, 0.5 * (1 + ph / eps + np.sin(np.pi * ph / eps) / np.pi)))
    Hz = np.where(np.abs(ph) <= eps, 0.5 * (-1.0) * (1 + np.cos(np.pi * ph / eps)) / eps, 0.0)
    assert np.allclose(s.hs[3, 3, :], H, atol=1e-14)
    assert np.allclose(s.hsz[3, 3, 3:-3], Hz[3:-3], atol=1e-9)
    assert np.allclose(s.rho, p.rho_f1 * (1 - s.hs) + p.rho_f2 * s.hs, rtol=1e-15)


def _assemble(s, dt):
    """Dense scaled operator and right-hand side of graddin (9_grad.f90:46-151)."""
    p = s.p
    nx, ny, nz = p.nx, p.ny, p.nz
    rho = s.rho
    def rf(axis, f, j, k, i):  # face density with end faces copied
        a = np.moveaxis(rho, axis, 0)
        n = a.shape[0]
        fc = min(max(f, 2), n)
        return 0.5 * (a[fc - 1] + a[fc - 2])
    N = nx * ny * nz
    idx = lambda i, j, k: (i - 1) + nx * ((j - 1) + ny * (k - 1))
    c = {}
    for k in range(1, nz + 1):
        for j in range(1, ny + 1):
            for i in range(1, nx + 2):
                c[("x", i, j, k)] = dt / p.dx**2 / (0.5 * (rho[min(max(i, 2), nx) - 1, j - 1, k - 1] + rho[min(max(i, 2), nx) - 2, j - 1, k - 1]))
    for k in range(1, nz + 1):
        for j in range(1, ny + 2):
            for i in range(1, nx + 1):
                c[("y", i, j, k)] = dt / p.dy**2 / (0.5 * (rho[i - 1, min(max(j, 2), ny) - 1, k - 1] + rho[i - 1, min(max(j, 2), ny) - 2, k - 1]))
    for k in range(1, nz + 2):
        for j in range(1, ny + 1):
            for i in range(1, nx + 1):
                c[("z", i, j, k)] = dt / p.dz**2 / (0.5 * (rho[i - 1, j - 1, min(max(k, 2), nz) - 1] + rho[i - 1, j - 1, min(max(k, 2), nz) - 2]))
    A = np.zeros((N, N))
    D = np.zeros(N)
    q = np.zeros(N)
    u, v, w = s.u, s.v, s.w
    for k in range(1, nz + 1):
        for j in range(1, ny + 1):
            for i in range(1, nx + 1):
                r = idx(i, j, k)
                D[r] = (c[("x", i + 1, j, k)] + c[("x", i, j, k)] + c[("y", i, j + 1, k)] + c[("y", i, j, k)]
                        + c[("z", i, j, k + 1)] + c[("z", i, j, k)])
                q[r] = (u[i, j, k] - u[i + 1, j, k]) / p.dx + (v[i, j, k] - v[i, j + 1, k]) / p.dy + (w[i, j, k] - w[i, j, k + 1]) / p.dz
    for k in range(1, nz + 1):
        for j in range(1, ny + 1):
            for i in range(1, nx + 1):
                r = idx(i, j, k)
                A[r, r] += D[r]
                for (di, dj, dk, key) in [(1, 0, 0, ("x", i + 1, j, k)), (-1, 0, 0, ("x", i, j, k)),
                                          (0, 1, 0, ("y", i, j + 1, k)), (0, -1, 0, ("y", i, j, k)),
                                          (0, 0, 1, ("z", i, j, k + 1)), (0, 0, -1, ("z", i, j, k))]:
                    ii, jj, kk = i + di, j + dj, k + dk
                    ii, jj, kk = min(max(ii, 1), nx), min(max(jj, 1), ny), min(max(kk, 1), nz)
                    A[r, idx(ii, jj, kk)] -= c[key]
    return A, D, q


def test_graddin_dense_solve():
    """The projection solves sum_f c_f (p_P - p_nb) = -div u* with Neumann
    walls; at 6x5x4 the dense least-squares solution (the operator is
    singular: constants) must match the CG result up to a constant."""
    p = O.Params(nx=6, ny=5, nz=4, dx=0.2, dy=0.2, dz=0.2)
    s = O.random_state(p, seed=5, width=0.5, height=0.4)
    s.run("contorno", 2)
    dt = 1e-3
    s.scal["dt"] = dt
    A, D, q = _assemble(s, dt)
    assert np.allclose(A, A.T, rtol=1e-13, atol=0)          # symmetric
    assert np.allclose(A @ np.ones(len(D)), 0, atol=1e-9 * np.max(D))  # constant null space
    s.prd1[...] = 0.0
    s.run("graddin")
    pcg = s.prd1[1:-1, 1:-1, 1:-1].ravel(order="F")
    pls = np.linalg.lstsq(A, q, rcond=None)[0]
    dcg = pcg - pcg.mean()
    dls = pls - pls.mean()
    assert np.max(np.abs(dcg - dls)) <= 1e-3 * np.max(np.abs(dls)), np.max(np.abs(dcg - dls))


def test_graddin_uniform_density_dct():
    """Uniform density: the operator is the Neumann Laplacian times dt/rho,
    diagonalised by the DCT-II."""
    from scipy.fft import dctn, idctn
    p = O.Params(nx=8, ny=6, nz=6, dx=0.1, dy=0.1, dz=0.1)
    s = O.random_state(p, seed=9, width=2.0, height=2.0)  # all water
    assert np.all(s.rho == p.rho_f2)
    s.run("contorno", 2)
    dt = 1e-3
    s.scal["dt"] = dt
    s.prd1[...] = 0.0
    s.run("graddin")
    q = ((s.u[1:-2, 1:-1, 1:-1] - s.u[2:-1, 1:-1, 1:-1]) / p.dx + (s.v[1:-1, 1:-2, 1:-1] - s.v[1:-1, 2:-1, 1:-1]) / p.dy
         + (s.w[1:-1, 1:-1, 1:-2] - s.w[1:-1, 1:-1, 2:-1]) / p.dz)
    c = dt / p.rho_f2 / p.dx**2
    Q = dctn(q, type=2, norm="ortho")
    lam = [2 - 2 * np.cos(np.pi * np.arange(n) / n) for n in (p.nx, p.ny, p.nz)]
    L = c * (lam[0][:, None, None] + lam[1][None, :, None] + lam[2][None, None, :])
    L[0, 0, 0] = 1.0
    P = Q / L
    P[0, 0, 0] = 0.0
    pex = idctn(P, type=2, norm="ortho")
    pcg = s.prd1[1:-1, 1:-1, 1:-1]
    err = np.max(np.abs((pcg - pcg.mean()) - pex)) / np.max(np.abs(pex))
    assert err < 1e-3, err


def test_contorno_noslip_mirror():
    """No-slip walls: tangential ghosts are the negated first interior value,
    normal ghosts mirror across the wall face (3_contorno.f90:74-78,132-136)."""
    p = O.Params()
    s = O.random_state(p, seed=3)
    s.run("contorno", 1)
    assert np.array_equal(s.u[2:-2, 0, 1:-1], -s.u[2:-2, 1, 1:-1])
    assert np.array_equal(s.v[1:-1, 0, 1:-1], s.v[1:-1, 2, 1:-1])
    assert np.array_equal(s.u[0, 1:-1, 1:-1], s.u[2, 1:-1, 1:-1])
    assert np.array_equal(s.w[1:-1, 1:-1, 0], s.w[1:-1, 1:-1, 2])
    s.run("contorno", 2)
    assert np.all(s.u[1, :, :] == 0) and np.all(s.u[p.nx + 1, :, :] == 0)
    assert np.all(s.w[:, :, 1] == 0) and np.all(s.w[:, :, p.nz + 1] == 0)


def test_cg_summation_order_sensitivity():
    """Scale of the reference CG's round-off sensitivity, which sets the GPU
    tolerance of whole steps: the same scaled CG (9_grad.f90:175-264) with
    numpy's pairwise sums instead of sequential ones stops after the same
    number of iterations on a cold-start dam-break solve, and its pressure
    differs from the oracle's by less than 2e-6 relative (but far more than
    the 1e-15 of a single operation)."""
    import scipy.sparse as sp
    from helpers import assemble_projection
    p = O.Params(nx=16, ny=8, nz=8, dx=1 / 16, dy=1 / 16, dz=1 / 16)
    st = O.State(p)
    st.ls[...] = O.dambreak_phi(p, 0.25, 0.25)
    st.scal["dt"] = 1e-3
    st.run("ls_ini")
    st.run("contorno", 1)
    st.run("level_set")
    st.scal["dt"] = 5e-4
    for cmd, arg in (("visco", 0), ("convdiff", 0), ("tempo", 1), ("contorno", 2)):
        st.run(cmd, arg)
    a = st.copy().run("graddin")
    A, D, q = assemble_projection(p, st.rho, st.u, st.v, st.w, 5e-4)
    sD = np.sqrt(D)
    As = sp.diags(1 / sD) @ A @ sp.diags(1 / sD)
    e = st.prd1[1:-1, 1:-1, 1:-1].ravel(order="F") * sD
    r = As @ e - q / sD
    pp = r.copy()
    am = float(np.sum(r * r))
    thr, it = 1e-4 / len(D), 0
    while abs(am) > thr and it < 10000:
        it += 1
        mp = As @ pp
        am = float(np.sum(r * r))
        al = am / float(np.sum(pp * mp))
        e = e - al * pp
        r = r - al * mp
        pp = r + float(np.sum(r * r)) / am * pp
    po = a.prd1[1:-1, 1:-1, 1:-1].ravel(order="F")
    d = np.max(np.abs(e / sD - po)) / np.max(np.abs(po))
    assert it == a.diag["cg_iters"]
    assert 1e-12 < d < 2e-6, d


# ---- est (4_plot.f90:500-525) -------------------------------------------------
EST = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "est", "*.npz")))


@pytest.mark.parametrize("path", EST, ids=[os.path.basename(p)[:-4] for p in EST])
def test_est_golden(path):
    """The oracle's max |div u| equals the reference est's maximum bit for
    bit (seeded velocities, a projected field, one raised wall face)."""
    d = np.load(path)
    nx, ny, nz = [int(x) for x in d["grid"]]
    dx, dy, dz = [float(x) for x in d["h"]]
    per = os.path.basename(path).startswith("est_per")
    p = O.Params(nx=nx, ny=ny, nz=nz, dx=dx, dy=dy, dz=dz, **(dict(ccx0=0, ccxf=0, cczf=1) if per else {}))
    s = O.State(p)
    s.u[...], s.v[...], s.w[...] = d["u"], d["v"], d["w"]
    assert s.est_div() == float(d["divmax"][0])
    assert float(np.max(np.abs(d["div"]))) == float(d["divmax"][0])


def test_est_closed_forms():
    """u = a x, v = b y, w = c z has div = a + b + c in every cell; raising
    one u face by delta moves the divergence of the two cells beside it by
    -+delta/dx and leaves every other cell alone.  Unequal spacings catch a
    swapped dx/dy/dz, a + b + c = 0 makes the raised face the maximum."""
    p = O.Params(nx=6, ny=5, nz=4, dx=0.125, dy=0.25, dz=0.5)
    s = O.State(p)
    xf = (np.arange(p.nx + 3) - 1.0) * p.dx  # u(i) sits at x = (i-1) dx
    yf = (np.arange(p.ny + 3) - 1.0) * p.dy
    zf = (np.arange(p.nz + 3) - 1.0) * p.dz
    a, b, c = 3.0, 1.0, 0.5
    s.u[...] = a * xf[:, None, None]
    s.v[...] = b * yf[None, :, None]
    s.w[...] = c * zf[None, None, :]
    assert s.est_div() == a + b + c
    s.u[...] = 1.0 * xf[:, None, None]
    s.v[...] = 1.0 * yf[None, :, None]
    s.w[...] = -2.0 * zf[None, None, :]
    assert s.est_div() == 0.0
    delta = 0.375
    s.u[4, 2, 3] += delta  # the face between cells i = 3 and i = 4 of row (j, k) = (2, 3)
    assert s.est_div() == delta / p.dx
    s.u[4, 2, 3] -= delta
    s.w[2, 2, p.nz + 1] -= delta  # top wall face of cell (2, 2, nz)
    assert s.est_div() == delta / p.dz


# ---- dam-break initial level set ----------------------------------------------
def _segment_distance(px, pz, ax, az, bx, bz):
    """Euclidean distance from points (px, pz) to the segment a-b."""
    ex, ez = bx - ax, bz - az
    t = np.clip(((px - ax) * ex + (pz - az) * ez) / (ex * ex + ez * ez), 0.0, 1.0)
    return np.hypot(px - (ax + t * ex), pz - (az + t * ez))


@pytest.mark.parametrize("width,height", [(0.25, 0.25), (0.3125, 0.40625), (0.2, 0.45)])
def test_dambreak_phi_distance(width, height):
    """The initial level set is the signed distance to the column's free
    faces: the distance from each cell centre to the two segments x = width,
    0 <= z <= height and z = height, 0 <= x <= width (the faces on walls are
    not interfaces), positive inside the column, the same on every y row.
    A swapped width/height, a centre offset of one cell or a Manhattan
    corner distance fails here."""
    p = O.Params(nx=32, ny=5, nz=24, dx=1 / 32, dy=0.1, dz=1 / 48)
    phi = O.dambreak_phi(p, width, height)
    x = (np.arange(1, p.nx + 1) - 0.5) * p.dx
    z = (np.arange(1, p.nz + 1) - 0.5) * p.dz
    X, Z = np.meshgrid(x, z, indexing="ij")
    d = np.minimum(_segment_distance(X, Z, width, 0.0, width, height),
                   _segment_distance(X, Z, 0.0, height, width, height))
    inside = (X < width) & (Z < height)
    exact = np.where(inside, d, -d)
    for j in range(p.ny):
        assert np.allclose(phi[:, j, :], exact, rtol=0, atol=1e-14)
    # positive exactly at the centres inside the column, zero on the free
    # faces (a centre on z = height when height/dz is a half integer)
    assert np.array_equal(np.sign(phi[:, 0, :]), np.sign(exact))
    assert np.any(inside) and np.any(~inside)


def test_dambreak_phi_eikonal():
    """|grad phi| = 1 by central differences wherever the stencil stays in one
    smooth region (off the inner bisector x - width = z - height and off
    the outer corner's edges), and the outside corner distance is the
    Euclidean one."""
    p = O.Params(nx=64, ny=2, nz=64, dx=1 / 64, dy=0.5, dz=1 / 64)
    width, height = 0.5, 0.5
    phi = O.dambreak_phi(p, width, height)[:, 0, :]
    h = p.dx
    gx = (phi[2:, 1:-1] - phi[:-2, 1:-1]) / (2 * h)
    gz = (phi[1:-1, 2:] - phi[1:-1, :-2]) / (2 * h)
    g = np.hypot(gx, gz)
    x = (np.arange(2, p.nx) - 0.5) * h
    z = (np.arange(2, p.nz) - 0.5) * h
    X, Z = np.meshgrid(x, z, indexing="ij")
    smooth = (np.abs((X - width) - (Z - height)) > 2 * h) & (np.abs(X - width) > 2 * h) & (np.abs(Z - height) > 2 * h)
    corner = (X > width) & (Z > height)
    # linear pieces (inside, beside the faces): central differences are exact
    assert np.allclose(g[smooth & ~corner], 1.0, atol=1e-12)
    # outside corner, phi = -r: |central difference| = 1 + O(h^2/r^2)
    r = np.hypot(X - width, Z - height)
    assert np.all(np.abs(g[smooth & corner] - 1.0) <= (h / r[smooth & corner]) ** 2)
    # outside corner: phi = -sqrt((x-w)^2 + (z-h)^2)
    i, k = 60, 50  # 0-based, both beyond the column
    xc, zc = (i + 0.5) * h, (k + 0.5) * h
    assert phi[i, k] == pytest.approx(-np.hypot(xc - width, zc - height), abs=1e-15)


def test_stencil_relres_matches_assembly():
    """The matrix-free residual of the large-grid tests equals the one of
    the sparse assembly (itself checked against the dense operator above)."""
    from helpers import assemble_projection, scaled_relres, stencil_relres
    p = O.Params(nx=12, ny=9, nz=7, dx=1 / 12, dy=1 / 10, dz=1 / 8)
    s = O.random_state(p, seed=4)
    rng = np.random.default_rng(1)
    pres = rng.standard_normal(s.prd1.shape)
    A, D, q = assemble_projection(p, s.rho, s.u, s.v, s.w, 1e-3)
    want = scaled_relres(A, D, q, pres[1:-1, 1:-1, 1:-1].ravel(order="F"))
    got = stencil_relres(p, s.rho, s.u, s.v, s.w, 1e-3, pres)
    assert got == pytest.approx(want, rel=1e-12)
