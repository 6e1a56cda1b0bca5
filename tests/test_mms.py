"""The manufactured-solution module (suli-openmp_amd/mms.py).

CPU: its sympy sources are checked against the reference's own discrete
operator (the oracle's convdiff, pinned bit for bit to the reference): with
the exact fields, Fu + tf_u - (dp/dx)/rho must vanish as the grid is refined,
at the first order of the upwind advection (and likewise for v and w).
GPU: the full time step converges to the manufactured solution at first
order in FP64 and in mixed precision, velocities, pressure and the
interface (the level set at exactly advected markers of the initial
sphere)."""
import numpy as np
import pytest
import sympy as sp

import mms
import oracle as O


def _residual(n):
    h = 1.0 / n
    p = O.Params(nx=n, ny=n, nz=n, dx=h, dy=h, dz=h, **mms.PARAMS)
    st = O.State(p)
    st.u[...] = mms.field_on_grid("u", st.u.shape, (0, 0, 0), "fcc", h)
    st.v[...] = mms.field_on_grid("v", st.v.shape, (0, 0, 0), "cfc", h)
    st.w[...] = mms.field_on_grid("w", st.w.shape, (0, 0, 0), "ccf", h)
    st.rho[...] = mms.RHO
    st.ls_nu[...] = mms.MU
    st.run("contorno", 1)
    st.run("visco")
    st.run("convdiff")
    out = []
    for F, tf, axis, kinds in (("Fu", "tf_u", mms.X, "fcc"), ("Fv", "tf_v", mms.Y, "cfc"), ("Fw", "tf_w", mms.Z, "ccf")):
        Fh = st.arr(F)
        src = mms.field_on_grid(tf, Fh.shape, (1, 1, 1), kinds, h)
        gp = sp.lambdify((mms.X, mms.Y, mms.Z), sp.diff(mms.P_EX, axis) / mms.RHO, "numpy")
        ax = [((np.arange(Fh.shape[a]) + 1) - (1.0 if kinds[a] == "f" else 0.5)) * h for a in range(3)]
        Xg, Yg, Zg = np.meshgrid(*ax, indexing="ij")
        g = np.broadcast_to(gp(Xg, Yg, Zg), Fh.shape)
        R = Fh + src - g
        if F == "Fw":  # the wall faces z = 0, 1 hold w = 0 and are overwritten by contorno
            R = R[:, :, 1:-1]
        out.append(float(np.sqrt(np.mean(R * R))))
    return np.array(out)


def test_sources_consistent_with_reference_operator():
    r16, r32 = _residual(16), _residual(32)
    print("residual 16^3", r16, "32^3", r32)
    # first order: halving h at least roughly halves the residual of every component
    assert np.all(r32 < 0.6 * r16)
    # and the residual is small against the size of the source terms (O(1..10))
    assert np.all(r32 < 1.0)


def test_exact_fields_satisfy_boundary_conditions():
    z = np.array([0.0, 1.0])
    x = np.linspace(0, 1, 7)
    X, Z = np.meshgrid(x, z, indexing="ij")
    assert np.allclose(mms.exact("w", X, 0.3, Z), 0.0, atol=1e-12)  # free slip: no normal flow
    dudz = sp.lambdify((mms.X, mms.Y, mms.Z), sp.diff(mms.U_EX, mms.Z), "numpy")
    assert np.allclose(dudz(X, 0.3, Z), 0.0, atol=1e-12)
    div = sp.simplify(sp.diff(mms.U_EX, mms.X) + sp.diff(mms.V_EX, mms.Y) + sp.diff(mms.W_EX, mms.Z))
    assert div == 0


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_time_step_converges_to_manufactured_solution(precision):
    print(f"mms {precision} 16^3 and 32^3", flush=True)
    a = mms.run(16, precision, t_end=0.05)
    b = mms.run(32, precision, t_end=0.05)
    print(a["errors"], b["errors"], flush=True)
    for f in ("u", "v", "w"):
        ea, eb = a["errors"][f]["l2"], b["errors"][f]["l2"]
        assert eb < ea and np.log2(ea / eb) > 0.7, (f, ea, eb)
    assert b["errors"]["p"]["l2"] < a["errors"]["p"]["l2"]
    assert abs(b["errors"]["volume"]["relative_drift"]) < 0.05


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_convergence_study_32_to_128(precision):
    """The survey's manufactured-solution study on 32^3, 64^3 and 128^3 to
    t = 0.1: every field (u, v, w, p and the interface) converges at first
    order, the order of the reference's upwind advection (measured 0.93-1.10
    from 32 to 64, 0.95-1.07 from 64 to 128; 0.97-1.05 from 128 to 256 in
    the full study), in both precisions."""
    res = []
    for n in (32, 64, 128):
        print(f"mms study {precision} {n}^3", flush=True)
        res.append(mms.run(n, precision, t_end=0.1))
        e = res[-1]["errors"]
        print({f: (round(e[f]["l2"], 6), round(e[f]["linf"], 6)) for f in ("u", "v", "w", "p", "phi")}, flush=True)
    for o in mms.orders(res):
        print(o["grids"], {k: round(v, 2) for k, v in o["order"].items()})
        for k, v in o["order"].items():
            assert 0.8 < v < 1.3, (o["grids"], k, v)
    assert res[-1]["errors"]["phi"]["linf"] < 2.5e-3
    assert abs(res[-1]["errors"]["volume"]["relative_drift"]) < 1e-3
