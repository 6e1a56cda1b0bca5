"""Manufactured-solution check of the time step on the unit cube.

Fields (wavenumber 2 pi, x and y periodic, free-slip walls at z = 0 and 1):

    u =  sin(2 pi x) cos(2 pi y) cos(2 pi z)
    v =  cos(2 pi x) sin(2 pi y) cos(2 pi z)
    w = -2 cos(2 pi x) cos(2 pi y) sin(2 pi z)          (div u = 0)
    p =  P0 cos(2 pi x) cos(2 pi y) cos(2 pi z)
    phi = 0.25 - |x - (0.5, 0.5, 0.5)|                   (sphere, radius 0.25)

One fluid (rho = 1, mu = 0.01 on both sides of the level set, no surface
tension, no gravity, no eddy viscosity), so the level set is advected by the
velocity without feeding back.  The steady momentum source makes (u, p) an
exact steady solution of the continuous equations that the time step
discretises (6_convdiff.f90:148-186 with the eddy viscosity zero: the
normal-stress terms carry only nut, the cross terms the molecular
viscosity interpolated to the edges), with the projection enforcing
continuity:

    tf = (u . grad) u - (1/rho) [viscous terms of convdiff] + grad p / rho

The sources are derived with sympy and evaluated at the staggered face
positions of tf_u(nx+1,ny,nz), tf_v, tf_w; they enter convdiff like the
reference's tf_u (6_convdiff.f90:156).  Starting from the exact fields, the
solver runs to a fixed time with dt = min(0.1 h, 0.2 h^2 rho/mu) (CFL 0.2,
viscous number 0.2; the explicit viscous terms are unstable above 0.25 in
two directions); the errors of u, v,
w (faces), p (cells, mean removed) and the interface volume are reported per
grid with the observed order log2(e(2h)/e(h)).

The level-set error is measured on the interface: 2000 markers spread over
the initial sphere are carried by the exact (steady) velocity field with
RK4 (400 steps, converged to 1e-12), and the numerical level set,
interpolated trilinearly to the markers, should vanish there; its RMS and
largest magnitude over the markers are the interface's L2 and Linf errors
(the level set is a distance: they are distances to the numerical
interface).

    python3 mms.py --grids 32 64 128 256 --precision mixed fp64
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np
import sympy as sp

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

RHO, MU, P0 = 1.0, 0.01, 0.5
X, Y, Z = sp.symbols("x y z", real=True)
_K = 2 * sp.pi
U_EX = sp.sin(_K * X) * sp.cos(_K * Y) * sp.cos(_K * Z)
V_EX = sp.cos(_K * X) * sp.sin(_K * Y) * sp.cos(_K * Z)
W_EX = -2 * sp.cos(_K * X) * sp.cos(_K * Y) * sp.sin(_K * Z)
P_EX = P0 * sp.cos(_K * X) * sp.cos(_K * Y) * sp.cos(_K * Z)


def _sources():
    """Symbolic tf_u, tf_v, tf_w (steady: no time derivative)."""
    u, v, w, p = U_EX, V_EX, W_EX, P_EX
    d = sp.diff
    adv = [u * d(f, X) + v * d(f, Y) + w * d(f, Z) for f in (u, v, w)]
    # convdiff's viscous terms with nut = 0 (molecular viscosity in the cross terms only)
    visc_u = MU * (d(d(u, Y) + d(v, X), Y) + d(d(u, Z) + d(w, X), Z))
    visc_v = MU * (d(d(v, X) + d(u, Y), X) + d(d(v, Z) + d(w, Y), Z))
    visc_w = MU * (d(d(w, X) + d(u, Z), X) + d(d(w, Y) + d(v, Z), Y))
    gp = [d(p, X), d(p, Y), d(p, Z)]
    return [sp.simplify(adv[q] - visc / RHO + gp[q] / RHO) for q, visc in enumerate((visc_u, visc_v, visc_w))]


TF_EX = _sources()
_LAMB = {name: sp.lambdify((X, Y, Z), e, "numpy") for name, e in
         zip(("u", "v", "w", "p", "tf_u", "tf_v", "tf_w"), (U_EX, V_EX, W_EX, P_EX, *TF_EX))}


def exact(name, x, y, z):
    out = _LAMB[name](x, y, z)
    return np.broadcast_to(out, np.broadcast(x, y, z).shape).astype(np.float64)


def _coords(n, h, kind):
    """Positions of an array with Fortran index 0.. along one axis: cell
    centres (i - 1/2) h or faces (i - 1) h."""
    i = np.arange(n)
    return (i - 0.5) * h if kind == "c" else (i - 1.0) * h


def field_on_grid(name, shape, lo, kinds, h):
    """Exact field on an array of Fortran shape `shape` whose first index is
    `lo` along each axis; kinds = ('f'|'c') per axis."""
    ax = [_coords(shape[a] + lo[a], h, kinds[a])[lo[a]:] for a in range(3)]
    Xg, Yg, Zg = np.meshgrid(*ax, indexing="ij")
    return np.asfortranarray(exact(name, Xg, Yg, Zg))


def sphere_markers(m=2000):
    """m points evenly spread over the initial sphere (Fibonacci lattice)."""
    k = np.arange(m) + 0.5
    zc = 1.0 - 2.0 * k / m
    r = np.sqrt(1.0 - zc * zc)
    th = math.pi * (1.0 + math.sqrt(5.0)) * k
    return np.stack([0.5 + 0.25 * r * np.cos(th), 0.5 + 0.25 * r * np.sin(th), 0.5 + 0.25 * zc], axis=1)


def advect_markers(xm, t_end, nsub):
    """Exact trajectories (RK4 in the steady analytic velocity) to t_end."""
    def vel(q):
        return np.stack([exact(c, q[:, 0], q[:, 1], q[:, 2]) for c in "uvw"], axis=1)
    dt = t_end / nsub
    q = xm.copy()
    for _ in range(nsub):
        k1 = vel(q)
        k2 = vel(q + 0.5 * dt * k1)
        k3 = vel(q + 0.5 * dt * k2)
        k4 = vel(q + dt * k3)
        q = q + dt / 6.0 * (k1 + 2 * k2 + 2 * k3 + k4)
    return q


def interp_cells(phi, q, h):
    """Trilinear interpolation of a cell array (centres (i - 1/2) h) at the
    points q; x and y periodic, z clamped to the first and last centres."""
    n = phi.shape
    s = q / h - 0.5
    i0 = np.floor(s).astype(int)
    f = s - i0
    out = np.zeros(len(q))
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                ii = (i0[:, 0] + dx) % n[0]
                jj = (i0[:, 1] + dy) % n[1]
                kk = np.clip(i0[:, 2] + dz, 0, n[2] - 1)
                wgt = (f[:, 0] if dx else 1 - f[:, 0]) * (f[:, 1] if dy else 1 - f[:, 1]) * (f[:, 2] if dz else 1 - f[:, 2])
                out += wgt * phi[ii, jj, kk]
    return out


def sphere_sdf(n, h):
    c = (np.arange(1, n + 1) - 0.5) * h
    Xg, Yg, Zg = np.meshgrid(c, c, c, indexing="ij")
    return np.asfortranarray(0.25 - np.sqrt((Xg - 0.5) ** 2 + (Yg - 0.5) ** 2 + (Zg - 0.5) ** 2))


PARAMS = dict(ccx0=0, ccxf=0, ccy0=0, ccyf=0, ccz0=1, cczf=1, m_turb=0, rho_f1=RHO, rho_f2=RHO, mi_f1=MU,
              mi_f2=MU, sigma=0.0, gx=0.0, gz=0.0)


def run(n, precision, t_end=0.1, tol=1e-10):
    import pysuli as S
    h = 1.0 / n
    prec = {"fp64": S.PREC_FP64, "mixed": S.PREC_MIXED}[precision]
    s = S.Solver(n, n, n, h, precision=prec, poisson=S.POISSON_FCG, precond=S.PRECOND_GMG, tol_rel=tol, **PARAMS)
    try:
        # exact initial state: faces (u along x, ...) and cells, ghosts by contorno
        s.upload("u", field_on_grid("u", s.shape("u"), (0, 0, 0), "fcc", h))
        s.upload("v", field_on_grid("v", s.shape("v"), (0, 0, 0), "cfc", h))
        s.upload("w", field_on_grid("w", s.shape("w"), (0, 0, 0), "ccf", h))
        s.upload("prd1", field_on_grid("p", s.shape("prd1"), (0, 0, 0), "ccc", h))
        s.upload("tf_u", field_on_grid("tf_u", s.shape("tf_u"), (1, 1, 1), "fcc", h))
        s.upload("tf_v", field_on_grid("tf_v", s.shape("tf_v"), (1, 1, 1), "cfc", h))
        s.upload("tf_w", field_on_grid("tf_w", s.shape("tf_w"), (1, 1, 1), "ccf", h))
        s.upload("ls", sphere_sdf(n, h))
        vol0 = s.level_set_ini()
        s.contorno(1)
        s.contorno(3)
        # advective CFL 0.2 (|u| <= 2) and, on fine grids, the explicit
        # viscous limit (mu/rho) dt/h^2 <= 0.2 of the cross-derivative terms
        dt = min(0.1 * h, 0.2 * h * h * RHO / MU)
        steps = max(1, int(math.ceil(t_end / dt)))
        dt = t_end / steps
        t0 = time.time()
        iters = 0
        for _ in range(steps):
            _, ps = s.step(dt)
            iters += sum(p["iters"] for p in ps)
        s.sync()
        wall = time.time() - t0
        err = {}
        for name, kinds in (("u", "fcc"), ("v", "cfc"), ("w", "ccf")):
            a = s.download(name)
            ex = field_on_grid(name, a.shape, (0, 0, 0), kinds, h)
            sl = [slice(1, n + 1)] * 3
            ax = "uvw".index(name)
            sl[ax] = slice(1, n + 1 if ax < 2 else n + 2)  # x, y periodic: faces 1..n; z: walls included
            e = (a - ex)[tuple(sl)]
            err[name] = {"l2": float(np.sqrt(np.mean(e * e))), "linf": float(np.abs(e).max())}
        p = s.download("prd1")[1:-1, 1:-1, 1:-1]
        pe = field_on_grid("p", (n + 2, n + 2, n + 2), (0, 0, 0), "ccc", h)[1:-1, 1:-1, 1:-1]
        e = (p - p.mean()) - (pe - pe.mean())
        err["p"] = {"l2": float(np.sqrt(np.mean(e * e))), "linf": float(np.abs(e).max())}
        vi, vn = s.volumes()
        err["volume"] = {"relative_drift": (vn - vol0) / vol0}
        markers = advect_markers(sphere_markers(), t_end, 400)
        ph = interp_cells(s.download("ls"), markers, h)
        err["phi"] = {"l2": float(np.sqrt(np.mean(ph * ph))), "linf": float(np.abs(ph).max())}
        return {"n": n, "precision": precision, "steps": steps, "dt": dt, "wall_s": wall,
                "poisson_iters_per_solve": iters / (2 * steps), "errors": err}
    finally:
        s.close()


def orders(results):
    """Observed order between consecutive grids of the same precision."""
    out = []
    for a, b in zip(results[:-1], results[1:]):
        if a["precision"] != b["precision"]:
            continue
        r = math.log(a["n"] / b["n"], 0.5)
        o = {}
        for f in ("u", "v", "w", "p", "phi"):
            for m in ("l2", "linf"):
                ea, eb = a["errors"][f][m], b["errors"][f][m]
                o[f"{f}_{m}"] = math.log(ea / eb, 2) / r if ea > 0 and eb > 0 else None
        out.append({"precision": a["precision"], "grids": [a["n"], b["n"]], "order": o})
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--grids", type=int, nargs="+", default=[32, 64, 128])
    ap.add_argument("--precision", nargs="+", default=["fp64", "mixed"], choices=["fp64", "mixed"])
    ap.add_argument("--t-end", type=float, default=0.1)
    ap.add_argument("--out", default=None)
    args = ap.parse_args()
    res = []
    for prec in args.precision:
        for n in args.grids:
            r = run(n, prec, args.t_end)
            e = r["errors"]
            print(f"{prec:5s} {n:4d}^3 steps {r['steps']:4d} {r['wall_s']:7.2f}s  "
                  + "  ".join(f"{f} {e[f]['l2']:.3e}/{e[f]['linf']:.3e}" for f in ("u", "v", "w", "p", "phi"))
                  + f"  vol {e['volume']['relative_drift']:+.2e}", flush=True)
            res.append(r)
    od = orders(res)
    for o in od:
        print(o["precision"], o["grids"], {k: (round(v, 2) if v is not None else None) for k, v in o["order"].items()})
    if args.out:
        with open(args.out, "w") as f:
            json.dump({"runs": res, "orders": od}, f, indent=1)


if __name__ == "__main__":
    main()
