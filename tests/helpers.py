"""Shared helpers of the parity tests: oracle states and GPU solvers built
from the same inputs, and array comparison with stated tolerances."""
import numpy as np

import oracle as O

CELL = ["rho", "ls_nu", "ls", "mod_ls", "kurv", "hs", "hsx", "hsy", "hsz", "ddlsdx", "ddlsdy", "ddlsdz"]
VEL = ["u", "v", "w"]


def rel_err(a, b):
    """max |a-b| / max(|b|, tiny), over the whole array."""
    a = np.asarray(a)
    b = np.asarray(b)
    sc = max(float(np.max(np.abs(b))) if b.size else 0.0, 1e-300)
    return float(np.max(np.abs(a - b))) / sc if b.size else 0.0


def consistent_state(p, seed=1, **kw):
    """Oracle state with derived level-set fields, eddy viscosity and ghost
    planes consistent, as they are after contorno(1) in the driver loop."""
    s = O.random_state(p, seed=seed, **kw)
    s.run("visco")
    s.run("contorno", 1)
    return s


def obstacle_state(p, seed=1, **kw):
    """consistent_state with a seeded bottom obstacle one to three cells high
    in the middle of the tank, clear of the side walls (ku, kv <= nz,
    kw <= nz+1 as 3_contorno.f90:741-760 clamps them), and columns below the
    bed beside it as obstaculo leaves them where the bed dips under z = 0:
    ku = kv = -1 with kw = 0 on the row after the bump, all three -1 on the
    row before it; p.obst_t must be nonzero."""
    assert p.obst_t
    s = O.random_state(p, seed=seed, **kw)
    rng = np.random.default_rng(seed)
    for name in ("ku", "kv", "kw"):
        a = s.arr(name)
        n0, n1 = a.shape
        top = 3 if name == "kw" else 2
        a[...] = 0.0
        a[n0 // 3: 2 * n0 // 3 + 1, 2: n1 - 2] = rng.integers(1, top + 1, size=(2 * n0 // 3 + 1 - n0 // 3, n1 - 4))
        a[2 * n0 // 3 + 1, 2: n1 - 2] = 0 if name == "kw" else -1
        a[n0 // 3 - 1, 2: n1 - 2] = -1
    s.run("visco")
    s.run("contorno", 1)
    return s


def gpu_from_state(state, **kw):
    """GPU solver with the state's parameters and fields; for a time scheme
    other than RK2, the scheme and its kept forces fu0, fu1 as well."""
    import pysuli
    sv = pysuli.Solver.from_oracle_params(state.p, **kw)
    pysuli.load_state(sv, state)
    if state.p.obst_t:
        sv.set_obstacle(state.ku, state.kv, state.kw)
    if (state.p.der, state.p.adv_type) != (1, 1):
        sv.set_advection(state.p.adv_type, state.p.der)
    if state.p.t_tempo != 1:
        sv.set_time_scheme(state.p.t_tempo)
        for n in ["fu0", "fv0", "fw0", "fu1", "fv1", "fw1"]:
            sv.upload(n, state.arr(n))
    return sv


def assemble_projection(p, rho, u, v, w, dt):
    """Sparse form of the projection system of graddin (9_grad.f90:46-116)
    for non-periodic walls: A p = sum over interior faces c_f (p_P - p_Q),
    c_f = dt/(h^2 rho_f) with rho_f the face average (end faces copied);
    D = sum of all six c_f (wall faces included, as the reference's scaling);
    q = -div u*.  Returns (A csr, D, q) over cells in Fortran order."""
    import scipy.sparse as sp
    nx, ny, nz = p.nx, p.ny, p.nz
    rho = np.asarray(rho)

    def face_rho(axis):
        a = np.moveaxis(rho, axis, 0)
        n = a.shape[0]
        f = np.empty((n + 1,) + a.shape[1:])
        f[1:n] = 0.5 * (a[1:] + a[:-1])
        f[0] = f[1]
        f[n] = f[n - 1]
        return np.moveaxis(f, 0, axis)

    cx = dt / p.dx**2 / face_rho(0)
    cy = dt / p.dy**2 / face_rho(1)
    cz = dt / p.dz**2 / face_rho(2)
    D = cx[1:] + cx[:-1] + cy[:, 1:] + cy[:, :-1] + cz[:, :, 1:] + cz[:, :, :-1]
    q = ((u[1:-2, 1:-1, 1:-1] - u[2:-1, 1:-1, 1:-1]) / p.dx + (v[1:-1, 1:-2, 1:-1] - v[1:-1, 2:-1, 1:-1]) / p.dy
         + (w[1:-1, 1:-1, 1:-2] - w[1:-1, 1:-1, 2:-1]) / p.dz)
    idx = np.arange(nx * ny * nz).reshape((nx, ny, nz), order="F")
    rows, cols, vals = [], [], []
    diag = np.zeros((nx, ny, nz))
    for axis, c in ((0, cx), (1, cy), (2, cz)):
        cin = np.take(c, np.arange(1, c.shape[axis] - 1), axis=axis)  # interior faces
        a = np.take(idx, np.arange(0, idx.shape[axis] - 1), axis=axis)
        b = np.take(idx, np.arange(1, idx.shape[axis]), axis=axis)
        rows += [a.ravel(), b.ravel()]
        cols += [b.ravel(), a.ravel()]
        vals += [-cin.ravel(), -cin.ravel()]
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[axis] = slice(0, -1)
        hi[axis] = slice(1, None)
        diag[tuple(lo)] += cin
        diag[tuple(hi)] += cin
    rows.append(idx.ravel())
    cols.append(idx.ravel())
    vals.append(diag.ravel())
    N = nx * ny * nz
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(N, N))
    return A, D.ravel(order="F"), q.ravel(order="F")


def scaled_relres(A, D, q, pcell):
    """||D^-1/2 (q - A p)|| / ||D^-1/2 q||, the solver's stopping measure."""
    s = 1.0 / np.sqrt(D)
    return float(np.linalg.norm(s * (q - A @ pcell)) / np.linalg.norm(s * q))


def dambreak_diagnostics(p, u, v, w, ls, rho):
    """Integral diagnostics of a dam-break state given in the reference
    extents: water volume (phi > 0 cells, cell count times dV), the front
    position as the reference writes it (4_plot.f90:364-375: scanning i from
    nx down to 2, the first sign change of ls along x gives dx*(i-0.5) +
    ls(i,j,k)), the kinetic energy sum 0.5 rho |u_c|^2 dV with the face
    velocities averaged to the cell centres, and max |u|."""
    nx, ny, nz = p.nx, p.ny, p.nz
    dV = p.dx * p.dy * p.dz
    front = None
    for i in range(nx, 1, -1):  # Fortran index i, arrays 0-based
        a, b = ls[i - 2], ls[i - 1]
        hit = np.argwhere((a * b) <= 0.0)
        if hit.size:
            # the reference loops k outer, j inner: first (k, j) in that order
            order = sorted(((kk, jj) for jj, kk in hit), key=lambda t: (t[0], t[1]))
            kk, jj = order[0]
            front = p.dx * (i - 0.5) + ls[i - 1, jj, kk]
            break
    uc = 0.5 * (u[1:-2, 1:-1, 1:-1] + u[2:-1, 1:-1, 1:-1])
    vc = 0.5 * (v[1:-1, 1:-2, 1:-1] + v[1:-1, 2:-1, 1:-1])
    wc = 0.5 * (w[1:-1, 1:-1, 1:-2] + w[1:-1, 1:-1, 2:-1])
    ke = float(np.sum(0.5 * rho * (uc * uc + vc * vc + wc * wc)) * dV)
    return {"volume": float(np.count_nonzero(ls > 0) * dV), "front": front, "ke": ke,
            "umax": float(max(np.abs(u).max(), np.abs(v).max(), np.abs(w).max()))}


def stencil_relres(p, rho, u, v, w, dt, pres):
    """||D^-1/2 (q - A p)|| / ||D^-1/2 q|| of the projection system of
    graddin (9_grad.f90:46-151, non-periodic walls) by array stencils,
    without assembling a matrix: for grids too large for the sparse
    assembly of assemble_projection.  pres holds the ghosted pressure
    (0:nx+1, 0:ny+1, 0:nz+1); only its interior is used (walls are Neumann:
    the wall faces carry no flux)."""
    rho = np.asarray(rho, dtype=np.float64)

    def face_rho(axis):
        a = np.moveaxis(rho, axis, 0)
        n = a.shape[0]
        f = np.empty((n + 1,) + a.shape[1:])
        f[1:n] = 0.5 * (a[1:] + a[:-1])
        f[0] = f[1]
        f[n] = f[n - 1]
        return np.moveaxis(f, 0, axis)

    pc = np.asarray(pres, dtype=np.float64)[1:-1, 1:-1, 1:-1]
    q = ((u[1:-2, 1:-1, 1:-1] - u[2:-1, 1:-1, 1:-1]) / p.dx + (v[1:-1, 1:-2, 1:-1] - v[1:-1, 2:-1, 1:-1]) / p.dy
         + (w[1:-1, 1:-1, 1:-2] - w[1:-1, 1:-1, 2:-1]) / p.dz)
    Ap = np.zeros_like(pc)
    D = np.zeros_like(pc)
    for axis, h in ((0, p.dx), (1, p.dy), (2, p.dz)):
        c = dt / h**2 / face_rho(axis)
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[axis], hi[axis] = slice(0, -1), slice(1, None)
        D += c[tuple(hi)] + c[tuple(lo)]
        cin = np.take(c, np.arange(1, c.shape[axis] - 1), axis=axis)  # interior faces
        flux = cin * (pc[tuple(lo)] - pc[tuple(hi)])  # c_f (p_low - p_high)
        Ap[tuple(lo)] += flux
        Ap[tuple(hi)] -= flux
        del c, cin, flux
    s = 1.0 / np.sqrt(D)
    return float(np.linalg.norm(s * (q - Ap)) / np.linalg.norm(s * q))
