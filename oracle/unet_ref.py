"""NumPy restatement of the learned Poisson preconditioner (test
infrastructure only: the tests use it to check the GPU network; the product
never imports it).

The network and its training loss have no counterpart in the reference,
whose projection (9_grad.f90:10-278) is Jacobi-scaled CG.  What this module
shares with the reference is the operator the loss is taken through: the
symmetrically scaled projection matrix of graddin, A_s = D^-1/2 A D^-1/2,
with face coefficients dt/(h^2 rho_f) (9_grad.f90:46-76), face densities
averaged from the two cells and copied at the end faces
(8_derivadas.f90:304-434), and the wall faces counted in the diagonal D
(9_grad.f90:80-87) while the mirrored ghost makes their off-diagonal term the
cell itself (9_grad.f90:99-116).  The tests pin that operator against a
sparse assembly of the reference's system, the convolution against a
scalar-loop definition, and the gradient against finite differences.

Arrays are channels-last, indexed [k, j, i, c], i.e. cell v = i + nx*(j +
ny*k) with the channel fastest.  The network has L levels (channels
min(16*2^l, 64)); the weight vector holds, layer by layer in the order of
arch(L), the kernel W[27][c_out][c_in] (tap = (dz+1)*9 + (dy+1)*3 + (dx+1))
followed by the bias[c_out]; L is recovered from its length.

With ``bf16=True`` the functions round at the points where the GPU stores
BF16 (layer outputs, pooled/upsampled tensors, weights of all layers but the
first and the last, backward gradient tensors) so that the remaining difference is the
accumulation order; with ``bf16=False`` everything is float64.
"""
from __future__ import annotations

import numpy as np

def channels(L):
    return [min(16 << l, 64) for l in range(L)]


def arch(L=3):
    """Layer (c_in, c_out) list of an L-level network in weight-vector order:
    encoder (a, b per level 0..L-2), bottom (a, b), decoder (levels L-2 down
    to 0), output."""
    C = channels(L)
    lay = [(2, C[0]), (C[0], C[0])]
    for l in range(1, L - 1):
        lay += [(C[l - 1], C[l]), (C[l], C[l])]
    lay += [(C[L - 2], C[L - 1]), (C[L - 1], C[L - 2])]
    for l in range(L - 2, 0, -1):
        lay.append((C[l], C[l - 1]))
    lay += [(C[0], C[0]), (C[0], 1)]
    return lay


LAYERS = arch(3)


def nparams(L=3) -> int:
    return sum(27 * ci * co + co for ci, co in arch(L))


def levels_of(nw: int) -> int:
    for L in range(2, 7):
        if nparams(L) == nw:
            return L
    raise ValueError(f"no architecture has {nw} parameters")


def unpack(wvec):
    """Split the flat weight vector into [(W[27,co,ci], b[co])]."""
    out, off = [], 0
    for ci, co in arch(levels_of(len(wvec))):
        W = np.asarray(wvec[off:off + 27 * co * ci], dtype=np.float64).reshape(27, co, ci)
        off += 27 * co * ci
        b = np.asarray(wvec[off:off + co], dtype=np.float64)
        off += co
        out.append((W, b))
    assert off == len(wvec)
    return out


def pack(layers):
    return np.concatenate([np.concatenate([W.ravel(), b.ravel()]) for W, b in layers])


def bf16(x):
    """Round float64 values to BF16 (through FP32, round to nearest even)."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    nan = (u & 0x7F800000) == 0x7F800000
    r = np.where(nan, u, r).astype(np.uint32)
    return r.view(np.float32).astype(np.float64)


def f32(x):
    return np.asarray(x, dtype=np.float32).astype(np.float64)


def taps():
    for t in range(27):
        yield t, t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1  # tap, dz, dy, dx


def conv(x, W, b=None):
    """3x3x3 convolution, zero padding: out[v,co] = b + sum_tap,ci x[v+tap,ci] W[tap,co,ci]."""
    nz, ny, nx, _ = x.shape
    xp = np.pad(x, ((1, 1), (1, 1), (1, 1), (0, 0)))
    out = np.zeros((nz, ny, nx, W.shape[1]))
    for t, dz, dy, dx in taps():
        out += xp[1 + dz:1 + dz + nz, 1 + dy:1 + dy + ny, 1 + dx:1 + dx + nx, :] @ W[t].T
    if b is not None:
        out += b
    return out


def conv_t(g, W):
    """Adjoint of conv in its input: din[u,ci] = sum_tap g[u-tap,co] W[tap,co,ci]."""
    nz, ny, nx, _ = g.shape
    gp = np.pad(g, ((1, 1), (1, 1), (1, 1), (0, 0)))
    out = np.zeros((nz, ny, nx, W.shape[2]))
    for t, dz, dy, dx in taps():
        out += gp[1 - dz:1 - dz + nz, 1 - dy:1 - dy + ny, 1 - dx:1 - dx + nx, :] @ W[t]
    return out


def conv_w(g, x):
    """Weight gradient of conv: dW[tap,co,ci] = sum_v g[v,co] x[v+tap,ci]."""
    nz, ny, nx, _ = x.shape
    xp = np.pad(x, ((1, 1), (1, 1), (1, 1), (0, 0)))
    dW = np.zeros((27, g.shape[3], x.shape[3]))
    G = g.reshape(-1, g.shape[3])
    for t, dz, dy, dx in taps():
        dW[t] = G.T @ xp[1 + dz:1 + dz + nz, 1 + dy:1 + dy + ny, 1 + dx:1 + dx + nx, :].reshape(-1, x.shape[3])
    return dW


def pool(x):
    nz, ny, nx, C = x.shape
    return x.reshape(nz // 2, 2, ny // 2, 2, nx // 2, 2, C).mean(axis=(1, 3, 5))


def up(x):
    return x.repeat(2, axis=0).repeat(2, axis=1).repeat(2, axis=2)


def down_sum(x):
    nz, ny, nx, C = x.shape
    return x.reshape(nz // 2, 2, ny // 2, 2, nx // 2, 2, C).sum(axis=(1, 3, 5))


def relu(x):
    return np.maximum(x, 0.0)


def inputs(r, rho, rho_f1, rho_f2, bf=False):
    """x0 = [r/rms(r), (log rho - l1)/l2], l1 the mean and l2 the half
    difference of the two fluids' log densities; r, rho as [k, j, i]."""
    r = f32(r) if bf else np.asarray(r, dtype=np.float64)
    rho = f32(rho) if bf else np.asarray(rho, dtype=np.float64)
    l1 = 0.5 * (np.log(rho_f1) + np.log(rho_f2))
    l2 = max(0.5 * abs(np.log(rho_f2) - np.log(rho_f1)), 1e-6)
    rms = np.sqrt(np.sum(r * r) / r.size)
    x0 = np.stack([r / rms if rms > 0 else 0 * r, (np.log(rho) - l1) / l2], axis=-1)
    return f32(x0) if bf else x0


def forward(wvec, x0, bf=False):
    """Network output z[k,j,i] and the activations (for backward)."""
    L = levels_of(len(wvec))
    lay = unpack(wvec)
    n = len(lay)
    q = bf16 if bf else (lambda a: a)
    Wm = [((bf16(W) if (bf and 0 < i < n - 1) else (f32(W) if bf else W)), (f32(b) if bf else b))
          for i, (W, b) in enumerate(lay)]
    ea = lambda l: 2 * l
    eb = lambda l: 2 * l + 1
    bot = 2 * (L - 1)
    dec = lambda l: 2 * L + (L - 2 - l)
    A = {"x0": x0, "W": Wm, "L": L}
    A["A0"] = q(relu(conv(x0, *Wm[ea(0)])))
    A["S0"] = q(relu(conv(A["A0"], *Wm[eb(0)])))
    for l in range(1, L):
        A[f"P{l}"] = q(pool(A[f"S{l - 1}"]))
        if l < L - 1:
            A[f"A{l}"] = q(relu(conv(A[f"P{l}"], *Wm[ea(l)])))
            A[f"S{l}"] = q(relu(conv(A[f"A{l}"], *Wm[eb(l)])))
        else:
            A[f"A{l}"] = q(relu(conv(A[f"P{l}"], *Wm[bot])))
            A["B"] = q(relu(conv(A[f"A{l}"], *Wm[bot + 1])))
    prev = A["B"]
    for l in range(L - 2, -1, -1):
        A[f"U{l}"] = q(up(prev) + A[f"S{l}"])
        A[f"D{l}"] = q(relu(conv(A[f"U{l}"], *Wm[dec(l)])))
        prev = A[f"D{l}"]
    z = conv(A["D0"], *Wm[n - 1])[..., 0]
    return (f32(z) if bf else z), A


def backward(dz, A, bf=False):
    """Gradient of a scalar with dL/dz = dz through the network; returns the
    flat gradient vector."""
    q = bf16 if bf else (lambda a: a)
    Wm, L = A["W"], A["L"]
    n = len(Wm)
    G = [None] * n
    ea = lambda l: 2 * l
    eb = lambda l: 2 * l + 1
    bot = 2 * (L - 1)
    dec = lambda l: 2 * L + (L - 2 - l)

    def layer(i, g, x):
        G[i] = (conv_w(g, x), g.reshape(-1, g.shape[-1]).sum(axis=0))

    dz = dz[..., None]
    layer(n - 1, dz, A["D0"])
    gD = {0: q(conv_t(dz, Wm[n - 1][0]) * (A["D0"] > 0))}
    gU = {}
    for l in range(0, L - 1):
        layer(dec(l), gD[l], A[f"U{l}"])
        gU[l] = q(conv_t(gD[l], Wm[dec(l)][0]))
        if l + 1 == L - 1:
            gB = q(down_sum(gU[l]) * (A["B"] > 0))
        else:
            gD[l + 1] = q(down_sum(gU[l]) * (A[f"D{l + 1}"] > 0))
    l = L - 1
    layer(bot + 1, gB, A[f"A{l}"])
    gA = q(conv_t(gB, Wm[bot + 1][0]) * (A[f"A{l}"] > 0))
    layer(bot, gA, A[f"P{l}"])
    gP = q(conv_t(gA, Wm[bot][0]))
    for l in range(L - 2, -1, -1):
        gS = q((gU[l] + up(gP) / 8.0) * (A[f"S{l}"] > 0))
        layer(eb(l), gS, A[f"A{l}"])
        gA = q(conv_t(gS, Wm[eb(l)][0]) * (A[f"A{l}"] > 0))
        if l > 0:
            layer(ea(l), gA, A[f"P{l}"])
            gP = q(conv_t(gA, Wm[ea(l)][0]))
        else:
            layer(ea(0), gA, A["x0"])
    return pack(G)


# ---- the scaled projection operator on the cell grid ------------------------

def face_coefs(rho, dx, dy, dz):
    """Face coefficients c_f = (1/h^2)/rho_f along x, y, z; rho as [k, j, i].
    Returns arrays with n+1 faces along their own axis (face f between cells
    f-1 and f, end faces copied from their neighbour)."""
    def along(axis, h):
        a = np.moveaxis(rho, axis, 0)
        n = a.shape[0]
        f = np.empty((n + 1,) + a.shape[1:])
        f[1:n] = 0.5 * (a[1:] + a[:-1])
        f[0] = f[1]
        f[n] = f[n - 1]
        return np.moveaxis(1.0 / (h * h) / f, 0, axis)
    return along(2, dx), along(1, dy), along(0, dz)  # x is the last axis


def apply_As(z, rho, dx, dy, dz):
    """A_s z = z - sum_f c_f/sqrt(D_P D_Q) z_Q, a wall face coupling the cell
    to itself (Neumann ghost)."""
    return apply_As_faces(z, *face_coefs(rho, dx, dy, dz))


def mg_faces(rho, dt, dx, dy, dz, level):
    """Face conductances of multigrid level `level` of the FCG preconditioner
    (no counterpart in the reference): on the finest grid dt/(h^2 rho_f) with
    the face densities of face_coefs (the graddin coefficients), rounded to
    FP32 and then BF16, zero on the wall faces (the Neumann operator without
    the ghost coupling); each coarser face is half the sum, in FP32, of the
    four fine faces on it.  rho as [k, j, i]; returns (cx, cy, cz) with n+1
    faces along their own axis."""
    rho = np.asarray(rho, dtype=np.float64)

    def fine(axis, h):
        a = np.moveaxis(rho, axis, 0)
        n = a.shape[0]
        f = np.zeros((n + 1,) + a.shape[1:])
        f[1:n] = (dt / (h * h)) / ((a[1:] + a[:-1]) * 0.5)
        return np.moveaxis(bf16(f), 0, axis).astype(np.float32)

    c = [fine(2, dx), fine(1, dy), fine(0, dz)]
    h = np.float32(0.5)
    for _ in range(level):
        cx, cy, cz = c
        # coarse face (K, J, I) of each direction: fine faces at 2I (2J, 2K)
        # along the face normal, the 2x2 fine faces over it summed in the
        # order (a, b), (a+1, b), (a, b+1), (a+1, b+1) of the two tangential
        # indices (x faces: j then k; y faces: i then k; z faces: i then j)
        x = cx[:, :, ::2]
        cx = h * (((x[0::2, 0::2] + x[0::2, 1::2]) + x[1::2, 0::2]) + x[1::2, 1::2])
        y = cy[:, ::2, :]
        cy = h * (((y[0::2, :, 0::2] + y[0::2, :, 1::2]) + y[1::2, :, 0::2]) + y[1::2, :, 1::2])
        zf = cz[::2]
        cz = h * (((zf[:, 0::2, 0::2] + zf[:, 0::2, 1::2]) + zf[:, 1::2, 0::2]) + zf[:, 1::2, 1::2])
        c = [cx, cy, cz]
    return tuple(np.asarray(a, dtype=np.float64) for a in c)


def block_mean(rho, level):
    """Mean of rho over the 2^level-cubes of cells (the density input of the
    network on multigrid level `level`)."""
    m = 1 << level
    nz, ny, nx = rho.shape
    return np.asarray(rho, dtype=np.float64).reshape(nz // m, m, ny // m, m, nx // m, m).mean(axis=(1, 3, 5))


def apply_As_faces(z, cx, cy, cz):
    """A_s z = z - sum_f c_f/sqrt(D_P D_Q) z_Q for face conductances cx, cy,
    cz (n+1 faces along their own axis, face f between cells f-1 and f),
    D = the sum of a cell's six faces, a wall face coupling the cell to itself
    (Neumann ghost)."""
    D = cx[:, :, 1:] + cx[:, :, :-1] + cy[:, 1:] + cy[:, :-1] + cz[1:] + cz[:-1]
    s = np.asarray(z, dtype=np.float64).copy()

    def term(c_hi, c_lo, axis):
        zp = np.moveaxis(z, axis, 0)
        Dp = np.moveaxis(D, axis, 0)
        chi = np.moveaxis(c_hi, axis, 0)
        clo = np.moveaxis(c_lo, axis, 0)
        # neighbour values with the mirrored ghost at the walls
        zn_hi = np.concatenate([zp[1:], zp[-1:]], axis=0)
        Dn_hi = np.concatenate([Dp[1:], Dp[-1:]], axis=0)
        zn_lo = np.concatenate([zp[:1], zp[:-1]], axis=0)
        Dn_lo = np.concatenate([Dp[:1], Dp[:-1]], axis=0)
        t = chi / np.sqrt(Dp * Dn_hi) * zn_hi + clo / np.sqrt(Dp * Dn_lo) * zn_lo
        return np.moveaxis(t, 0, axis)

    s -= term(cx[:, :, 1:], cx[:, :, :-1], 2)
    s -= term(cy[:, 1:], cy[:, :-1], 1)
    s -= term(cz[1:], cz[:-1], 0)
    return s


def loss_and_grad(wvec, r, rho, rho_f1, rho_f2, dx, dy, dz, bf=False, faces=None):
    """L = 1 - (r.s)^2/(|r|^2 |s|^2), s = A_s z(r, rho); returns (L, dL/dw, z).
    A_s is the scaled projection operator of rho, or that of the face
    conductances `faces` = (cx, cy, cz) when given (a multigrid level)."""
    r64 = f32(r) if bf else np.asarray(r, dtype=np.float64)
    rho64 = f32(rho) if bf else np.asarray(rho, dtype=np.float64)
    x0 = inputs(r, rho, rho_f1, rho_f2, bf)
    z, A = forward(wvec, x0, bf)
    if faces is None:
        faces = face_coefs(rho64, dx, dy, dz)
    s = apply_As_faces(z, *faces)
    rs, ss, rr = float(np.sum(r64 * s)), float(np.sum(s * s)), float(np.sum(r64 * r64))
    L = 1.0 - rs * rs / (rr * ss)
    t = -2.0 * rs / (rr * ss) * r64 + 2.0 * rs * rs / (rr * ss * ss) * s
    if bf:
        t = f32(t)
    dz_ = apply_As_faces(t, *faces)
    if bf:
        dz_ = f32(dz_)
    return L, backward(dz_, A, bf), z


def loss(wvec, r, rho, rho_f1, rho_f2, dx, dy, dz):
    x0 = inputs(r, rho, rho_f1, rho_f2)
    z, _ = forward(wvec, x0)
    s = apply_As(z, np.asarray(rho, dtype=np.float64), dx, dy, dz)
    rs, ss, rr = float(np.sum(r * s)), float(np.sum(s * s)), float(np.sum(r * r))
    return 1.0 - rs * rs / (rr * ss)
