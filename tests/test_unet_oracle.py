"""Pins of the NumPy learned-preconditioner oracle (CPU only).

The convolution is checked against its scalar-loop definition; the data and
weight gradients of convolution, pooling and upsampling against their
adjoint identities; the scaled operator against a sparse assembly of the
reference's projection system (9_grad.f90:46-116); the whole gradient against
central finite differences; and the BF16 rounding against torch's
float32 -> bfloat16 conversion."""
import numpy as np
import pytest
import torch

import unet_ref as U
from helpers import assemble_projection
from oracle import Params

RNG = np.random.default_rng(7)


def test_nparams():
    assert U.nparams() == 27 * 6704 + 225  # 181233
    w = RNG.standard_normal(U.nparams())
    assert np.array_equal(U.pack(U.unpack(w)), w)


def test_conv_matches_loop_definition():
    nz, ny, nx, ci, co = 3, 4, 5, 2, 3
    x = RNG.standard_normal((nz, ny, nx, ci))
    W = RNG.standard_normal((27, co, ci))
    b = RNG.standard_normal(co)
    out = U.conv(x, W, b)
    ref = np.zeros_like(out)
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                for c in range(co):
                    s = b[c]
                    for dz in (-1, 0, 1):
                        for dy in (-1, 0, 1):
                            for dx in (-1, 0, 1):
                                kk, jj, ii = k + dz, j + dy, i + dx
                                if 0 <= kk < nz and 0 <= jj < ny and 0 <= ii < nx:
                                    t = (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1)
                                    s += W[t, c] @ x[kk, jj, ii]
                    ref[k, j, i, c] = s
    assert np.allclose(out, ref, rtol=1e-13, atol=1e-13)


def test_conv_adjoints():
    x = RNG.standard_normal((4, 6, 8, 3))
    W = RNG.standard_normal((27, 5, 3))
    g = RNG.standard_normal((4, 6, 8, 5))
    lhs = np.sum(U.conv(x, W) * g)
    assert np.isclose(lhs, np.sum(x * U.conv_t(g, W)), rtol=1e-12)
    assert np.isclose(lhs, np.sum(W * U.conv_w(g, x)), rtol=1e-12)


def test_pool_up_adjoints():
    x = RNG.standard_normal((4, 6, 8, 3))
    y = RNG.standard_normal((2, 3, 4, 3))
    assert np.isclose(np.sum(U.pool(x) * y), np.sum(x * U.up(y) / 8.0), rtol=1e-12)
    assert np.isclose(np.sum(U.up(y) * x), np.sum(y * U.down_sum(x)), rtol=1e-12)
    c = RNG.standard_normal((1, 1, 1, 2))
    assert np.allclose(U.pool(U.up(c)), c)


def _two_phase_rho(nz, ny, nx, seed=3):
    g = np.random.default_rng(seed)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    phi = 0.35 * nx - np.hypot(i - 0.3 * nx, k - 0.2 * nz) + 0.5 * g.standard_normal((nz, ny, nx))
    h = np.clip(0.5 + phi / 3.0, 0, 1)
    return 1.2 + (1000.0 - 1.2) * h


@pytest.mark.parametrize("dims", [(8, 6, 4), (12, 8, 8)])
def test_scaled_operator_matches_reference_assembly(dims):
    nx, ny, nz = dims
    p = Params(nx=nx, ny=ny, nz=nz, dx=0.1, dy=0.07, dz=0.05)
    rho = _two_phase_rho(nz, ny, nx)  # [k, j, i]
    z = RNG.standard_normal((nz, ny, nx))
    rho_f = np.transpose(rho, (2, 1, 0))  # (nx, ny, nz), Fortran cell order
    A, D, _ = assemble_projection(
        p, rho_f, np.zeros((nx + 3, ny + 2, nz + 2)), np.zeros((nx + 2, ny + 3, nz + 2)),
        np.zeros((nx + 2, ny + 2, nz + 3)), 0.37)
    zf = np.transpose(z, (2, 1, 0)).ravel(order="F")
    s = 1.0 / np.sqrt(D)
    ref = s * (A @ (s * zf))
    got = np.transpose(U.apply_As(z, rho, p.dx, p.dy, p.dz), (2, 1, 0)).ravel(order="F")
    assert np.allclose(got, ref, rtol=1e-12, atol=1e-12)


def test_gradient_matches_finite_differences():
    nz, ny, nx = 4, 8, 8
    w = RNG.standard_normal(U.nparams()) * 0.2
    r = RNG.standard_normal((nz, ny, nx))
    rho = _two_phase_rho(nz, ny, nx)
    args = (r, rho, 1.2, 1000.0, 0.1, 0.1, 0.1)
    L0, g, _ = U.loss_and_grad(w, *args)
    assert 0.0 <= L0 <= 1.0
    for seed in range(3):
        d = np.random.default_rng(100 + seed).standard_normal(w.size)
        eps = 1e-7
        fd = (U.loss(w + eps * d, *args) - U.loss(w - eps * d, *args)) / (2 * eps)
        assert np.isclose(fd, g @ d, rtol=1e-5, atol=1e-9), (fd, g @ d)
    # per layer: a bias and a kernel entry of every layer
    off = 0
    for ci, co in U.arch(3):
        for q in (off + 5 % (27 * ci * co), off + 27 * ci * co):
            e = np.zeros_like(w)
            e[q] = 1e-6
            fd = (U.loss(w + e, *args) - U.loss(w - e, *args)) / 2e-6
            assert np.isclose(fd, g[q], rtol=1e-4, atol=1e-8), (q, fd, g[q])
        off += 27 * ci * co + co


def test_loss_vanishes_for_exact_inverse():
    nz, ny, nx = 4, 4, 4
    rho = _two_phase_rho(nz, ny, nx)
    z = RNG.standard_normal((nz, ny, nx))
    r = U.apply_As(z, rho, 0.1, 0.1, 0.1)
    s = U.apply_As(z, rho, 0.1, 0.1, 0.1)
    # with z the exact solution, s = r and the loss vanishes
    rs, ss, rr = np.sum(r * s), np.sum(s * s), np.sum(r * r)
    assert abs(1 - rs * rs / (rr * ss)) < 1e-14


def test_bf16_matches_torch():
    x = np.concatenate([RNG.standard_normal(10000) * 10.0 ** RNG.integers(-6, 6, 10000),
                        [1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -1.0 - 2.0 ** -8, 0.0, 65504.0]])
    ref = torch.tensor(x, dtype=torch.float32).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(U.bf16(x), ref)
    assert U.bf16(np.array([1.0 + 2.0 ** -8]))[0] == 1.0
    assert U.bf16(np.array([1.0 + 3 * 2.0 ** -8]))[0] == 1.0 + 2.0 ** -6


def test_bf16_forward_close_to_fp64():
    nz, ny, nx = 4, 8, 8
    w = RNG.standard_normal(U.nparams()) * 0.1
    r = RNG.standard_normal((nz, ny, nx))
    rho = _two_phase_rho(nz, ny, nx)
    z64, _ = U.forward(w, U.inputs(r, rho, 1.2, 1000.0))
    z16, _ = U.forward(w, U.inputs(r, rho, 1.2, 1000.0, bf=True), bf=True)
    assert np.linalg.norm(z16 - z64) / np.linalg.norm(z64) < 3e-2


@pytest.mark.parametrize("L", [2, 4])
def test_gradient_matches_finite_differences_other_depths(L):
    nz, ny, nx = 8, 8, 16
    assert U.levels_of(U.nparams(L)) == L
    w = np.random.default_rng(L).standard_normal(U.nparams(L)) * 0.1
    r = RNG.standard_normal((nz, ny, nx))
    rho = _two_phase_rho(nz, ny, nx)
    args = (r, rho, 1.2, 1000.0, 0.1, 0.1, 0.1)
    _, g, _ = U.loss_and_grad(w, *args)
    for seed in range(2):
        d = np.random.default_rng(200 + seed).standard_normal(w.size)
        eps = 1e-7
        fd = (U.loss(w + eps * d, *args) - U.loss(w - eps * d, *args)) / (2 * eps)
        assert np.isclose(fd, g @ d, rtol=1e-5, atol=1e-9), (fd, g @ d)


# ---- the network as a multigrid level's coarse solver -------------------------

def _dense_As_faces(cx, cy, cz):
    """The matrix of apply_As_faces, column by column."""
    nz, ny, nx = cx.shape[0], cx.shape[1], cx.shape[2] - 1
    n = nx * ny * nz
    M = np.empty((n, n))
    for q in range(n):
        e = np.zeros(n)
        e[q] = 1.0
        M[:, q] = U.apply_As_faces(e.reshape(nz, ny, nx), cx, cy, cz).ravel()
    return M


def test_multigrid_faces_closed_forms():
    """Uniform density: every interior face of level l is 2^l times the
    BF16-rounded fine conductance, wall faces are zero.  Density layered along
    one axis: a coarse face normal to it equals twice the fine face at twice
    its index (so the normal index is the fine face 2I, not 2I-1)."""
    nz, ny, nx = 8, 8, 16
    dt, h = 1e-3, 1 / 16
    c0 = U.bf16(dt / (h * h) / 700.0)
    rho = np.full((nz, ny, nx), 700.0)
    for lev in range(3):
        cx, cy, cz = U.mg_faces(rho, dt, h, h, h, lev)
        for c, ax in ((cx, 2), (cy, 1), (cz, 0)):
            c = np.moveaxis(c, ax, 0)
            assert np.all(c[0] == 0) and np.all(c[-1] == 0)
            np.testing.assert_allclose(c[1:-1], (2.0 ** lev) * c0, rtol=1e-6)
    g = np.random.default_rng(0)
    for ax in range(3):
        prof = 1.2 + 998.8 * g.random(rho.shape[ax])
        shp = [1, 1, 1]
        shp[ax] = -1
        lay = np.broadcast_to(prof.reshape(shp), rho.shape).copy()
        f0 = U.mg_faces(lay, dt, h, h, h, 0)[2 - ax]
        f1 = U.mg_faces(lay, dt, h, h, h, 1)[2 - ax]
        f0n, f1n = np.moveaxis(f0, ax, 0), np.moveaxis(f1, ax, 0)
        np.testing.assert_allclose(f1n, 2.0 * f0n[::2, ::2, ::2], rtol=1e-6)


def test_multigrid_level_operator_invariants():
    """On every level the scaled operator of the coarsened faces is symmetric,
    positive semi-definite with the single null vector sqrt(D) (Neumann
    walls), so the network's coarse problem is the multigrid one."""
    g = np.random.default_rng(1)
    nz, ny, nx = 4, 4, 8
    rho = np.where(g.random((nz, ny, nx)) < 0.4, 1000.0, 1.2)
    for lev in range(2):
        cx, cy, cz = U.mg_faces(rho, 2e-4, 1 / 8, 1 / 8, 1 / 8, lev)
        M = _dense_As_faces(cx, cy, cz)
        np.testing.assert_allclose(M, M.T, atol=1e-12)
        ev = np.linalg.eigvalsh(0.5 * (M + M.T))
        assert ev[0] > -1e-12 and ev[1] > 1e-8
        D = cx[:, :, 1:] + cx[:, :, :-1] + cy[:, 1:] + cy[:, :-1] + cz[1:] + cz[:-1]
        assert np.abs(M @ np.sqrt(D).ravel()).max() < 1e-12 * np.sqrt(D).max()
        nz, ny, nx = nz // 2, ny // 2, nx // 2


def test_block_mean_of_linear_density():
    k, j, i = np.meshgrid(np.arange(8), np.arange(4), np.arange(8), indexing="ij")
    rho = 3.0 + 2.0 * i + 0.5 * j - 1.5 * k
    for lev in (1, 2):
        m = 1 << lev
        K, J, I = np.meshgrid(np.arange(8 // m), np.arange(4 // m), np.arange(8 // m), indexing="ij")
        c = lambda a: a * m + 0.5 * (m - 1)  # noqa: E731  (block centre)
        np.testing.assert_allclose(U.block_mean(rho, lev), 3.0 + 2.0 * c(I) + 0.5 * c(J) - 1.5 * c(K))


def test_face_loss_gradient_matches_finite_differences():
    """The loss through a multigrid level's operator (faces given) has the
    gradient of its finite differences."""
    g = np.random.default_rng(2)
    nz, ny, nx = 4, 4, 8
    rho = np.where(g.random((2 * nz, 2 * ny, 2 * nx)) < 0.5, 1000.0, 1.2)
    faces = U.mg_faces(rho, 2e-4, 1 / 16, 1 / 16, 1 / 16, 1)
    rhoe = U.block_mean(rho, 1)
    r = g.standard_normal((nz, ny, nx))
    w = 0.2 * g.standard_normal(U.nparams(2))
    L0, grad, _ = U.loss_and_grad(w, r, rhoe, 1.2, 1000.0, 0, 0, 0, faces=faces)
    for q in g.choice(w.size, 12, replace=False):
        e = np.zeros_like(w)
        e[q] = 1e-6
        lp = U.loss_and_grad(w + e, r, rhoe, 1.2, 1000.0, 0, 0, 0, faces=faces)[0]
        lm = U.loss_and_grad(w - e, r, rhoe, 1.2, 1000.0, 0, 0, 0, faces=faces)[0]
        fd = (lp - lm) / 2e-6
        assert abs(fd - grad[q]) <= 1e-5 * max(1.0, abs(fd)), (q, fd, grad[q])
