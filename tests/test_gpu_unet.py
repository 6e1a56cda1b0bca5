"""GPU tests of the learned preconditioner: the BF16 MFMA network and its
backward pass against the NumPy oracle (oracle/unet_ref.py), the training
loop, FCG(4) with the network to the FP64 tolerance, the fallback path, and
the ablation trained vs untrained vs plain Jacobi.

Tolerances: with the oracle rounding to BF16 at the same points as the GPU
(bf16=True) the remaining differences are FP32 accumulation order and the
rounding flips it causes: network output within 1e-2 relative L2, weight
gradient within 5e-2 (per layer 0.1); against the plain FP64 oracle the
gradient direction within cosine 0.98."""
import numpy as np
import pytest

import pysuli as S
import unet_ref as U

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def two_phase(nz, ny, nx, seed=3):
    g = np.random.default_rng(seed)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    phi = 0.35 * nx - np.hypot(i - 0.3 * nx, k - 0.2 * nz) + 0.5 * g.standard_normal((nz, ny, nx))
    return 1.2 + (1000.0 - 1.2) * np.clip(0.5 + phi / 3.0, 0, 1)


def solver(nx, ny, nz, **kw):
    kw.setdefault("precision", S.PREC_MIXED)
    kw.setdefault("poisson", S.POISSON_FCG)
    kw.setdefault("precond", S.PRECOND_GMG)
    kw.setdefault("tol_rel", 1e-8)
    return S.Solver(nx, ny, nz, 1.0 / nx, **kw)


# grids below 2^18 voxels take the row-group kernels, 64^3 the LDS-tiled ones
@pytest.mark.parametrize("L,dims", [(3, (32, 16, 16)), (3, (20, 12, 8)), (5, (32, 16, 16)), (2, (12, 8, 8)),
                                    (2, (64, 64, 64))])
def test_network_and_gradient_match_oracle(L, dims):
    nx, ny, nz = dims
    print(f"unet parity L={L} dims={dims}", flush=True)
    s = solver(nx, ny, nz)
    try:
        s.unet_configure(L)
        s.unet_init(11)
        w = s.unet_get_weights().astype(np.float64)
        assert w.size == U.nparams(L) == s.unet_nparams()
        r = np.random.default_rng(5).standard_normal((nz, ny, nx))
        rho = two_phase(nz, ny, nx)
        loss, g, z = s.unet_loss_grad(r, rho)
        lb, gb, zb = U.loss_and_grad(w, r, rho, 1.2, 1000.0, 1 / nx, 1 / nx, 1 / nx, bf=True)
        _, g64, _ = U.loss_and_grad(w, r, rho, 1.2, 1000.0, 1 / nx, 1 / nx, 1 / nx)
        assert rel(z, zb) < 1e-2
        assert abs(loss - lb) < 1e-3
        assert rel(g, gb) < 5e-2
        off = 0
        for ci, co in U.arch(L):
            n = 27 * ci * co + co
            assert rel(g[off:off + n], gb[off:off + n]) < 0.1, (ci, co)
            off += n
        assert g @ g64 / (np.linalg.norm(g) * np.linalg.norm(g64)) > 0.98
    finally:
        s.close()


def test_weights_roundtrip_and_errors():
    s = solver(16, 8, 8)
    try:
        n = s.unet_nparams()
        assert n == U.nparams(3)
        w = np.random.default_rng(0).standard_normal(n).astype(np.float32)
        s.unet_set_weights(w)
        assert np.array_equal(s.unet_get_weights(), w)
        with pytest.raises(S.SuliError):
            s.unet_set_weights(w[:-1])
        with pytest.raises(S.SuliError):
            s.unet_configure(7)
        # the learned preconditioner needs weights: a fresh configuration has none
        s.unet_configure(4)
        s.init_dambreak(0.3, 0.3)
        s.set_solver(S.POISSON_FCG, S.PRECOND_LEARNED)
        with pytest.raises(S.SuliError, match="no weights"):
            s.graddin(1e-3)
        with pytest.raises(S.SuliError, match="empty"):
            s.unet_init(0)
            s.unet_train(1, 1e-3)
    finally:
        s.close()


def _dambreak(nx, L):
    ny = nz = nx // 2
    s = solver(nx, ny, nz)
    s.init_dambreak(0.3, 0.35)
    dt = 0.2 / nx / (2 * np.sqrt(9.81 * 0.35))
    s.contorno(1)
    s.contorno(3)
    s.unet_configure(L)
    return s, dt


def _stage_state(s, dt):
    """Advance to the projection of a first RK stage; returns prd1 to restore."""
    s.level_set(dt)
    s.contorno(3)
    s.visco()
    s.convdiff()
    s.tempo(1, 0.5 * dt)
    s.contorno(2)
    return s.download("prd1")


def _solve(s, pc, pr0, dt):
    s.upload("prd1", pr0)
    s.set_solver(S.POISSON_FCG, pc)
    return s.graddin(0.5 * dt)


def test_training_ablation():
    """Trained on the solver's own residuals (GMG solves, then the learned
    solver's trajectories), the network cuts the FCG iterations well below
    Jacobi and below the untrained network, converges to the FP64 tolerance
    without falling back, and the training loss drops."""
    print("unet ablation 32x16x16 L=4", flush=True)
    s, dt = _dambreak(32, 4)
    try:
        s.unet_init(1)
        s.unet_capture(3000, 1)
        for _ in range(6):
            s.step(dt)
        s.unet_capture(0)
        n0 = s.unet_dataset_size()
        assert n0 > 20
        first = s.unet_train(1, 1e-3, 0)
        loss = s.unet_train(1500, 1e-3, 0)
        assert loss < 0.5 * first
        for rnd in range(4):
            s.set_solver(S.POISSON_FCG, S.PRECOND_LEARNED)
            s.unet_capture(3000, 2)
            for _ in range(2):
                s.step(dt)
            s.unet_capture(0)
            loss = s.unet_train(1000, 1e-3 * 0.5 ** (rnd / 2), rnd + 1)
        assert s.unet_dataset_size() > n0
        trained = s.unet_get_weights()
        pr0 = _stage_state(s, dt)
        jac = _solve(s, S.PRECOND_JACOBI, pr0, dt)
        lrn = _solve(s, S.PRECOND_LEARNED, pr0, dt)
        s.unet_init(1)  # the same network untrained
        unt = _solve(s, S.PRECOND_LEARNED, pr0, dt)
        s.unet_set_weights(trained)
        print(f"jacobi {jac}\nlearned {lrn}\nuntrained {unt}", flush=True)
        for st in (jac, lrn, unt):
            assert st["converged"] == 1 and st["res"] <= 1e-8
        assert lrn["fallbacks"] == 0
        assert lrn["iters"] < 0.5 * jac["iters"]
        assert unt["fallbacks"] >= 1 or unt["iters"] > 2 * lrn["iters"]
    finally:
        s.close()


# ---- the network as the coarse solver of a multigrid level --------------------

@pytest.mark.parametrize("entry,L,prec", [(1, 3, S.PREC_MIXED), (2, 2, S.PREC_MIXED), (1, 2, S.PREC_FP64)])
def test_level_samples_and_gradient_match_oracle(entry, L, prec):
    """Samples captured at multigrid level `entry` hold that level's operator
    (the BF16 finest faces coarsened, as oracle.unet_ref.mg_faces, bit for
    bit) and block-mean density; the loss, gradient and output on such a
    sample match the oracle through the same face operator."""
    print(f"unet level samples entry={entry} L={L} prec={prec}", flush=True)
    nx = 32
    s = solver(nx, nx // 2, nx // 2, precision=prec)
    try:
        s.init_dambreak(0.3, 0.35)
        dt = 0.2 / nx / (2 * np.sqrt(9.81 * 0.35))
        s.contorno(1)
        s.contorno(3)
        s.unet_configure(L)
        s.unet_set_entry(entry)
        assert s.unet_entry() == entry
        s.unet_init(3)
        pr0 = _stage_state(s, dt)
        s.unet_capture(40, 1)
        st = _solve(s, S.PRECOND_GMG, pr0, dt)
        s.unet_capture(0)
        assert st["converged"] == 1
        n = s.unet_dataset_size()
        assert n >= 2
        assert s.unet_dataset_dims() == (nx >> entry, (nx // 2) >> entry, (nx // 2) >> entry, entry)
        rho = s.download("rho").transpose(2, 1, 0)
        ref = U.mg_faces(rho, 0.5 * dt, 1 / nx, 1 / nx, 1 / nx, entry)
        w = s.unet_get_weights().astype(np.float64)
        for q in (0, n - 1):
            r, rhoe, faces = s.unet_dataset_get(q)
            np.testing.assert_allclose(rhoe, U.block_mean(rho, entry), rtol=1e-6)
            for a, b in zip(faces, ref):
                assert np.array_equal(a.astype(np.float64), b)
            assert np.all(np.isfinite(r)) and np.abs(r).max() > 0
            loss, g, z = s.unet_sample_loss_grad(q)
            lb, gb, zb = U.loss_and_grad(w, r, rhoe, 1.2, 1000.0, 0, 0, 0, bf=True, faces=faces)
            assert rel(z, zb) < 1e-2
            assert abs(loss - lb) < 1e-3
            assert rel(g, gb) < 5e-2
        # the entry level must exist and the grid must fit the network there
        s.unet_set_entry(4)
        s.set_solver(S.POISSON_FCG, S.PRECOND_LEARNED)
        with pytest.raises(S.SuliError):
            _solve(s, S.PRECOND_LEARNED, pr0, dt)
        with pytest.raises(S.SuliError):
            s.unet_set_entry(5)
    finally:
        s.close()


def test_coarse_entry_ablation():
    """The network as the coarse solver of level 1: trained on the level's
    problems met by multigrid and then by its own solves, FCG converges to
    the FP64 tolerance without fallback in far fewer iterations than Jacobi;
    the same V-cycle without the network (all weights zero: no coarse
    correction) stalls or needs many more."""
    print("unet coarse-entry ablation 64x32x32 entry=1 L=3", flush=True)
    s, dt = _dambreak(64, 3)
    try:
        s.unet_set_entry(1)
        s.unet_init(1)
        s.unet_capture(3000, 1)
        for _ in range(6):
            s.step(dt)
        s.unet_capture(0)
        n0 = s.unet_dataset_size()
        assert n0 > 20
        first = s.unet_train(1, 1e-3, 0)
        loss = s.unet_train(1500, 1e-3, 0)
        assert loss < 0.5 * first
        for rnd in range(3):
            s.set_solver(S.POISSON_FCG, S.PRECOND_LEARNED)
            s.unet_capture(3000, 2)
            for _ in range(2):
                s.step(dt)
            s.unet_capture(0)
            loss = s.unet_train(800, 1e-3 * 0.5 ** (rnd / 2), rnd + 1)
        assert s.unet_dataset_size() > n0
        trained = s.unet_get_weights()
        pr0 = _stage_state(s, dt)
        jac = _solve(s, S.PRECOND_JACOBI, pr0, dt)
        lrn = _solve(s, S.PRECOND_LEARNED, pr0, dt)
        s.unet_set_weights(np.zeros_like(trained))
        zero = _solve(s, S.PRECOND_LEARNED, pr0, dt)
        s.unet_set_weights(trained)
        print(f"jacobi {jac}\nlearned {lrn}\nno network {zero}", flush=True)
        for st in (jac, lrn, zero):
            assert st["converged"] == 1 and st["res"] <= 1e-8
        assert lrn["fallbacks"] == 0
        assert lrn["iters"] < 0.25 * jac["iters"]
        assert zero["fallbacks"] >= 1 or zero["iters"] > 2 * lrn["iters"]
    finally:
        s.close()


def test_shipped_default_meets_target():
    """The default shipped network at 256x128x128 on a column geometry it was
    not trained on (seed 0): the projection converges to the FP64 relative
    residual 1e-8 without fallback in at least 3x less time than the plain
    (Jacobi) FCG, and in at most a quarter of its iterations."""
    from train_unet import Tank

    print("shipped default learned preconditioner 256x128x128 seed 0", flush=True)
    tank = Tank(256)
    s = tank.s
    try:
        tank.reset(0)
        tank.steps(4, S.PRECOND_GMG)
        dt = tank.dt
        pr0 = _stage_state(s, dt)
        S.use_learned(s, S.shipped_learned()[0])
        res = {}
        for name, pc in (("jacobi", S.PRECOND_JACOBI), ("learned", S.PRECOND_LEARNED)):
            runs = [_solve(s, pc, pr0, dt) for _ in range(3)]
            res[name] = (sorted(r["ms"] for r in runs)[1], runs[-1])
        print({k: (round(v[0], 2), v[1]["iters"], v[1]["fallbacks"]) for k, v in res.items()}, flush=True)
        for _, st in res.values():
            assert st["converged"] == 1 and st["res"] <= 1e-8
        assert res["learned"][1]["fallbacks"] == 0
        assert res["learned"][1]["iters"] <= 0.25 * res["jacobi"][1]["iters"]
        assert res["jacobi"][0] >= 3.0 * res["learned"][0]
    finally:
        s.close()
