"""Checkpoint and restart on the GPU: a run resumed from a binary checkpoint
continues bit for bit (FP64 and mixed precision); the reference's restart
file (restart_salva, 11_diversos.f90:303-329) written by the flang-built
reference is read back exactly, and the file this library writes is read by
the reference's restart_ini (11_diversos.f90:232-301) to the same arrays."""
import os

import numpy as np
import pytest

import oracle as O
import pysuli
from helpers import consistent_state, gpu_from_state

pytestmark = pytest.mark.gpu

STATE = ["u", "v", "w", "prd1", "rho", "ls_nu", "ls", "mod_ls", "hsx", "hsy", "hsz", "hs", "kurv"]


def _tank(prec):
    p = O.Params(nx=32, ny=16, nz=16, dx=1 / 32, dy=1 / 32, dz=1 / 32)
    sv = pysuli.Solver.from_oracle_params(p, precision=prec, poisson=pysuli.POISSON_FCG,
                                          precond=pysuli.PRECOND_GMG, tol_rel=1e-8)
    return p, sv


def _columns(p, seed=3):
    """Bottom-obstacle columns for the tank: a seeded bump one to three cells
    high in the middle, and a strip of -1 columns (obstaculo's value for a
    column without obstacle, 3_contorno.f90:741-760) beside it."""
    rng = np.random.default_rng(seed)
    out = []
    for n0, n1, top in ((p.nx + 3, p.ny + 2, 2), (p.nx + 2, p.ny + 3, 2), (p.nx + 2, p.ny + 2, 3)):
        k = np.zeros((n0, n1), dtype=np.int32)
        k[n0 // 3: 2 * n0 // 3, 2: n1 - 2] = rng.integers(1, top + 1, size=(2 * n0 // 3 - n0 // 3, n1 - 4))
        k[2 * n0 // 3: 2 * n0 // 3 + 2, 2: n1 - 2] = -1
        out.append(k)
    return out


@pytest.mark.parametrize("obst", [False, True], ids=["plain", "obstacle"])
@pytest.mark.parametrize("prec", [pysuli.PREC_FP64, pysuli.PREC_MIXED])
def test_checkpoint_resumes_bit_for_bit(tmp_path, prec, obst):
    """Resume in a fresh context: with an obstacle, the checkpoint alone
    installs the columns (the fresh context never sees suli_set_obstacle);
    without one, loading removes the obstacle the fresh context was given."""
    print(f"checkpoint resume 32x16x16 prec={prec} obstacle={obst}", flush=True)
    dt = 1e-3
    path = str(tmp_path / "run.ckpt")
    p, sv = _tank(prec)
    try:
        if obst:
            sv.set_obstacle(*_columns(p))
        sv.init_dambreak(0.3, 0.3)
        sv.contorno(1)
        for _ in range(3):
            sv.step(dt)
        sv.checkpoint_save(path)
        for _ in range(3):
            sv.step(dt)
        want = {n: sv.download(n) for n in STATE}
        vol_want = sv.volumes()
    finally:
        sv.close()
    p, sv = _tank(prec)
    try:
        if not obst:  # a checkpoint without an obstacle removes the context's
            sv.set_obstacle(*_columns(p, seed=5))
        sv.checkpoint_load(path)
        for _ in range(3):
            sv.step(dt)
        for n in STATE:
            assert np.array_equal(sv.download(n), want[n]), n
        assert sv.volumes() == vol_want
    finally:
        sv.close()
    # another grid or precision refuses it
    other = pysuli.Solver(16, 16, 16, 1 / 16, precision=prec)
    try:
        with pytest.raises(pysuli.SuliError, match="differs"):
            other.checkpoint_load(path)
    finally:
        other.close()


def _ref_params():
    p = O.Params()
    if not O.ref_available(p):
        pytest.skip("flang reference harness not built here")
    return p


def test_restart_reads_reference_file(tmp_path):
    print("restart: reference file -> GPU (8x6x5, FP64)", flush=True)
    p = _ref_params()
    s = consistent_state(p, seed=13, perturb=0.4)
    s.scal["vol_ini"], s.scal["vol_ins"] = 0.0123, 0.0124
    os.makedirs(tmp_path / "dados")
    O.ref_run(s.copy(), "restart_salva", workdir=str(tmp_path))
    sv = pysuli.Solver.from_oracle_params(p, precision=pysuli.PREC_FP64)
    try:
        sv.restart_read(str(tmp_path / "dados" / "arquivorestart"))
        for n in STATE:
            assert np.array_equal(sv.download(n), s.arr(n)), n
        assert sv.volumes()[1] == 0.0124
    finally:
        sv.close()


def test_reference_reads_restart_written(tmp_path):
    print("restart: GPU file -> reference restart_ini (8x6x5, FP64)", flush=True)
    p = _ref_params()
    s = consistent_state(p, seed=17, perturb=0.4)
    sv = gpu_from_state(s, precision=pysuli.PREC_FP64)
    try:
        sv.step(1e-3)
        got = {n: sv.download(n) for n in STATE}
        os.makedirs(tmp_path / "dados")
        sv.restart_write(str(tmp_path / "dados" / "arquivorestart"))
        vol = sv.volumes()
    finally:
        sv.close()
    r = O.State(p)
    O.ref_run(r, "restart_ini", workdir=str(tmp_path))
    for n in STATE:
        assert np.array_equal(r.arr(n), got[n]), n
    assert r.scal["vol_ins"] == vol[1]
