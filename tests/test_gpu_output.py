"""Field output on the GPU (suli_frame_record, suli_write_frame): the record
built by k_frame_record from the resident fields against the reference's own
file (golden fixture, bit for bit in FP64 mode, record markers included) and
against the oracle (oracle.plot_record) on the fields the GPU holds after a
time step, at a grid size with partial tiles; FP32 fields within 1e-6 of
each array's largest magnitude."""
import os

import numpy as np
import pytest

import oracle as O
import pysuli
from helpers import gpu_from_state
from test_output import fixture_state, record_of_file, sizes

pytestmark = pytest.mark.gpu


def test_record_matches_reference_file(tmp_path):
    print("frame record vs reference file (8x6x5, FP64)", flush=True)
    s, d = fixture_state()
    want = record_of_file(d["record_file"])
    sv = gpu_from_state(s, precision=pysuli.PREC_FP64)
    try:
        got = sv.frame_record(1)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
        n1, n = sizes(s.p)
        assert np.array_equal(sv.frame_record(0), want[:6 * n1 + 2 * n])
        path = tmp_path / "campos_00011"
        sv.write_frame(str(path), 1)
        assert open(path, "rb").read() == bytes(d["record_file"])
        with pytest.raises(pysuli.SuliError):
            sv.frame_record(2)
    finally:
        sv.close()


@pytest.mark.parametrize("prec", [pysuli.PREC_FP64, pysuli.PREC_MIXED])
def test_record_after_a_step_matches_oracle(prec):
    """A dam break on 40x24x20 (partial 64x4 blocks in x and y) advanced one
    step on the GPU; the oracle builds the record from the fields the GPU
    holds."""
    print(f"frame record after a step 40x24x20 prec={prec}", flush=True)
    p = O.Params(nx=40, ny=24, nz=20, dx=1 / 40, dy=1 / 40, dz=1 / 40)
    sv = pysuli.Solver.from_oracle_params(p, precision=prec, poisson=pysuli.POISSON_FCG,
                                          precond=pysuli.PRECOND_GMG, tol_rel=1e-8)
    try:
        sv.init_dambreak(0.35, 0.3)
        sv.contorno(1)
        sv.step(2e-3)
        s = O.State(p)
        for name in ("u", "v", "w", "ls", "prd1", "nut"):
            s.arr(name)[...] = sv.download(name)
        want = O.plot_record(s, 1)
        got = sv.frame_record(1)
        assert got.size == want.size
        if prec == pysuli.PREC_FP64:
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
        else:
            n1, n = sizes(p)
            bounds = np.cumsum([0] + [n1] * 6 + [n] * 6 + [(p.nx + 1) * p.ny * p.nz, p.nx * (p.ny + 1) * p.nz,
                                                       p.nx * p.ny * (p.nz + 1)])
            for a, b in zip(bounds[:-1], bounds[1:]):
                sc = max(float(np.abs(want[a:b]).max()), 1e-30)
                assert float(np.abs(got[a:b] - want[a:b]).max()) <= 1e-6 * sc, (a, b)
    finally:
        sv.close()
