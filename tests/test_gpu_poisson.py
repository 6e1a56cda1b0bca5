"""Mixed-precision projection (FP32 FCG inside FP64 iterative refinement):
the FP64 true relative residual reaches the requested tolerance, computed
independently here from the downloaded fields (scipy.sparse assembly of the
reference operator), and the pressure agrees with a direct solve."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

import oracle as O
import pysuli
from helpers import assemble_projection, consistent_state, gpu_from_state, scaled_relres

pytestmark = pytest.mark.gpu


def _stage_state(p, seed):
    st = consistent_state(p, seed=seed, vel=0.02)
    st.scal["dt"] = 1e-3
    st.run("convdiff")
    st.run("tempo", 1)
    st.run("contorno", 2)
    return st


@pytest.mark.parametrize("precision", [pysuli.PREC_FP64, pysuli.PREC_MIXED])
@pytest.mark.parametrize("precond", [pysuli.PRECOND_JACOBI, pysuli.PRECOND_GMG])
@pytest.mark.parametrize("tol", [1e-8, 1e-10])
def test_fcg_true_residual(precision, precond, tol):
    p = O.Params(nx=32, ny=16, nz=16, dx=1 / 32, dy=1 / 32, dz=1 / 32)
    st = _stage_state(p, seed=21)
    sv = gpu_from_state(st, precision=precision, poisson=pysuli.POISSON_FCG, precond=precond, tol_rel=tol)
    info = sv.graddin(1e-3)
    pres = sv.download("prd1")
    rho = sv.download("rho")
    u, v, w = sv.download("u"), sv.download("v"), sv.download("w")
    sv.close()
    A, D, q = assemble_projection(p, rho, u, v, w, 1e-3)
    pc = pres[1:-1, 1:-1, 1:-1].ravel(order="F")
    rr = scaled_relres(A, D, q, pc)
    print(f"precision {precision} precond {precond} tol {tol}: {info} host relres {rr:.3e}")
    assert info["converged"] == 1, info
    assert rr <= 1.5 * tol, rr
    # against a direct solve (one cell pinned: the operator's null space is constants)
    N = A.shape[0]
    keep = np.arange(1, N)
    x = np.zeros(N)
    x[keep] = spla.spsolve(A[keep][:, keep].tocsc(), q[keep])
    dd = (pc - pc.mean()) - (x - x.mean())
    assert np.max(np.abs(dd)) <= 1e-5 * np.max(np.abs(x - x.mean())), np.max(np.abs(dd))


def test_gmg_beats_jacobi_iterations():
    p = O.Params(nx=64, ny=32, nz=32, dx=1 / 64, dy=1 / 64, dz=1 / 64)
    st = _stage_state(p, seed=22)
    st.prd1[...] = 0.0  # cold start
    its = {}
    for pc in (pysuli.PRECOND_JACOBI, pysuli.PRECOND_GMG):
        sv = gpu_from_state(st, precision=pysuli.PREC_MIXED, poisson=pysuli.POISSON_FCG, precond=pc, tol_rel=1e-8)
        its[pc] = sv.graddin(1e-3)
        sv.close()
    print(its)
    assert its[pysuli.PRECOND_GMG]["converged"] and its[pysuli.PRECOND_JACOBI]["converged"]
    assert its[pysuli.PRECOND_GMG]["iters"] * 5 < its[pysuli.PRECOND_JACOBI]["iters"]
