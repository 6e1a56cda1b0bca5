"""The large sloshing tank at its full size: 2 x 1 x 1 m on 512x256x256
(33.5 M cells), water 0.4 m deep, density ratio 1000, horizontal forcing
0.05 g sin(2 pi t/1.2 s + phase), mixed precision, multigrid FCG to 1e-8.

A few steps at that size: every projection converges, the FP64 residual of
the last one, recomputed here from the downloaded fields by array stencils,
meets the tolerance; the interface volume drifts by less than 1e-6; and
the default learned preconditioner (trained on dam breaks at 256x128x128)
solves the next projection of this unseen tank to the same tolerance, its
fallbacks to multigrid counted and printed."""
import math

import numpy as np
import pytest

import oracle as O
import pysuli
from helpers import stencil_relres

pytestmark = pytest.mark.gpu

NX, NY, NZ = 512, 256, 256


def test_sloshing_512_converges_and_learned_solves_unseen():
    print(f"sloshing {NX}x{NY}x{NZ}, ratio 1000", flush=True)
    h = 2.0 / NX
    depth = 0.4
    sv = pysuli.Solver(NX, NY, NZ, h, precision=pysuli.PREC_MIXED, poisson=pysuli.POISSON_FCG,
                       precond=pysuli.PRECOND_GMG, tol_rel=1e-8, rho_f1=1.0, rho_f2=1000.0)
    try:
        z = (np.arange(1, NZ + 1) - 0.5) * h
        sv.upload("ls", np.broadcast_to(depth - z, (NX, NY, NZ)).copy(order="F"))
        vol0 = sv.level_set_ini()
        sv.contorno(1)
        sv.contorno(3)
        sv.set_forcing(0.05 * 9.81, 1.2, 0.7)
        dt = 0.2 * h / math.sqrt(9.81 * depth)
        for step in range(1, 6):
            _, ps = sv.step(dt)
            print(f"  step {step}: " + ", ".join(f"{q['iters']} it {q['ms']:.1f} ms relres {q['res']:.2e}"
                                                  for q in ps), flush=True)
            assert all(q["converged"] and q["res"] <= 1e-8 for q in ps), ps
        drift = (sv.volumes()[1] - vol0) / vol0
        print(f"  volume drift after 5 steps {drift:.2e}", flush=True)
        assert abs(drift) < 1e-6
        # the next projection: multigrid, then the learned default from the same input
        sv.level_set(dt)
        sv.contorno(3)
        sv.visco()
        sv.convdiff()
        sv.tempo(1, 0.5 * dt)
        sv.contorno(2)
        p0 = sv.download("prd1")
        fields = {n: sv.download(n) for n in ("u", "v", "w", "rho")}
        p = O.Params(nx=NX, ny=NY, nz=NZ, dx=h, dy=h, dz=h)
        sv.set_guess(0)
        res = {}
        models = pysuli.shipped_learned()
        cases = [("gmg", pysuli.PRECOND_GMG, None)] + ([("learned", pysuli.PRECOND_LEARNED, models[0])]
                                                        if models else [])
        for name, pc, meta in cases:
            if meta is not None:
                pysuli.use_learned(sv, meta)
            sv.upload("prd1", p0)
            sv.set_solver(pysuli.POISSON_FCG, pc)
            st = sv.graddin(0.5 * dt)
            rr = stencil_relres(p, fields["rho"], fields["u"], fields["v"], fields["w"], 0.5 * dt,
                                sv.download("prd1"))
            res[name] = (st, rr)
            print(f"  {name}: {st['iters']} iterations, {st['outer']} passes, {st['fallbacks']} fallbacks, "
                  f"{st['ms']:.1f} ms, solver relres {st['res']:.2e}, host relres {rr:.2e}", flush=True)
            assert st["converged"] == 1 and rr <= 1.5e-8, (name, st, rr)
        assert res["gmg"][0]["fallbacks"] == 0
    finally:
        sv.close()
