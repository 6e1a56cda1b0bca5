"""End-to-end drivers on the GPU: the Fortran time loop over the C ABI
(suli-openmp_amd/principal_gpu, the reference's CALL sequence) and the
sloshing tank with its time-dependent body force."""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import pysuli as S

pytestmark = pytest.mark.gpu
PKG = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "suli-openmp_amd")


def test_fortran_driver_runs_the_time_loop():
    exe = os.path.join(PKG, "principal_gpu")
    if not os.path.exists(exe):
        pytest.skip("principal_gpu not built (flang absent)")
    print("fortran driver 64x32x32, 20 steps", flush=True)
    out = subprocess.run([exe, "64", "32", "32", "20", "1", "0.25", "0.25", "10"], capture_output=True, text=True,
                         timeout=240)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]
    steps = re.findall(r"step\s+(\d+) t=\s*(\S+) vol_err=\s*(\S+) maxdiv=\s*(\S+) poisson_iters=\s*(\d+)", out.stdout)
    assert len(steps) == 20
    for it, t, vol_err, maxdiv, iters in steps:
        assert abs(float(vol_err)) < 1e-3 and float(maxdiv) < 1e-3 and 0 < int(iters) < 50
    frames = re.findall(r"frame\s+(\d+) max\|u\|=\s*(\S+)", out.stdout)
    assert [int(f[0]) for f in frames] == [10, 20] and all(0 < float(f[1]) < 5 for f in frames)
    assert "Mcell-steps/s=" in out.stdout


def test_sloshing_forcing_drives_the_tank():
    """A flat free surface at rest stays at rest without forcing and is set
    in motion by gx(t); the volume is conserved either way."""
    print("sloshing 64x32x32", flush=True)
    sys.path.insert(0, PKG)
    nx, ny, nz = 64, 32, 32
    h = 2.0 / nx
    speeds = []
    for amp in (0.0, 0.05 * 9.81):
        s = S.Solver(nx, ny, nz, h, precision=S.PREC_MIXED, poisson=S.POISSON_FCG, precond=S.PRECOND_GMG,
                     tol_rel=1e-8)
        try:
            z = (np.arange(1, nz + 1) - 0.5) * h
            s.upload("ls", np.broadcast_to(0.4 - z, (nx, ny, nz)).copy(order="F"))
            vol0 = s.level_set_ini()
            s.contorno(1)
            s.contorno(3)
            s.set_forcing(amp, 1.2, np.pi / 2)
            dt = 0.2 * h / np.sqrt(9.81 * 0.4)
            for _ in range(20):
                s.step(dt)
            speeds.append(float(np.abs(s.download("u")).max()))
            assert abs(s.volumes()[1] - vol0) / vol0 < 1e-6
        finally:
            s.close()
    print("max |u| without / with forcing:", speeds, flush=True)
    assert speeds[0] < 1e-6 and speeds[1] > 0.01  # measured 5e-11 and 0.030
