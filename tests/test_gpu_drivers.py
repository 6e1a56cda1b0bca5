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


def test_fortran_driver_runs_the_time_loop(tmp_path):
    exe = os.path.join(PKG, "principal_gpu")
    if not os.path.exists(exe):
        pytest.skip("principal_gpu not built (flang absent)")
    print("fortran driver 64x32x32, 20 steps", flush=True)
    out = subprocess.run([exe, "64", "32", "32", "20", "1", "0.25", "0.25", "10"], capture_output=True, text=True,
                         timeout=240, cwd=str(tmp_path))
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]
    steps = re.findall(r"step\s+(\d+) t=\s*(\S+) vol_err=\s*(\S+) maxdiv=\s*(\S+) poisson_iters=\s*(\d+)", out.stdout)
    assert len(steps) == 20
    for it, t, vol_err, maxdiv, iters in steps:
        assert abs(float(vol_err)) < 1e-3 and float(maxdiv) < 1e-3 and 0 < int(iters) < 50
    frames = re.findall(r"frame\s+(\d+) max\|u\|=\s*(\S+)", out.stdout)
    assert [int(f[0]) for f in frames] == [10, 20] and all(0 < float(f[1]) < 5 for f in frames)
    assert "Mcell-steps/s=" in out.stdout
    # field output: the initial record and one per frame, each one Fortran
    # record of the t_plot = 1 layout, and the ParaView index over them
    nx, ny, nz = 64, 32, 32
    n1, n = (nx + 1) * (ny + 1) * (nz + 1), nx * ny * nz
    nval = 6 * n1 + 6 * n + (nx + 1) * ny * nz + nx * (ny + 1) * nz + nx * ny * (nz + 1)
    for q in (10, 11, 12):
        raw = open(tmp_path / "arquivos" / f"campos_{q:05d}", "rb").read()
        assert len(raw) == 4 * nval + 8
        assert np.frombuffer(raw[:4], "<i4")[0] == 4 * nval == np.frombuffer(raw[-4:], "<i4")[0]
        assert np.all(np.isfinite(np.frombuffer(raw[4:-4], "<f4")))
    import xml.etree.ElementTree as ET
    text = open(tmp_path / "arquivos" / "visu.xdmf").read()
    grids = ET.fromstring(text.split("\n", 2)[2]).find("Domain").find("Grid").findall("Grid")
    assert [g.get("Name") for g in grids] == ["campos_00010", "campos_00011", "campos_00012"]


@pytest.mark.parametrize("t_tempo", [0, 2, 3])
def test_fortran_driver_time_schemes(tmp_path, t_tempo):
    """The driver with t_tempo = 0 (Euler), 2 (RK3), 3 (AB2): ntt stages per
    step (2_iniciais.f90:85-102), volume and divergence as with RK2."""
    exe = os.path.join(PKG, "principal_gpu")
    if not os.path.exists(exe):
        pytest.skip("principal_gpu not built (flang absent)")
    print(f"fortran driver 64x32x32, 10 steps, t_tempo {t_tempo}", flush=True)
    out = subprocess.run([exe, "64", "32", "32", "10", "1", "0.25", "0.25", "0", str(t_tempo)],
                         capture_output=True, text=True, timeout=240, cwd=str(tmp_path))
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]
    steps = re.findall(r"step\s+(\d+) t=\s*(\S+) vol_err=\s*(\S+) maxdiv=\s*(\S+) poisson_iters=\s*(\d+)", out.stdout)
    assert len(steps) == 10
    for it, t, vol_err, maxdiv, iters in steps:
        assert abs(float(vol_err)) < 1e-3 and float(maxdiv) < 1e-3 and 0 < int(iters) < 50
    m = re.search(r"poisson_iters/solve=\s*(\S+)", out.stdout)
    assert m, out.stdout[-500:]


def test_fortran_driver_obstacle(tmp_path):
    """The driver with the reference's obstacle type 7 (a hump rising at
    x = 0.88 m, 3_contorno.f90:639-652) from suli_obstacle_columns: the dam
    break runs over it.  The projection does not see the obstacle (the
    reference zeroes the velocities of its columns afterwards, 3_contorno.f90:
    266-291), so the divergence next to it is not small: only bounded."""
    exe = os.path.join(PKG, "principal_gpu")
    if not os.path.exists(exe):
        pytest.skip("principal_gpu not built (flang absent)")
    print("fortran driver 64x32x32, 10 steps, obst_t 7", flush=True)
    out = subprocess.run([exe, "64", "32", "32", "10", "1", "0.25", "0.25", "0", "1", "7"],
                         capture_output=True, text=True, timeout=240, cwd=str(tmp_path))
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]
    steps = re.findall(r"step\s+(\d+) t=\s*(\S+) vol_err=\s*(\S+) maxdiv=\s*(\S+) poisson_iters=\s*(\d+)", out.stdout)
    assert len(steps) == 10
    for it, t, vol_err, maxdiv, iters in steps:
        assert abs(float(vol_err)) < 1e-2 and float(maxdiv) < 10.0 and 0 < int(iters) < 50


@pytest.mark.parametrize("mms_t", [1, 2])
def test_fortran_driver_manufactured_solution(tmp_path, mms_t):
    """The driver's manufactured-solution run (13_mms.f90 through the shim's
    mms_i, termo_fonte1/2, mms and contorno's planes; FP64, no projection)
    against the oracle's run of the same case: the errors of mms() per step."""
    import math

    import oracle as O
    exe = os.path.join(PKG, "principal_gpu")
    if not os.path.exists(exe):
        pytest.skip("principal_gpu not built (flang absent)")
    nx, ny, nz, ts = 16, 12, 10, 5
    print(f"fortran driver manufactured solution mms_t {mms_t}, {nx}x{ny}x{nz}, {ts} steps", flush=True)
    out = subprocess.run([exe, str(nx), str(ny), str(nz), str(ts), "1", "0.25", "0.25", "0", "1", "0", str(mms_t)],
                         capture_output=True, text=True, timeout=240, cwd=str(tmp_path))
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]
    rows = re.findall(r"mms\s+(\d+) t=\s*(\S+) err=\s*(\S+)\s+(\S+)\s+(\S+)\s+(\S+)", out.stdout)
    assert len(rows) == ts
    got = np.array([[float(x) for x in r[2:]] for r in rows])
    p = O.Params(nx=nx, ny=ny, nz=nz, dx=2 * math.pi / nx, dy=2 * math.pi / ny, dz=2.5 / nz, ccx0=3, ccxf=3,
                 ccy0=3, ccyf=3, ccz0=3, cczf=3, m_turb=0, gx=0.0, gz=0.0, mms_t=mms_t, mms_a=0.0, mms_h0=2.0,
                 mms_coef=1.0 if mms_t == 2 else 0.0)
    _, errs = O.weng_run(p, 0.05 * p.dx, ts)
    ref = np.array(errs[1:])
    print("driver errors", got[-1], "oracle", ref[-1])
    assert np.allclose(got[:, :3], ref[:, :3], rtol=1e-9, atol=1e-14)
    if mms_t == 2:
        assert np.allclose(got[:, 3], ref[:, 3], rtol=1e-9, atol=1e-14)


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
