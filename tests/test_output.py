"""Field output (plot_f, 4_plot.f90:260-363) and its ParaView index (visu,
12_paraview_3D.f90:1-179), on the CPU.

The oracle's record (oracle.plot_record) is pinned bit for bit against the
file the flang-built reference writes (tests/golden/output/g8_frame.npz, made
by tests/golden/output/gen_output.py; and a live run where the reference is
built here) and by closed forms; the library's XDMF writer (host code, no
GPU) is checked to be well-formed XML with the reference's grids, offsets and
dimensions."""
import math
import os
import xml.etree.ElementTree as ET

import numpy as np
import pytest

import oracle as O
import pysuli
from helpers import consistent_state

GOLD = os.path.join(os.path.dirname(__file__), "golden", "output", "g8_frame.npz")


def fixture_state():
    d = np.load(GOLD)
    nx, ny, nz, ccx0, ccxf, ccy0, ccyf, ccz0, cczf = [int(x) for x in d["__params"]]
    dx, dy, dz = [float(x) for x in d["__h"]]
    p = O.Params(nx=nx, ny=ny, nz=nz, dx=dx, dy=dy, dz=dz, ccx0=ccx0, ccxf=ccxf, ccy0=ccy0, ccyf=ccyf,
                 ccz0=ccz0, cczf=cczf)
    s = O.State(p)
    for name, _ in O.ARRAYS:  # the fixture predates the RK3/AB2 arrays: those stay zero
        if "in_" + name in d:
            s.arr(name)[...] = d["in_" + name]
    s.scal = dict(zip(O.SCALARS, [float(x) for x in d["__scal_in"]]))
    return s, d


def record_of_file(raw):
    raw = bytes(raw)
    n = np.frombuffer(raw[:4], "<i4")[0]
    assert n == len(raw) - 8 and np.frombuffer(raw[-4:], "<i4")[0] == n
    return np.frombuffer(raw[4:-4], "<f4")


def sizes(p):
    n1 = (p.nx + 1) * (p.ny + 1) * (p.nz + 1)
    n = p.nx * p.ny * p.nz
    return n1, n


def test_oracle_record_matches_reference_fixture():
    s, d = fixture_state()
    want = record_of_file(d["record_file"])
    got = O.plot_record(s, 1)
    assert got.dtype == np.float32 and got.size == want.size
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    n1, n = sizes(s.p)
    assert np.array_equal(O.plot_record(s, 0), got[:6 * n1 + 2 * n])


@pytest.mark.skipif(not O.ref_available(O.Params(nx=12, ny=8, nz=6, dx=1 / 12, dy=1 / 12, dz=1 / 12, ccx0=0,
                                                  ccxf=0, cczf=1)), reason="flang reference harness not built here")
def test_oracle_record_matches_live_reference(tmp_path):
    p = O.Params(nx=12, ny=8, nz=6, dx=1 / 12, dy=1 / 12, dz=1 / 12, ccx0=0, ccxf=0, cczf=1)
    s = consistent_state(p, seed=5, perturb=0.3)
    os.makedirs(tmp_path / "arquivos")
    O.ref_run(s.copy(), "plot", workdir=str(tmp_path))
    want = record_of_file(open(tmp_path / "arquivos" / "campos_00011", "rb").read())
    assert np.array_equal(O.plot_record(s, 1).view(np.uint32), want.view(np.uint32))


def test_oracle_record_closed_forms():
    """Coordinates (i-1) h; node averages of a linear field are its node
    values; a solid-body rotation about z has vorticity (0, 0, 2 omega);
    a uniform eddy viscosity gives uniform face values."""
    p = O.Params(nx=8, ny=6, nz=4, dx=0.25, dy=0.25, dz=0.25)
    s = O.State(p)
    om = 1.5
    # u on x faces at (i dx - dx... ) reference index i -> x = (i-1) dx, y = (j-0.5) dy
    i = np.arange(p.nx + 3)[:, None, None]
    j = np.arange(p.ny + 2)[None, :, None]
    s.u[...] = -om * ((j - 0.5) * p.dy - 0.7) + 0 * i
    iv = np.arange(p.nx + 2)[:, None, None]
    s.v[...] = om * ((iv - 0.5) * p.dx - 0.4) + 0 * np.arange(p.ny + 3)[None, :, None]
    s.w[...] = 0.0
    s.nut[...] = 3.0
    r = O.plot_record(s, 1)
    n1, n = sizes(p)
    X = r[:n1].reshape(p.nx + 1, p.ny + 1, p.nz + 1, order="F")
    assert np.allclose(X[:, 0, 0], np.arange(p.nx + 1) * p.dx, atol=1e-6)
    Z = r[2 * n1:3 * n1].reshape(p.nx + 1, p.ny + 1, p.nz + 1, order="F")
    assert np.allclose(Z[0, 0, :], np.arange(p.nz + 1) * p.dz, atol=1e-6)
    uaux = r[3 * n1:4 * n1].reshape(p.nx + 1, p.ny + 1, p.nz + 1, order="F")
    jn = np.arange(1, p.ny + 2)[None, :, None]  # node j sits at y = (j-1) dy
    assert np.allclose(uaux, np.broadcast_to(-om * ((jn - 1) * p.dy - 0.7), uaux.shape), atol=1e-6)
    o = 6 * n1
    vortk = r[o + 5 * n:o + 6 * n]
    assert np.allclose(vortk, 2 * om, rtol=1e-6)
    assert np.allclose(r[o + 3 * n:o + 5 * n], 0.0, atol=1e-6)
    assert np.all(r[o + 6 * n:] == 3.0)
    assert np.all(r[o + n:o + 2 * n] == 0.0)  # kaux: no obstacles


def test_xdmf_matches_reference_layout(tmp_path):
    s, d = fixture_state()
    ref = str(d["xdmf"])
    dt = s.scal["dt"]
    last = 20000 // math.ceil(1.0 / dt) + 10  # filen of visu with ts = ceiling(20/0.001), dt_frame = 1
    path = tmp_path / "visu.xdmf"
    pysuli.write_xdmf(str(path), s.p.nx, s.p.ny, s.p.nz, 10, last, 1, 1.0)
    text = open(path).read()
    root = ET.fromstring(text.split("\n", 2)[2])  # past the XML declaration and DOCTYPE
    grids = root.find("Domain").find("Grid").findall("Grid")
    assert [g.get("Name") for g in grids] == [f"campos_{q:05d}" for q in range(10, last + 1)]
    for g in grids:
        items = [e for e in g.iter("DataItem") if e.get("Format") == "Binary"]
        assert all(e.text.strip() == g.get("Name") for e in items)
    assert O.xdmf_layout(text) == O.xdmf_layout(ref)
    assert len(O.xdmf_layout(text)) == 15 * (last - 9)  # 3 coordinates, 3 + 1 + 1 + 1 + 3 + 3 fields
    # t_plot = 0: velocity, level set and obstacles only
    pysuli.write_xdmf(str(path), s.p.nx, s.p.ny, s.p.nz, 10, 10, 0, 1.0)
    assert len(O.xdmf_layout(open(path).read())) == 8
