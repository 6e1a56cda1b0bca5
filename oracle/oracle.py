"""Python side of the CPU oracle (test infrastructure only).

Holds a SuLi-IC solver state in the reference's own array extents and
Fortran order (0_modules.f90:87-135,177-201,71-73), runs the plain-C FP64
restatement (liboracle.so, suli_oracle.c) on it, and reads/writes the raw
state files of the flang-built reference harness (oracle/ref/harness.f90).

Only tests/, the smoke check and the CPU-baseline leg of the benchmark
import this module; the product never does.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

# Arrays of the state buffer, in the order of suli_oracle.c:or_bind and of the
# reference harness state file.  Extent kinds: U/V/W/P ghosted velocity and
# pressure arrays, C cell arrays, FX/FY/FZ face arrays without ghosts.
ARRAYS = [
    ("u", "U"), ("v", "V"), ("w", "W"), ("prd1", "P"),
    ("rho", "C"), ("ls_nu", "C"), ("ls", "C"), ("mod_ls", "C"), ("kurv", "C"),
    ("hs", "C"), ("hsx", "C"), ("hsy", "C"), ("hsz", "C"),
    ("ddlsdx", "C"), ("ddlsdy", "C"), ("ddlsdz", "C"),
    ("nut", "C"), ("xnut", "FX"), ("ynut", "FY"), ("znut", "FZ"),
    ("Fu", "FX"), ("Fv", "FY"), ("Fw", "FZ"),
    ("u0", "FX"), ("v0", "FY"), ("w0", "FZ"),
    ("ub", "U"), ("vb", "V"), ("wb", "W"),
    ("fu0", "FX"), ("fv0", "FY"), ("fw0", "FZ"), ("fu1", "FX"), ("fv1", "FY"), ("fw1", "FZ"),
    ("ku", "KU"), ("kv", "KV"), ("kw", "KW"),
]
SCALARS = ["dt", "t", "vol_ini", "vol_ins", "rho_f1", "rho_f2", "mi_f1", "mi_f2", "sigma", "alpha1"]


def extent_shape(kind: str, nx: int, ny: int, nz: int) -> tuple:
    """Fortran shape of an array kind (0_modules.f90:92-133)."""
    return {
        "U": (nx + 3, ny + 2, nz + 2),
        "V": (nx + 2, ny + 3, nz + 2),
        "W": (nx + 2, ny + 2, nz + 3),
        "P": (nx + 2, ny + 2, nz + 2),
        "C": (nx, ny, nz),
        "FX": (nx + 1, ny, nz),
        "FY": (nx, ny + 1, nz),
        "FZ": (nx, ny, nz + 1),
        "KU": (nx + 3, ny + 2), "KV": (nx + 2, ny + 3), "KW": (nx + 2, ny + 2),  # obstacle columns
    }[kind]


@dataclasses.dataclass
class Params:
    """Grid, boundary codes, model switches and fluid properties."""
    nx: int = 8
    ny: int = 6
    nz: int = 5
    dx: float = 0.125
    dy: float = 0.125
    dz: float = 0.125
    ccx0: int = 2
    ccxf: int = 2
    ccy0: int = 2
    ccyf: int = 2
    ccz0: int = 2
    cczf: int = 2
    m_turb: int = 1
    csmag: float = 0.1
    gx: float = 0.0
    gz: float = 9.81
    chezy: float = 44.5
    rho_f1: float = 1.2
    rho_f2: float = 1000.0
    mi_f1: float = 1.8e-5
    mi_f2: float = 1.0e-3
    sigma: float = 0.0728
    alpha1: float = 1.5
    t_tempo: int = 1  # 0 Euler, 1 RK2, 2 RK3, 3 AB2 (0_modules.f90:24)
    der: int = 1       # advection differences: 1 upwind, 2 central, 3 upwind 2nd order (0_modules.f90:26)
    adv_type: int = 1  # advection form: 1 classic, 2 rotational, 3 antisymmetric (0_modules.f90:27)
    obst_t: int = 0    # nonzero: the state's ku, kv, kw hold a bottom obstacle (0_modules.f90:29)

    @property
    def dx1(self) -> float:
        return max(self.dx, self.dy, self.dz)

    @property
    def dt1(self) -> float:
        return 0.1 * self.dx1  # 0_modules.f90:198


class _OrState(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
        ("ccx0", C.c_int), ("ccxf", C.c_int), ("ccy0", C.c_int), ("ccyf", C.c_int),
        ("ccz0", C.c_int), ("cczf", C.c_int),
        ("m_turb", C.c_int),
        ("csmag", C.c_double), ("gx", C.c_double), ("gz", C.c_double), ("chezy", C.c_double),
        ("rho_f1", C.c_double), ("rho_f2", C.c_double), ("mi_f1", C.c_double), ("mi_f2", C.c_double),
        ("sigma", C.c_double), ("alpha1", C.c_double),
        ("dt", C.c_double), ("t", C.c_double), ("vol_ini", C.c_double), ("vol_ins", C.c_double),
        ("dt1", C.c_double), ("a_dt", C.c_double * 3), ("tt", C.c_int), ("t_tempo", C.c_int), ("it", C.c_int),
        ("der", C.c_int), ("adv_type", C.c_int), ("obst_t", C.c_int),
        ("cg_iters", C.c_int), ("cg_res2", C.c_double), ("reinit_iters", C.c_int),
        ("mass_corr_iters", C.c_int), ("cg_max_it", C.c_int),
    ] + [(name, C.POINTER(C.c_double)) for name, _ in ARRAYS] + [
        ("tf_u", C.POINTER(C.c_double)), ("tf_v", C.POINTER(C.c_double)), ("tf_w", C.POINTER(C.c_double)),
    ]


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            subprocess.run(["make", "-s", "-C", HERE], check=True)
        L = C.CDLL(path)
        P = C.POINTER(_OrState)
        for fn in ["or_heaviside", "or_mod_ls1", "or_reinic_weno", "or_level_set", "or_level_set_ini_tail",
                   "or_visco", "or_convdiff", "or_tempo", "or_graddin", "or_posdin"]:
            getattr(L, fn).argtypes = [P]
            getattr(L, fn).restype = None
        L.or_contorno.argtypes = [P, C.c_int]
        L.or_contorno.restype = None
        L.or_est_div.argtypes = [P]
        L.or_est_div.restype = C.c_double
        L.or_stage.argtypes = [P, C.c_int, C.c_double]
        L.or_stage.restype = None
        L.or_step.argtypes = [P, C.c_double]
        L.or_step.restype = None
        L.or_state_len.argtypes = [C.c_int, C.c_int, C.c_int]
        L.or_state_len.restype = C.c_int64
        L.or_bind.argtypes = [P, C.POINTER(C.c_double)]
        L.or_bind.restype = None
        dp = C.POINTER(C.c_double)
        L.or_weno1.argtypes = [dp, dp, C.c_int, C.c_double, dp, C.c_int64, C.c_int]
        L.or_weno1.restype = None
        L.or_der_weno.argtypes = [P, dp, dp, dp, dp, dp, dp, dp]
        L.or_der_weno.restype = C.c_int
        _LIB = L
    return _LIB


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class State:
    """A full solver state in reference layout, runnable on the C oracle."""

    def __init__(self, p: Params):
        self.p = p
        n = 0
        self.offsets = {}
        for name, kind in ARRAYS:
            shp = extent_shape(kind, p.nx, p.ny, p.nz)
            self.offsets[name] = (n, shp)
            n += int(np.prod(shp))
        self.buf = np.zeros(n, dtype=np.float64)
        self.scal = {k: 0.0 for k in SCALARS}
        self.scal.update(rho_f1=p.rho_f1, rho_f2=p.rho_f2, mi_f1=p.mi_f1, mi_f2=p.mi_f2,
                         sigma=p.sigma, alpha1=p.alpha1)
        self.tf = None  # optional (tf_u, tf_v, tf_w) source arrays
        self.diag = {}
        self.cg_max_it = 0  # graddin iteration cap (0: the reference's 10000)

    def __getattr__(self, name):
        offs = self.__dict__.get("offsets")
        if offs is not None and name in offs:
            o, shp = offs[name]
            return self.buf[o:o + int(np.prod(shp))].reshape(shp, order="F")
        raise AttributeError(name)

    def arr(self, name) -> np.ndarray:
        return getattr(self, name)

    def copy(self) -> "State":
        s = State(self.p)
        s.buf[:] = self.buf
        s.scal = dict(self.scal)
        s.tf = None if self.tf is None else tuple(a.copy() for a in self.tf)
        s.cg_max_it = self.cg_max_it
        return s

    # -- reference harness state files --------------------------------------
    def to_file(self, path: str) -> None:
        with open(path, "wb") as f:
            np.array([self.p.nx, self.p.ny, self.p.nz], dtype="<i4").tofile(f)
            np.array([self.scal[k] for k in SCALARS], dtype="<f8").tofile(f)
            self.buf.astype("<f8").tofile(f)

    def from_file(self, path: str) -> None:
        with open(path, "rb") as f:
            hdr = np.fromfile(f, dtype="<i4", count=3)
            assert tuple(hdr) == (self.p.nx, self.p.ny, self.p.nz), hdr
            sc = np.fromfile(f, dtype="<f8", count=len(SCALARS))
            self.scal = dict(zip(SCALARS, [float(x) for x in sc]))
            self.buf[:] = np.fromfile(f, dtype="<f8", count=self.buf.size)

    # -- the C oracle ---------------------------------------------------------
    def _struct(self) -> _OrState:
        p = self.p
        st = _OrState()
        for f in ["nx", "ny", "nz", "dx", "dy", "dz", "ccx0", "ccxf", "ccy0", "ccyf", "ccz0", "cczf",
                  "m_turb", "csmag", "gx", "gz", "chezy", "t_tempo", "der", "adv_type", "obst_t"]:
            setattr(st, f, getattr(p, f))
        for k in ["rho_f1", "rho_f2", "mi_f1", "mi_f2", "sigma", "alpha1", "dt", "t", "vol_ini", "vol_ins"]:
            setattr(st, k, self.scal[k])
        st.dt1 = p.dt1
        st.cg_max_it = self.cg_max_it
        lib().or_bind(C.byref(st), _dptr(self.buf))
        if self.tf is not None:
            self._tf_keep = [np.ascontiguousarray(a, dtype=np.float64).ravel(order="F") for a in self.tf]
            st.tf_u, st.tf_v, st.tf_w = [_dptr(a) for a in self._tf_keep]
        return st

    def _pull(self, st: _OrState) -> None:
        for k in ["dt", "t", "vol_ini", "vol_ins"]:
            self.scal[k] = float(getattr(st, k))
        self.diag = dict(cg_iters=st.cg_iters, cg_res2=st.cg_res2, reinit_iters=st.reinit_iters,
                         mass_corr_iters=st.mass_corr_iters)

    def run(self, cmd: str, arg: int = 0, tt: int = 1, it: int = 1) -> "State":
        """Run one oracle routine (same command names as the reference harness);
        `it` is the step counter that tempo sees (AB2 takes Euler at it = 1)."""
        st = self._struct()
        st.tt = tt
        st.it = it
        L = lib()
        if cmd == "ls_ini":
            L.or_level_set_ini_tail(C.byref(st))
        elif cmd == "level_set":
            L.or_level_set(C.byref(st))
        elif cmd == "heaviside":
            L.or_heaviside(C.byref(st))
        elif cmd == "mod_ls1":
            L.or_mod_ls1(C.byref(st))
        elif cmd == "reinic":
            L.or_reinic_weno(C.byref(st))
        elif cmd == "visco":
            L.or_visco(C.byref(st))
        elif cmd == "convdiff":
            L.or_convdiff(C.byref(st))
        elif cmd == "tempo":
            st.tt = arg
            L.or_tempo(C.byref(st))
        elif cmd == "contorno":
            L.or_contorno(C.byref(st), arg)
        elif cmd == "graddin":
            L.or_graddin(C.byref(st))
        elif cmd == "posdin":
            L.or_posdin(C.byref(st))
        elif cmd == "stage":  # like the harness, leave dt at its base value afterwards
            L.or_stage(C.byref(st), arg, self.scal["dt"])
            st.dt = self.scal["dt"]
        elif cmd == "step":
            for it in range(1, max(arg, 1) + 1):
                st.it = it
                st.t = it * st.dt  # 1_principal.f90:65
                L.or_step(C.byref(st), st.dt)
        else:
            raise ValueError(cmd)
        self._pull(st)
        return self

    def est_div(self) -> float:
        st = self._struct()
        return float(lib().or_est_div(C.byref(st)))


def weno1(phi: np.ndarray, h: float, ihs: int):
    """WENO5 one-sided derivatives of one pencil (14_level_set.f90:511-610)."""
    phi = np.ascontiguousarray(phi, dtype=np.float64)
    n = phi.size
    dp = np.zeros(n)
    dn = np.zeros(n)
    lib().or_weno1(_dptr(dp), _dptr(dn), n, h, _dptr(phi), 1, ihs)
    return dp, dn


# -- the flang-built reference --------------------------------------------------
REF_DIR = os.path.join(HERE, "_ref")


def ref_tag(p: Params) -> str:
    return (f"g{p.nx}x{p.ny}x{p.nz}_h{p.dx:g}_{p.dy:g}_{p.dz:g}_bc{p.ccx0}{p.ccxf}{p.ccy0}{p.ccyf}{p.ccz0}{p.cczf}"
            f"_t{p.m_turb}_cs{p.csmag:g}_g{p.gx:g}_{p.gz:g}" + (f"_tt{p.t_tempo}" if p.t_tempo != 1 else "")
            + (f"_adv{p.adv_type}{p.der}" if (p.der, p.adv_type) != (1, 1) else "")
            + (f"_ob{p.obst_t}" if p.obst_t else ""))


def ref_build(p: Params, quiet: bool = True, variant: str = "") -> str:
    """Build (if needed) the reference harness for this configuration;
    variant "bench" / "bench_rel" are the CPU-baseline builds of
    build_ref.sh (REF_VARIANT)."""
    tag = ref_tag(p) + (f"_{variant}" if variant else "")
    exe = os.path.join(REF_DIR, tag, "harness")
    srcs = [os.path.join(HERE, "ref", f) for f in ("harness.f90", "build_ref.sh")]
    if os.path.exists(exe) and os.path.getmtime(exe) >= max(os.path.getmtime(f) for f in srcs):
        return exe
    args = [os.path.join(HERE, "ref", "build_ref.sh"), tag, str(p.nx), str(p.ny), str(p.nz),
            repr(float(p.dx)), repr(float(p.dy)), repr(float(p.dz)),
            str(p.ccx0), str(p.ccxf), str(p.ccy0), str(p.ccyf), str(p.ccz0), str(p.cczf),
            str(p.m_turb), repr(float(p.csmag)), repr(float(p.gx)), repr(float(p.gz)), str(p.t_tempo),
            str(p.der), str(p.adv_type), str(p.obst_t)]
    subprocess.run(args, check=True, stdout=subprocess.DEVNULL if quiet else None,
                   env=dict(os.environ, REF_VARIANT=variant))
    return exe


def ref_available(p: Params) -> bool:
    return os.path.exists(os.path.join(REF_DIR, ref_tag(p), "harness"))


def ref_run(state: State, cmd: str, arg: int | None = None, workdir: str | None = None) -> str:
    """Run the flang-built reference harness on a copy of `state` (in place)."""
    exe = ref_build(state.p)
    wd = workdir or os.path.join(REF_DIR, ref_tag(state.p))
    fin = os.path.join(wd, f"in_{os.getpid()}.bin")
    fout = os.path.join(wd, f"out_{os.getpid()}.bin")
    state.to_file(fin)
    argv = [exe, fin, fout, cmd] + ([str(arg)] if arg is not None else [])
    res = subprocess.run(argv, check=True, capture_output=True, text=True, cwd=wd)
    state.from_file(fout)
    os.remove(fin)
    os.remove(fout)
    return res.stdout


# -- synthetic states -------------------------------------------------------------
def dambreak_phi(p: Params, width: float, height: float) -> np.ndarray:
    """Signed distance (positive in water) to a water column filling
    x <= width, z <= height across the whole y extent, at cell centres
    x_i = (i-0.5)dx (14_level_set.f90:53-63).  The column's faces on the
    walls x=0, y=0, y=Ly, z=0 are not interfaces, so the distance is the
    2-D distance to the free faces x=width and z=height."""
    x = (np.arange(1, p.nx + 1) - 0.5) * p.dx
    z = (np.arange(1, p.nz + 1) - 0.5) * p.dz
    X, Z = np.meshgrid(x, z, indexing="ij")
    ax = X - width
    az = Z - height
    inside = (ax <= 0) & (az <= 0)
    d_in = np.minimum(-ax, -az)
    d_out = -np.sqrt(np.maximum(ax, 0) ** 2 + np.maximum(az, 0) ** 2)
    phi2 = np.where(inside, d_in, d_out)
    return np.repeat(phi2[:, None, :], p.ny, axis=1)


def random_state(p: Params, seed: int = 0, vel: float = 0.05, width: float = 0.4, height: float = 0.3,
                 perturb: float = 0.0) -> State:
    """A dam-break level set with smooth random velocities and pressure, for
    routine-level parity tests.  The level-set derived fields are filled by the
    oracle's level_set_ini tail (heaviside, volume, mod_ls1)."""
    rng = np.random.default_rng(seed)
    s = State(p)
    phi = dambreak_phi(p, width, height)
    if perturb:
        x = (np.arange(1, p.nx + 1) - 0.5) * p.dx
        y = (np.arange(1, p.ny + 1) - 0.5) * p.dy
        X, Y = np.meshgrid(x, y, indexing="ij")
        phi = phi + perturb * p.dx * (np.sin(2 * np.pi * X / (p.nx * p.dx)) * np.cos(2 * np.pi * Y / (p.ny * p.dy)))[:, :, None]
    s.ls[...] = phi
    for name in ["u", "v", "w"]:
        a = s.arr(name)
        a[...] = vel * rng.standard_normal(a.shape)
    s.prd1[...] = 100.0 * rng.standard_normal(s.prd1.shape)
    s.scal["dt"] = 1e-3
    s.run("ls_ini")
    return s


# -- field output ------------------------------------------------------------------
def plot_record(s: State, t_plot: int = 1) -> np.ndarray:
    """The FP32 record that plot_f writes to arquivos/campos_NNNNN
    (4_plot.f90:260-363; 4_plot.f90:186-187 for t_plot = 0 and :211-213 for
    t_plot = 1, mms_t = 0, no obstacles), as a flat array: the FP64 values of
    the reference in its evaluation order, rounded once to FP32."""
    p = s.p
    nx, ny, nz = p.nx, p.ny, p.nz

    def axis(n, h):  # x1(0) = -dx, x1(i) = x1(i-1) + dx (4_plot.f90:264-266,275-277)
        t = np.empty(n + 2)
        x = -h
        t[0] = x
        for i in range(1, n + 2):
            x = x + h
            t[i] = x
        return t

    shp = (nx + 1, ny + 1, nz + 1)
    X = np.broadcast_to(axis(nx, p.dx)[1:, None, None], shp)
    Y = np.broadcast_to(axis(ny, p.dy)[None, 1:, None], shp)
    Z = np.broadcast_to(axis(nz, p.dz)[None, None, 1:], shp)
    u, v, w = s.u, s.v, s.w  # reference indices from 0
    a, b = slice(1, None), slice(0, -1)  # node index i and i-1 over i = 1..n+1
    U = u[1:nx + 2, 0:ny + 2, 0:nz + 2]
    V = v[0:nx + 2, 1:ny + 2, 0:nz + 2]
    W = w[0:nx + 2, 0:ny + 2, 1:nz + 2]
    # (4_plot.f90:269-272)
    uaux = (((U[:, a, a] + U[:, b, a]) + U[:, a, b]) + U[:, b, b]) * 0.25
    vaux = (((V[a, :, a] + V[b, :, a]) + V[a, :, b]) + V[b, :, b]) * 0.25
    waux = (((W[a, a, :] + W[a, b, :]) + W[b, a, :]) + W[b, b, :]) * 0.25
    parts = [X, Y, Z, uaux, vaux, waux, s.ls, np.zeros((nx, ny, nz))]
    if t_plot == 1:
        c = (slice(1, nx + 1), slice(1, ny + 1), slice(1, nz + 1))
        sh = lambda arr, di, dj, dk: arr[1 + di:nx + 1 + di, 1 + dj:ny + 1 + dj, 1 + dk:nz + 1 + dk]  # noqa: E731
        # (4_plot.f90:190-201)
        dudy = (sh(u, 0, 1, 0) - sh(u, 0, 0, 0)) / p.dy
        dudz = (sh(u, 0, 0, 1) - sh(u, 0, 0, 0)) / p.dz
        dvdx = (sh(v, 1, 0, 0) - sh(v, 0, 0, 0)) / p.dx
        dvdz = (sh(v, 0, 0, 1) - sh(v, 0, 0, 0)) / p.dz
        dwdx = (sh(w, 1, 0, 0) - sh(w, 0, 0, 0)) / p.dx
        dwdy = (sh(w, 0, 1, 0) - sh(w, 0, 0, 0)) / p.dy

        def interp_cf(arr, axis):  # 8_derivadas.f90:306-323 and its y, z versions
            m = np.moveaxis(arr, axis, 0)
            mid = (m[1:] + m[:-1]) * 0.5
            return np.moveaxis(np.concatenate([mid[:1], mid, mid[-1:]], axis=0), 0, axis)

        nut = s.nut
        parts += [s.prd1[c], dwdy - dvdz, dudz - dwdx, dvdx - dudy,
                  interp_cf(nut, 0), interp_cf(nut, 1), interp_cf(nut, 2)]
    return np.concatenate([np.asarray(x, dtype=np.float64).ravel(order="F").astype(np.float32) for x in parts])


def xdmf_layout(text: str) -> list:
    """(Seek, Dimensions) of every binary DataItem of an XDMF index, in order,
    from the reference's visu output (list-directed lines that its compiler
    may break inside a tag) or from a well-formed file."""
    import re
    flat = re.sub(r"\s+", " ", text.replace("\n ", ""))
    out = []
    for m in re.finditer(r"<DataItem[^>]*Format=\"Binary\"[^>]*>", flat):
        tag = m.group(0)
        seek = int(re.search(r"Seek=\"\s*(\d+)\s*\"", tag).group(1))
        dims = tuple(int(t) for t in re.search(r"Dimensions=\"([^\"]*)\"", tag).group(1).split())
        out.append((seek, dims))
    return out
