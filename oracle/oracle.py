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
    mms_t: int = 0     # manufactured solution: 0 none, 1 steady, 2 unsteady (0_modules.f90:37)
    mms_a: float = 0.0     # its surface amplitude a, mean depth h0 and pressure scale coef
    mms_h0: float = 2.0    # (module mms_m, 0_modules.f90:208-212)
    mms_coef: float = 0.0
    wave_t: int = 0    # wave inflow at x = 0: 0 none, 1 Stokes I, 2 Stokes II, 5 Stokes V (0_modules.f90:35)
    wave_a: float = 0.03   # its amplitude, period and still depth (waves_coef, 3_contorno.f90:1086-1088)
    wave_p: float = 2.5
    wave_h0: float = 1.0

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
        ("mms_t", C.c_int), ("bp", C.POINTER(C.c_double) * 24),
        ("wave_t", C.c_int), ("wave_a", C.c_double), ("wave_p", C.c_double), ("wave_h0", C.c_double),
        ("wc", C.c_double * 14),
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
        L.or_stage_noproj.argtypes = [P, C.c_int, C.c_double]
        L.or_stage_noproj.restype = None
        for fn in ["or_waves_coef", "or_boundary_waves"]:
            getattr(L, fn).argtypes = [P]
            getattr(L, fn).restype = None
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
        self.planes = None  # optional boundary planes of wall code 3 (name -> array, PLANE_NAMES)
        self.wc = np.zeros(14)  # wave coefficients of waves_coef (l_w, n_w, f_w, c_w, aeta1..5, avel1..5)
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
        s.planes = None if self.planes is None else {k: a.copy() for k, a in self.planes.items()}
        s.wc = self.wc.copy()
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
        st.mms_t = p.mms_t
        st.wave_t, st.wave_a, st.wave_p, st.wave_h0 = p.wave_t, p.wave_a, p.wave_p, p.wave_h0
        for q in range(14):
            st.wc[q] = self.wc[q]
        if self.planes is None and p.wave_t:
            self.planes = {n: np.zeros(plane_shape(p, n)) for n in PLANE_NAMES}
        if self.planes is not None:
            self._bp_keep = [np.ascontiguousarray(self.planes[n], dtype=np.float64).ravel(order="F")
                             for n in PLANE_NAMES]
            for q, a in enumerate(self._bp_keep):
                assert a.size == int(np.prod(plane_shape(p, PLANE_NAMES[q])))
                st.bp[q] = _dptr(a)
        elif any(c == 3 for c in (p.ccx0, p.ccxf, p.ccy0, p.ccyf, p.ccz0, p.cczf)) or p.ccx0 == 4:
            raise ValueError("wall code 3 (and 4 at x = 0) needs the boundary planes (State.planes)")
        return st

    def _pull(self, st: _OrState) -> None:
        for k in ["dt", "t", "vol_ini", "vol_ins"]:
            self.scal[k] = float(getattr(st, k))
        self.wc = np.array([st.wc[q] for q in range(14)])
        if self.planes is not None:  # the C routines may write the planes (boundary_waves)
            for n, flat in zip(PLANE_NAMES, self._bp_keep):
                self.planes[n] = flat.reshape(plane_shape(self.p, n), order="F")
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
        elif cmd == "waves_coef":
            L.or_waves_coef(C.byref(st))
        elif cmd == "boundary_waves":
            L.or_boundary_waves(C.byref(st))
        elif cmd == "stage_noproj":  # the stage of the manufactured-solution mode
            L.or_stage_noproj(C.byref(st), tt, self.scal["dt"])
            st.dt = self.scal["dt"]
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
            + (f"_ob{p.obst_t}" if p.obst_t else "")
            + (f"_mms{p.mms_t}_a{p.mms_a:g}_h{p.mms_h0:g}_c{p.mms_coef:g}" if p.mms_t else "")
            + (f"_wave{p.wave_t}" if p.wave_t else ""))


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
    env = dict(os.environ, REF_VARIANT=variant)
    if p.mms_t:
        env["REF_MMS"] = f"{p.mms_t} {float(p.mms_a)!r} {float(p.mms_h0)!r} {float(p.mms_coef)!r}"
    if p.wave_t:
        env["REF_WAVE"] = str(p.wave_t)
    subprocess.run(args, check=True, stdout=subprocess.DEVNULL if quiet else None, env=env)
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


# boundary planes of the prescribed-velocity walls (wall code 3), in the order
# of their declaration (0_modules.f90:96-120), with their extents as
# (first-index length, second-index length); each starts at index 0
PLANE_NAMES = ("bxx0", "bxx1", "bxy0", "bxz0", "bxxf", "bxxf1", "bxyf", "bxzf",
               "byx0", "byy0", "byy1", "byz0", "byxf", "byyf", "byyf1", "byzf",
               "bzx0", "bzy0", "bzz0", "bzz1", "bzxf", "bzyf", "bzzf", "bzzf1")


def plane_shape(p: Params, name: str) -> tuple:
    nx, ny, nz = p.nx, p.ny, p.nz
    wall, comp = name[1], name[2]
    # the two in-plane axes of the wall, then the extra face of the component
    ax = {"x": ("y", "z"), "y": ("x", "z"), "z": ("x", "y")}[wall]
    n = {"x": nx, "y": ny, "z": nz}
    return tuple(n[a] + (3 if a == comp else 2) for a in ax)


def ref_run_mms(state: State, cmd: str, arg: int | None = None) -> tuple:
    """Run an mms command of the reference harness (mms_src, mms_run) on a copy
    of `state` (in place); returns its stdout and the arrays of OUT.mms
    (tf_u, tf_v, tf_w, tf_px, tf_py, tf_pz, the boundary planes, erro_u,
    erro_v, erro_w, erro_p), each in its module extents."""
    p = state.p
    exe = ref_build(p)
    wd = os.path.join(REF_DIR, ref_tag(p))
    os.makedirs(os.path.join(wd, "dados"), exist_ok=True)  # mms_i opens dados/erros_mms.txt
    fin = os.path.join(wd, f"in_{os.getpid()}.bin")
    fout = os.path.join(wd, f"out_{os.getpid()}.bin")
    state.to_file(fin)
    argv = [exe, fin, fout, cmd] + ([str(arg)] if arg is not None else [])
    res = subprocess.run(argv, check=True, capture_output=True, text=True, cwd=wd)
    state.from_file(fout)
    raw = np.fromfile(fout + ".mms", dtype="<f8")
    for f in (fin, fout, fout + ".mms"):
        os.remove(f)
    nx, ny, nz = p.nx, p.ny, p.nz
    shapes = [("tf_u", (nx + 1, ny, nz)), ("tf_v", (nx, ny + 1, nz)), ("tf_w", (nx, ny, nz + 1)),
              ("tf_px", (nx + 1, ny, nz)), ("tf_py", (nx, ny + 1, nz)), ("tf_pz", (nx, ny, nz + 1))]
    shapes += [(n, plane_shape(p, n)) for n in PLANE_NAMES]
    shapes += [("erro_u", (nx + 1, ny, nz)), ("erro_v", (nx, ny + 1, nz)), ("erro_w", (nx, ny, nz + 1)),
               ("erro_p", (nx, ny, nz))]
    out, q = {}, 0
    for n, sh in shapes:
        m = int(np.prod(sh))
        out[n] = raw[q:q + m].reshape(sh, order="F")
        q += m
    if q != raw.size:
        raise RuntimeError(f"mms output holds {raw.size} values, expected {q}")
    return res.stdout, out


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


# -- the manufactured solution of Weng (2009), 13_mms.f90 -------------------------
# The solution is a free-surface flow over a depth h(x, y[, t]): level set
# phi = h - z, velocities built from sin/cos products of x, y and z/h.
# mms_t = 1 (steady, 13_mms.f90:38-111):
#   h = a sin x sin y + h0
#   u = sin x cos y sin(pi z/(2h)) - cos x sin y (cos(2 pi z/h) - 1)
#   v = -cos x sin y sin(pi z/(2h)) + sin x cos y (cos(2 pi z/h) - 1)
#   w = -a (sin^2 x cos^2 y - cos^2 x sin^2 y) (sin(2 pi z/h) - (2 pi z/h) cos(2 pi z/h) + 2 pi) / (2 pi)
# mms_t = 2 (unsteady, :114-207), with X = x - pi, Y = y - pi:
#   h = a cos(X/2) cos(Y/2) cos t + h0,  c = cos(pi/2 (1 - z/h))
#   u = sin Y cos^2(X/2) c sin t,  v = -sin X cos^2(Y/2) c sin t,  w = -a cos(X/2) cos(Y/2) c sin t
#   p = -coef cos X cos Y cos(pi z/(2h))
def _weng_xy(p: Params, x, y):
    return (x - np.pi, y - np.pi) if p.mms_t == 2 else (x, y)


def weng_h(p: Params, x, y, t: float):
    X, Y = _weng_xy(p, x, y)
    a, h0 = p.mms_a, p.mms_h0
    if p.mms_t == 1:
        return a * np.sin(X) * np.sin(Y) + h0
    return a * np.cos(X * 0.5) * np.cos(Y * 0.5) * np.cos(t) + h0


def weng_vel(p: Params, comp: str, x, y, z, t: float):
    """Velocity component comp ('u', 'v', 'w') of the manufactured solution at
    the points (x, y, z), time t."""
    X, Y = _weng_xy(p, x, y)
    h = weng_h(p, x, y, t)
    a = p.mms_a
    if p.mms_t == 1:
        s4, c2 = np.sin(np.pi * 0.5 * z / h), np.cos(np.pi * 2.0 * z / h) - 1.0
        if comp == "u":
            return np.sin(X) * np.cos(Y) * s4 - np.cos(X) * np.sin(Y) * c2
        if comp == "v":
            return -np.cos(X) * np.sin(Y) * s4 + np.sin(X) * np.cos(Y) * c2
        r = 2.0 * np.pi * z / h
        return (-a * (np.sin(X) ** 2 * np.cos(Y) ** 2 - np.cos(X) ** 2 * np.sin(Y) ** 2)
                * (np.sin(r) - r * np.cos(r) + 2.0 * np.pi) / (2.0 * np.pi))
    c = np.cos(0.5 * np.pi * (1.0 - z / h)) * np.sin(t)
    if comp == "u":
        return np.sin(Y) * np.cos(X * 0.5) * np.cos(X * 0.5) * c
    if comp == "v":
        return -np.sin(X) * np.cos(Y * 0.5) * np.cos(Y * 0.5) * c
    return -a * np.cos(X * 0.5) * np.cos(Y * 0.5) * c


def weng_p(p: Params, x, y, z, t: float):
    X, Y = _weng_xy(p, x, y)
    return -p.mms_coef * np.cos(X) * np.cos(Y) * np.cos(0.5 * np.pi * z / weng_h(p, x, y, t))


def _grid(p: Params, off, rng):
    """Coordinates ((i - ox) dx, ...) of the index ranges rng = ((i0, i1), ...)."""
    ax = [(np.arange(r0, r1 + 1) - o) * h for (r0, r1), o, h in zip(rng, off, (p.dx, p.dy, p.dz))]
    return np.meshgrid(*ax, indexing="ij")


def weng_init(st: State, t: float = 0.0) -> None:
    """mms_i (13_mms.f90:12-212): the level set and the velocities at the
    interior faces (u(1:nx+1,1:ny,1:nz), ...), for mms_t = 2 also the pressure
    of the interior cells; ghosts are left to contorno."""
    p = st.p
    nx, ny, nz = p.nx, p.ny, p.nz
    X, Y, Z = _grid(p, (0.5, 0.5, 0.5), ((1, nx), (1, ny), (1, nz)))
    st.ls[...] = weng_h(p, X, Y, t) - Z
    st.u[1:nx + 2, 1:ny + 1, 1:nz + 1] = weng_vel(p, "u", *_grid(p, (1.0, 0.5, 0.5), ((1, nx + 1), (1, ny), (1, nz))), t)
    st.v[1:nx + 1, 1:ny + 2, 1:nz + 1] = weng_vel(p, "v", *_grid(p, (0.5, 1.0, 0.5), ((1, nx), (1, ny + 1), (1, nz))), t)
    st.w[1:nx + 1, 1:ny + 1, 1:nz + 2] = weng_vel(p, "w", *_grid(p, (0.5, 0.5, 1.0), ((1, nx), (1, ny), (1, nz + 1))), t)
    if p.mms_t == 2:
        st.prd1[1:nx + 1, 1:ny + 1, 1:nz + 1] = weng_p(p, X, Y, Z, t)


# mms_bc (13_mms.f90:468-816): per plane, the component whose formula it takes
# and the index of the plane along the wall normal (the reference evaluates
# the formula at the plane's own staggered position).  In the unsteady case
# bzy0 and bzyf take the formula of w at the positions of v (:777-781, :795-799).
_WENG_PLANES = {
    # name: (wall axis, component, normal index, extent lower bounds)
    "bxx0": ("x", "u", 0), "bxx1": ("x", "u", 1), "bxxf": ("x", "u", "n1"), "bxxf1": ("x", "u", "n1+1"),
    "bxy0": ("x", "v", 0), "bxz0": ("x", "w", 0), "bxyf": ("x", "v", "n+1"), "bxzf": ("x", "w", "n+1"),
    "byy0": ("y", "v", 0), "byy1": ("y", "v", 1), "byyf": ("y", "v", "n1"), "byyf1": ("y", "v", "n1+1"),
    "byx0": ("y", "u", 0), "byz0": ("y", "w", 0), "byxf": ("y", "u", "n+1"), "byzf": ("y", "w", "n+1"),
    "bzz0": ("z", "w", 0), "bzz1": ("z", "w", 1), "bzzf": ("z", "w", "n1"), "bzzf1": ("z", "w", "n1+1"),
    "bzx0": ("z", "u", 0), "bzy0": ("z", "v", 0), "bzxf": ("z", "u", "n+1"), "bzyf": ("z", "v", "n+1"),
}


def weng_planes(p: Params, t: float, planes: dict | None = None) -> dict:
    """The 24 boundary planes of mms_bc at time t, each in its module extents
    (entries the reference's loops leave out stay as they were in `planes`,
    zero if new).  For mms_t = 1 the planes are set only at t = 0."""
    out = {n: np.zeros(plane_shape(p, n)) for n in PLANE_NAMES} if planes is None else planes
    if p.mms_t == 1 and t != 0.0:
        return out
    n = {"x": p.nx, "y": p.ny, "z": p.nz}
    for name, (wall, comp, norm) in _WENG_PLANES.items():
        nn = n[wall]
        q = {0: 0, 1: 1, "n1": nn + 1, "n1+1": nn + 2, "n+1": nn + 1}[norm]
        # position offsets of the component's staggered grid: its own axis at
        # the face (index - 1), the others at cell centres (index - 0.5)
        face_axis = {"u": "x", "v": "y", "w": "z"}[comp]
        off = {ax: (1.0 if ax == face_axis else 0.5) for ax in "xyz"}
        coord = {}
        ax_in = [a for a in "xyz" if a != wall]
        # the loops run over 0..n+1 of both in-plane axes (13_mms.f90:485-486, ...)
        g = np.meshgrid(*[np.arange(0, n[a] + 2) for a in ax_in], indexing="ij")
        for a, gi in zip(ax_in, g):
            coord[a] = (gi - off[a]) * {"x": p.dx, "y": p.dy, "z": p.dz}[a]
        coord[wall] = (q - off[wall]) * {"x": p.dx, "y": p.dy, "z": p.dz}[wall] + 0.0 * g[0]
        f = "w" if (p.mms_t == 2 and name in ("bzy0", "bzyf")) else comp
        val = weng_vel(p, f, coord["x"], coord["y"], coord["z"], t)
        out[name][:val.shape[0], :val.shape[1]] = val
    return out


_WENG_SRC = {}


def weng_sources(p: Params, t: float) -> tuple:
    """Exact momentum sources of the manufactured solution at time t on the
    face grids of u, v, w (tf_u(nx+1,ny,nz), ...): [dq/dt +] u.grad q -
    nu (lap q + d/dq_axis (div u)-like terms) with nu = 0.00100798/998, the
    viscous form of convdiff (6_convdiff.f90:150-186):
        tf_u = [u_t] + u u_x + v u_y + w u_z - nu (2 u_xx + u_yy + u_zz + v_xy + w_xz)
    and likewise for v, w (termo_fonte1, termo_fonte2: 13_mms.f90:819-1322).
    Derived symbolically (sympy) from the fields of weng_vel; the reference's
    hand-derived expressions agree with these for a = 0 (its setting,
    0_modules.f90:208) and deviate for a != 0."""
    key = (p.mms_t, float(p.mms_a), float(p.mms_h0))
    if key not in _WENG_SRC:
        import sympy as sp
        x, y, z, t_ = sp.symbols("x y z t")
        a, h0 = sp.Float(p.mms_a), sp.Float(p.mms_h0)
        if p.mms_t == 1:
            X, Y = x, y
            h = a * sp.sin(X) * sp.sin(Y) + h0
            u = sp.sin(X) * sp.cos(Y) * sp.sin(sp.pi / 2 * z / h) - sp.cos(X) * sp.sin(Y) * (sp.cos(2 * sp.pi * z / h) - 1)
            v = -sp.cos(X) * sp.sin(Y) * sp.sin(sp.pi / 2 * z / h) + sp.sin(X) * sp.cos(Y) * (sp.cos(2 * sp.pi * z / h) - 1)
            r = 2 * sp.pi * z / h
            w = -a * (sp.sin(X) ** 2 * sp.cos(Y) ** 2 - sp.cos(X) ** 2 * sp.sin(Y) ** 2) * (
                sp.sin(r) - r * sp.cos(r) + 2 * sp.pi) / (2 * sp.pi)
        else:
            X, Y = x - sp.pi, y - sp.pi
            h = a * sp.cos(X / 2) * sp.cos(Y / 2) * sp.cos(t_) + h0
            c = sp.cos(sp.pi / 2 * (1 - z / h)) * sp.sin(t_)
            u = sp.sin(Y) * sp.cos(X / 2) ** 2 * c
            v = -sp.sin(X) * sp.cos(Y / 2) ** 2 * c
            w = -a * sp.cos(X / 2) * sp.cos(Y / 2) * c
        nu = sp.Float(0.00100798 / 998.0)
        V = (u, v, w)
        xs = (x, y, z)
        fns = []
        for d, q in enumerate(V):
            adv = sum(V[e] * sp.diff(q, xs[e]) for e in range(3))
            visc = sum(sp.diff(q, xs[e], 2) for e in range(3)) + sum(sp.diff(V[e], xs[d], xs[e]) for e in range(3))
            src = adv - nu * visc + (sp.diff(q, t_) if p.mms_t == 2 else 0)
            fns.append(sp.lambdify((x, y, z, t_), src, "numpy"))
        _WENG_SRC[key] = fns
    fu, fv, fw = _WENG_SRC[key]
    nx, ny, nz = p.nx, p.ny, p.nz
    one = lambda f, g: f(*g, t) * np.ones_like(g[0])  # noqa: E731
    return (one(fu, _grid(p, (1.0, 0.5, 0.5), ((1, nx + 1), (1, ny), (1, nz)))),
            one(fv, _grid(p, (0.5, 1.0, 0.5), ((1, nx), (1, ny + 1), (1, nz)))),
            one(fw, _grid(p, (0.5, 0.5, 1.0), ((1, nx), (1, ny), (1, nz + 1)))))


def _interp_cf(a, axis):
    """Centre to face average with the end faces copied (8_derivadas.f90:304-434)."""
    m = np.moveaxis(a, axis, 0)
    mid = (m[1:] + m[:-1]) * 0.5
    return np.moveaxis(np.concatenate([mid[:1], mid, mid[-1:]], axis=0), 0, axis)


def weng_errors(st: State, t: float, erro_p: np.ndarray) -> tuple:
    """mms() (13_mms.f90:218-465): RMS errors of u, v, w over the faces whose
    interpolated level set is positive and of the pressure over the positive
    cells.  erro_p (nx,ny,nz) is updated in place as the reference's module
    array: only the wet cells are rewritten, so a cell that dries keeps its
    last value and still enters the sum.  For mms_t = 1 the reference's exact
    pressure is unset; the pressure error is returned as NaN."""
    p = st.p
    nx, ny, nz = p.nx, p.ny, p.nz
    ls = st.ls
    res = []
    for comp, ax, arr, off, rng in (
            ("u", 0, st.u[1:nx + 2, 1:ny + 1, 1:nz + 1], (1.0, 0.5, 0.5), ((1, nx + 1), (1, ny), (1, nz))),
            ("v", 1, st.v[1:nx + 1, 1:ny + 2, 1:nz + 1], (0.5, 1.0, 0.5), ((1, nx), (1, ny + 1), (1, nz))),
            ("w", 2, st.w[1:nx + 1, 1:ny + 1, 1:nz + 2], (0.5, 0.5, 1.0), ((1, nx), (1, ny), (1, nz + 1)))):
        wet = _interp_cf(ls, ax) > 0.0
        e = np.where(wet, (weng_vel(p, comp, *_grid(p, off, rng), t) - arr) ** 2, 0.0)
        res.append(np.sqrt(e.sum() / wet.sum()))
    wet = ls > 0.0
    if p.mms_t == 2:
        pm = weng_p(p, *_grid(p, (0.5, 0.5, 0.5), ((1, nx), (1, ny), (1, nz))), t)
        erro_p[wet] = ((pm - st.prd1[1:nx + 1, 1:ny + 1, 1:nz + 1]) ** 2)[wet]
        res.append(np.sqrt(erro_p.sum() / wet.sum()))
    else:
        res.append(float("nan"))
    return tuple(float(r) for r in res)


def weng_run(p: Params, dt: float, nsteps: int) -> tuple:
    """The manufactured-solution run of 1_principal.f90:39-108 with mms_t > 0
    from t = 0: mms_i and level_set_ini's tail (14_level_set.f90:178-218),
    contorno(1), contorno(3), then per step t = it dt, the sources (mms_t = 1:
    once, at it = 1; mms_t = 2: every step), level_set, contorno(3) and the
    stages without the projection (visco, convdiff, tempo, contorno(2));
    mms() after each step.  Returns the final state and the errors (u, v, w,
    p) at t = 0 and after every step."""
    st = State(p)
    st.scal["dt"] = dt
    st.scal["t"] = 0.0
    weng_init(st, 0.0)
    st.planes = weng_planes(p, 0.0)
    st.run("ls_ini")
    st.run("contorno", 1)
    st.run("contorno", 3)
    erro_p = np.zeros((p.nx, p.ny, p.nz))
    errs = [weng_errors(st, 0.0, erro_p)]
    for it in range(1, nsteps + 1):
        t = it * dt
        st.scal["t"] = t
        if (p.mms_t == 1 and it == 1) or p.mms_t == 2:
            st.tf = weng_sources(p, t)
        st.run("level_set")
        st.run("contorno", 3)
        for tt in range(1, or_ntt(p.t_tempo) + 1):
            st.planes = weng_planes(p, t, st.planes)
            st.run("stage_noproj", tt=tt, it=it)
        errs.append(weng_errors(st, t, erro_p))
    return st, errs


def or_ntt(t_tempo: int) -> int:
    return 3 if t_tempo == 2 else 2 if t_tempo == 1 else 1
