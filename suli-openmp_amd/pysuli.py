"""ctypes binding of libsuli.so (include/suli.h) for Python callers.

Mirrors the reference's calling convention: the solver state lives on the
GPU; whole arrays in the reference's extents and Fortran order are exchanged
with upload()/download(); the time step is driven by one call per reference
subroutine (1_principal.f90:74-92).  Every call raises SuliError when the
library reports a nonzero status, and loading fails loudly when libsuli.so
is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsuli.so")

PREC_FP64, PREC_MIXED = 0, 1
POISSON_REF, POISSON_FCG = 0, 1
PRECOND_JACOBI, PRECOND_GMG, PRECOND_LEARNED = 0, 1, 2

FIELDS = ["u", "v", "w", "prd1", "rho", "ls_nu", "ls", "mod_ls", "kurv", "hs", "hsx", "hsy", "hsz",
          "ddlsdx", "ddlsdy", "ddlsdz", "nut", "xnut", "ynut", "znut", "Fu", "Fv", "Fw", "u0", "v0", "w0",
          "tf_u", "tf_v", "tf_w", "fu0", "fv0", "fw0", "fu1", "fv1", "fw1"]
FIELD_ID = {n: q for q, n in enumerate(FIELDS)}
KIND = {"u": "U", "v": "V", "w": "W", "prd1": "P", "xnut": "FX", "ynut": "FY", "znut": "FZ",
        "Fu": "FX", "Fv": "FY", "Fw": "FZ", "u0": "FX", "v0": "FY", "w0": "FZ", "tf_u": "FX", "tf_v": "FY",
        "tf_w": "FZ", "fu0": "FX", "fv0": "FY", "fw0": "FZ", "fu1": "FX", "fv1": "FY", "fw1": "FZ"}
TIME_EULER, TIME_RK2, TIME_RK3, TIME_AB2 = 0, 1, 2, 3  # t_tempo, 0_modules.f90:24
# boundary planes of the prescribed-velocity walls (wall code 3), field ids
# 64.. in the declaration order of 0_modules.f90:96-120
PLANES = ["bxx0", "bxx1", "bxy0", "bxz0", "bxxf", "bxxf1", "bxyf", "bxzf",
          "byx0", "byy0", "byy1", "byz0", "byxf", "byyf", "byyf1", "byzf",
          "bzx0", "bzy0", "bzz0", "bzz1", "bzxf", "bzyf", "bzzf", "bzzf1"]
FIELD_ID.update({n: 64 + q for q, n in enumerate(PLANES)})


def plane_shape(name: str, nx: int, ny: int, nz: int) -> tuple:
    """Module extents of a boundary plane: the wall's two in-plane axes, with
    one more entry along the axis of the plane's velocity component."""
    n = {"x": nx, "y": ny, "z": nz}
    wall, comp = name[1], name[2]
    return tuple(n[a] + (3 if a == comp else 2) for a in "xyz" if a != wall)


def extent_shape(kind: str, nx: int, ny: int, nz: int) -> tuple:
    """Fortran shape of the reference arrays (0_modules.f90:92-133)."""
    return {
        "U": (nx + 3, ny + 2, nz + 2), "V": (nx + 2, ny + 3, nz + 2), "W": (nx + 2, ny + 2, nz + 3),
        "P": (nx + 2, ny + 2, nz + 2), "C": (nx, ny, nz), "FX": (nx + 1, ny, nz),
        "FY": (nx, ny + 1, nz), "FZ": (nx, ny, nz + 1),
    }[kind]


class SuliError(RuntimeError):
    pass


class Params(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
        ("ccx0", C.c_int32), ("ccxf", C.c_int32), ("ccy0", C.c_int32), ("ccyf", C.c_int32),
        ("ccz0", C.c_int32), ("cczf", C.c_int32),
        ("m_turb", C.c_int32),
        ("rho_f1", C.c_double), ("rho_f2", C.c_double), ("mi_f1", C.c_double), ("mi_f2", C.c_double),
        ("sigma", C.c_double), ("alpha1", C.c_double),
        ("csmag", C.c_double), ("gx", C.c_double), ("gz", C.c_double), ("chezy", C.c_double),
        ("precision", C.c_int32), ("poisson", C.c_int32), ("precond", C.c_int32), ("max_it", C.c_int32),
        ("tol_abs", C.c_double), ("tol_rel", C.c_double), ("device", C.c_int32),
    ]


class PoissonStats(C.Structure):
    _fields_ = [("iters", C.c_int32), ("outer", C.c_int32), ("fallbacks", C.c_int32),
                ("converged", C.c_int32), ("res0", C.c_double), ("res", C.c_double), ("ms", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class LsStats(C.Structure):
    _fields_ = [("vol_ini", C.c_double), ("vol_ins", C.c_double), ("reinit_iters", C.c_int32),
                ("mass_corr_iters", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_LIB = None


def lib():
    """Load libsuli.so (built in-tree); raises if it is absent."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise SuliError(f"{LIB_PATH} is not built; run `make -C {HERE}` or __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        i32, f64, p = C.c_int, C.c_double, C.POINTER
        sig = {
            "suli_init": ([p(Params)], i32), "suli_finalize": ([], i32),
            "suli_last_error": ([C.c_char_p, i32], i32), "suli_version": ([], C.c_char_p),
            "suli_upload": ([i32, p(f64), C.c_int64], i32), "suli_download": ([i32, p(f64), C.c_int64], i32),
            "suli_field_len": ([i32], C.c_int64),
            "suli_level_set_ini": ([p(f64)], i32), "suli_init_dambreak": ([f64, f64, p(f64)], i32),
            "suli_level_set": ([f64, p(LsStats)], i32), "suli_contorno": ([i32], i32),
            "suli_visco": ([], i32), "suli_convdiff": ([], i32), "suli_tempo": ([i32, f64], i32),
            "suli_graddin": ([f64, p(PoissonStats)], i32), "suli_posdin": ([f64], i32),
            "suli_est_div": ([p(f64)], i32), "suli_convdiff_forces": ([], i32),
            "suli_est_div_noalloc": ([C.c_void_p] * 3 + [i32] * 3 + [f64] * 3 + [C.c_void_p, C.c_void_p], i32),
            "suli_stage": ([i32, f64, p(PoissonStats)], i32),
            "suli_step": ([f64, p(LsStats), p(PoissonStats)], i32), "suli_sync": ([], i32),
            "suli_set_vol_ini": ([f64], i32), "suli_get_volumes": ([p(f64), p(f64)], i32),
            "suli_set_solver": ([i32, i32, f64, i32], i32),
            "suli_set_mg": ([C.c_int64, i32, i32], i32),
            "suli_set_fast_stencils": ([i32], i32),
            "suli_precond_apply": ([p(f64), p(f64), C.c_int64], i32),
            "suli_fcg_readbacks": ([p(C.c_int64), p(C.c_int64)], i32),
            "suli_set_forcing": ([f64, f64, f64], i32), "suli_set_time": ([f64], i32),
            "suli_prof_enable": ([i32], i32), "suli_prof_select": ([C.c_char_p, i32], i32),
            "suli_prof_count": ([], i32), "suli_prof_get": ([i32, C.c_char_p, i32, p(C.c_int64), p(f64)], i32),
            "suli_unet_nparams": ([], C.c_int64), "suli_unet_configure": ([i32], i32), "suli_unet_levels": ([], i32),
            "suli_unet_set_entry": ([i32], i32), "suli_unet_entry": ([], i32), "suli_unet_set_fcg_m": ([i32], i32),
            "suli_unet_init": ([C.c_uint64], i32),
            "suli_unet_set_weights": ([p(C.c_float), C.c_int64], i32),
            "suli_unet_get_weights": ([p(C.c_float), C.c_int64], i32),
            "suli_unet_capture": ([i32, i32], i32), "suli_unet_clear": ([], i32), "suli_unet_dataset_size": ([], i32),
            "suli_unet_train": ([i32, f64, C.c_uint64, p(f64)], i32),
            "suli_unet_loss_grad": ([p(f64), p(f64), p(f64), p(C.c_float), p(C.c_float)], i32),
            "suli_unet_dataset_dims": ([p(i32), p(i32), p(i32), p(i32)], i32),
            "suli_unet_dataset_get": ([i32, p(C.c_float), p(C.c_float), p(C.c_float)], i32),
            "suli_unet_sample_loss_grad": ([i32, p(f64), p(C.c_float), p(C.c_float)], i32),
            "suli_frame_values": ([i32, p(C.c_int64)], i32),
            "suli_frame_record": ([i32, p(C.c_float), C.c_int64], i32),
            "suli_write_frame": ([C.c_char_p, i32], i32),
            "suli_checkpoint_save": ([C.c_char_p], i32), "suli_checkpoint_load": ([C.c_char_p], i32),
            "suli_set_guess": ([i32], i32),
            "suli_set_time_scheme": ([i32], i32), "suli_time_stages": ([], i32),
            "suli_set_advection": ([i32, i32], i32),
            "suli_set_obstacle": ([C.c_void_p, C.c_void_p, C.c_void_p], i32),
            "suli_obstacle_columns": ([i32, i32, i32, i32, f64, f64, f64, C.c_void_p, C.c_void_p, C.c_void_p], i32),
            "suli_restart_write": ([C.c_char_p], i32), "suli_restart_read": ([C.c_char_p], i32),
            "suli_write_xdmf": ([C.c_char_p, i32, i32, i32, i32, i32, i32, f64], i32),
            "suli_mms_set": ([i32, f64, f64, f64], i32), "suli_mms_i": ([], i32), "suli_mms_bc": ([], i32),
            "suli_termo_fonte1": ([], i32), "suli_termo_fonte2": ([], i32), "suli_mms": ([p(f64)], i32),
            "suli_waves_coef": ([i32, f64, f64, f64, p(f64)], i32), "suli_boundary_waves": ([], i32),
        }
        for name, (args, res) in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _LIB = L
    return _LIB


def _check(rc: int, what: str):
    if rc != 0:
        buf = C.create_string_buffer(1024)
        lib().suli_last_error(buf, 1024)
        raise SuliError(f"{what} failed ({rc}): {buf.value.decode(errors='replace')}")


class Solver:
    """One SuLi-IC state resident on the GPU."""

    def __init__(self, nx, ny, nz, dx, dy=None, dz=None, *, ccx0=2, ccxf=2, ccy0=2, ccyf=2, ccz0=2, cczf=2,
                 m_turb=1, rho_f1=1.2, rho_f2=1000.0, mi_f1=1.8e-5, mi_f2=1.0e-3, sigma=0.0728, alpha1=1.5,
                 csmag=0.1, gx=0.0, gz=9.81, chezy=44.5, precision=PREC_FP64, poisson=POISSON_REF,
                 precond=PRECOND_JACOBI, max_it=0, tol_abs=0.0, tol_rel=0.0, device=0):
        dy = dx if dy is None else dy
        dz = dx if dz is None else dz
        self.p = Params(nx, ny, nz, dx, dy, dz, ccx0, ccxf, ccy0, ccyf, ccz0, cczf, m_turb, rho_f1, rho_f2,
                        mi_f1, mi_f2, sigma, alpha1, csmag, gx, gz, chezy, precision, poisson, precond,
                        max_it, tol_abs, tol_rel, device)
        _check(lib().suli_init(C.byref(self.p)), "suli_init")
        self.nx, self.ny, self.nz = nx, ny, nz

    @classmethod
    def from_oracle_params(cls, op, **kw):
        f = {k: getattr(op, k) for k in ["nx", "ny", "nz", "dx", "dy", "dz", "ccx0", "ccxf", "ccy0", "ccyf",
                                         "ccz0", "cczf", "m_turb", "rho_f1", "rho_f2", "mi_f1", "mi_f2",
                                         "sigma", "alpha1", "csmag", "gx", "gz", "chezy"]}
        f.update(kw)
        nx, ny, nz, dx = f.pop("nx"), f.pop("ny"), f.pop("nz"), f.pop("dx")
        return cls(nx, ny, nz, dx, **f)

    def close(self):
        lib().suli_finalize()

    def shape(self, name):
        if name in PLANES:
            return plane_shape(name, self.nx, self.ny, self.nz)
        return extent_shape(KIND.get(name, "C"), self.nx, self.ny, self.nz)

    def upload(self, name, arr):
        a = np.asfortranarray(np.asarray(arr, dtype=np.float64))
        assert a.shape == self.shape(name), (name, a.shape, self.shape(name))
        flat = a.ravel(order="F")
        _check(lib().suli_upload(FIELD_ID[name], flat.ctypes.data_as(C.POINTER(C.c_double)), flat.size),
               f"suli_upload({name})")

    def download(self, name):
        shp = self.shape(name)
        flat = np.zeros(int(np.prod(shp)), dtype=np.float64)
        _check(lib().suli_download(FIELD_ID[name], flat.ctypes.data_as(C.POINTER(C.c_double)), flat.size),
               f"suli_download({name})")
        return flat.reshape(shp, order="F")

    # -- reference subroutines ------------------------------------------------
    def level_set_ini(self):
        v = C.c_double()
        _check(lib().suli_level_set_ini(C.byref(v)), "suli_level_set_ini")
        return v.value

    def set_vol_ini(self, v):
        _check(lib().suli_set_vol_ini(v), "suli_set_vol_ini")

    def volumes(self):
        a, b = C.c_double(), C.c_double()
        _check(lib().suli_get_volumes(C.byref(a), C.byref(b)), "suli_get_volumes")
        return a.value, b.value

    def init_dambreak(self, width, height):
        v = C.c_double()
        _check(lib().suli_init_dambreak(width, height, C.byref(v)), "suli_init_dambreak")
        return v.value

    def level_set(self, dt):
        st = LsStats()
        _check(lib().suli_level_set(dt, C.byref(st)), "suli_level_set")
        return st.as_dict()

    def contorno(self, nlock):
        _check(lib().suli_contorno(nlock), "suli_contorno")

    def visco(self):
        _check(lib().suli_visco(), "suli_visco")

    def convdiff(self):
        _check(lib().suli_convdiff(), "suli_convdiff")

    def convdiff_forces(self):
        _check(lib().suli_convdiff_forces(), "suli_convdiff_forces")

    def tempo(self, tt, dt):
        _check(lib().suli_tempo(tt, dt), "suli_tempo")

    def graddin(self, dt):
        st = PoissonStats()
        _check(lib().suli_graddin(dt, C.byref(st)), "suli_graddin")
        return st.as_dict()

    def posdin(self, dt):
        _check(lib().suli_posdin(dt), "suli_posdin")

    def est_div(self):
        v = C.c_double()
        _check(lib().suli_est_div(C.byref(v)), "suli_est_div")
        return v.value

    def stage(self, tt, dt_base):
        st = PoissonStats()
        _check(lib().suli_stage(tt, dt_base, C.byref(st)), "suli_stage")
        return st.as_dict()

    def step(self, dt):
        ls = LsStats()
        ps = (PoissonStats * 3)()
        _check(lib().suli_step(dt, C.byref(ls), ps), "suli_step")
        return ls.as_dict(), [ps[q].as_dict() for q in range(lib().suli_time_stages())]

    def sync(self):
        _check(lib().suli_sync(), "suli_sync")

    def set_forcing(self, amplitude, period, phase=0.0):
        _check(lib().suli_set_forcing(amplitude, period, phase), "suli_set_forcing")

    def set_time(self, t):
        _check(lib().suli_set_time(t), "suli_set_time")

    # -- the manufactured solution of 13_mms.f90 --------------------------------
    def mms_set(self, mms_t, a=0.0, h0=2.0, coef=0.0):
        _check(lib().suli_mms_set(mms_t, a, h0, coef), "suli_mms_set")

    def mms_i(self):
        _check(lib().suli_mms_i(), "suli_mms_i")

    def mms_bc(self):
        _check(lib().suli_mms_bc(), "suli_mms_bc")

    def termo_fonte(self, mms_t):
        fn = lib().suli_termo_fonte1 if mms_t == 1 else lib().suli_termo_fonte2
        _check(fn(), f"suli_termo_fonte{mms_t}")

    def mms(self):
        e = (C.c_double * 4)()
        _check(lib().suli_mms(e), "suli_mms")
        return tuple(e)

    # -- wave inflow (3_contorno.f90:952-1205) -----------------------------------
    def waves_coef(self, wave_t, a_w=0.03, p_w=2.5, h0_f=1.0):
        """Returns l_w, n_w, f_w, c_w, aeta1..5, avel1..5."""
        c = (C.c_double * 14)()
        _check(lib().suli_waves_coef(wave_t, a_w, p_w, h0_f, c), "suli_waves_coef")
        return np.array(c[:])

    def boundary_waves(self):
        _check(lib().suli_boundary_waves(), "suli_boundary_waves")

    def set_solver(self, poisson, precond, tol_rel=0.0, max_it=0):
        _check(lib().suli_set_solver(poisson, precond, tol_rel, max_it), "suli_set_solver")

    def set_mg(self, fuse_max_cells=-1, coarse_sweeps=-1, fuse_fine=-1):
        """Multigrid tuning: fuse the half sweeps of levels of at most
        fuse_max_cells cells (0: off); sweeps on the coarsest level; scaling,
        prolongation and z.q inside the neighbouring half sweeps (1; 2, the
        default: also the finest level's iterate and right-hand side in FP16
        in the FCG iteration's V-cycles; 0: separate launches); -1 keeps."""
        _check(lib().suli_set_mg(int(fuse_max_cells), int(coarse_sweeps), int(fuse_fine)), "suli_set_mg")

    def set_fast_stencils(self, on):
        """Momentum predictor in LDS tiles (1, default) or with every stencil
        value read from memory (0); the same values either way."""
        _check(lib().suli_set_fast_stencils(int(on)), "suli_set_fast_stencils")

    def fcg_readbacks(self):
        """(published, copied): host reads of the FCG control block that found
        the finish kernel's copy in pinned memory, and those that copied."""
        a, b = C.c_int64(), C.c_int64()
        _check(lib().suli_fcg_readbacks(C.byref(a), C.byref(b)), "suli_fcg_readbacks")
        return a.value, b.value

    def precond_apply(self, r):
        """z = M r with the current preconditioner and the operator of the last
        FCG graddin; r as a [k, j, i] array of the scaled residual."""
        r = np.ascontiguousarray(r, dtype=np.float64).ravel()
        n = self.nx * self.ny * self.nz
        if r.size != n:
            raise SuliError(f"precond_apply: expected {n} cells")
        z = np.empty(n, dtype=np.float64)
        P = C.POINTER
        _check(lib().suli_precond_apply(r.ctypes.data_as(P(C.c_double)), z.ctypes.data_as(P(C.c_double)), n),
               "suli_precond_apply")
        return z.reshape(self.nz, self.ny, self.nx)

    # ---- learned preconditioner -------------------------------------------
    def unet_nparams(self):
        return int(lib().suli_unet_nparams())

    def unet_configure(self, levels):
        _check(lib().suli_unet_configure(levels), "suli_unet_configure")

    def unet_levels(self):
        return int(lib().suli_unet_levels())

    def unet_set_entry(self, level):
        _check(lib().suli_unet_set_entry(level), "suli_unet_set_entry")

    def unet_entry(self):
        return int(lib().suli_unet_entry())

    def unet_set_fcg_m(self, m):
        _check(lib().suli_unet_set_fcg_m(m), "suli_unet_set_fcg_m")

    def unet_init(self, seed):
        _check(lib().suli_unet_init(seed), "suli_unet_init")

    def unet_set_weights(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        _check(lib().suli_unet_set_weights(w.ctypes.data_as(C.POINTER(C.c_float)), w.size),
               "suli_unet_set_weights")

    def unet_get_weights(self):
        w = np.empty(self.unet_nparams(), dtype=np.float32)
        _check(lib().suli_unet_get_weights(w.ctypes.data_as(C.POINTER(C.c_float)), w.size),
               "suli_unet_get_weights")
        return w

    def unet_capture(self, max_samples, every=0):
        _check(lib().suli_unet_capture(max_samples, every), "suli_unet_capture")

    def unet_clear(self):
        _check(lib().suli_unet_clear(), "suli_unet_clear")

    def unet_dataset_size(self):
        return int(lib().suli_unet_dataset_size())

    def unet_train(self, iters, lr, seed=0):
        v = C.c_double()
        _check(lib().suli_unet_train(iters, lr, seed, C.byref(v)), "suli_unet_train")
        return v.value

    def unet_loss_grad(self, r, rho):
        """(loss, gradient, z) for one sample; r, rho as [k, j, i] arrays."""
        r = np.ascontiguousarray(r, dtype=np.float64).ravel()
        rho = np.ascontiguousarray(rho, dtype=np.float64).ravel()
        n = self.nx * self.ny * self.nz
        if r.size != n or rho.size != n:
            raise SuliError(f"unet_loss_grad: expected {n} cells")
        loss = C.c_double()
        g = np.empty(self.unet_nparams(), dtype=np.float32)
        z = np.empty(n, dtype=np.float32)
        P = C.POINTER
        _check(lib().suli_unet_loss_grad(r.ctypes.data_as(P(C.c_double)), rho.ctypes.data_as(P(C.c_double)),
                                         C.byref(loss), g.ctypes.data_as(P(C.c_float)),
                                         z.ctypes.data_as(P(C.c_float))), "suli_unet_loss_grad")
        return loss.value, g, z.reshape(self.nz, self.ny, self.nx)

    def frame_record(self, t_plot=1):
        """The FP32 record of a field-output file (plot_f) as a flat array."""
        n = C.c_int64()
        _check(lib().suli_frame_values(t_plot, C.byref(n)), "suli_frame_values")
        buf = np.empty(n.value, dtype=np.float32)
        _check(lib().suli_frame_record(t_plot, buf.ctypes.data_as(C.POINTER(C.c_float)), buf.size),
               "suli_frame_record")
        return buf

    def write_frame(self, path, t_plot=1):
        _check(lib().suli_write_frame(os.fsencode(path), t_plot), "suli_write_frame")

    def set_guess(self, mode):
        _check(lib().suli_set_guess(mode), "suli_set_guess")

    def set_time_scheme(self, t_tempo):
        """Euler, RK2, RK3 or AB2 (t_tempo 0..3, 11_diversos.f90:10-228)."""
        _check(lib().suli_set_time_scheme(t_tempo), "suli_set_time_scheme")

    def time_stages(self):
        return lib().suli_time_stages()

    def set_obstacle(self, ku, kv, kw):
        """Bottom-obstacle columns ku (nx+3, ny+2), kv (nx+2, ny+3), kw (nx+2, ny+2)
        in the reference's index order; None for all three removes it."""
        if ku is None and kv is None and kw is None:
            _check(lib().suli_set_obstacle(None, None, None), "suli_set_obstacle")
            return
        shapes = [(self.nx + 3, self.ny + 2), (self.nx + 2, self.ny + 3), (self.nx + 2, self.ny + 2)]
        arrs = []
        for a, shp in zip((ku, kv, kw), shapes):
            a = np.asarray(a)
            assert a.shape == shp, (a.shape, shp)
            arrs.append(np.ascontiguousarray(np.rint(a).astype(np.int32).ravel(order="F")))
        _check(lib().suli_set_obstacle(*[a.ctypes.data for a in arrs]), "suli_set_obstacle")

    def set_advection(self, adv_type, der):
        """classico (adv_type 1, der 1..3), rotacional (2, 1), antissim (3, 1)."""
        _check(lib().suli_set_advection(adv_type, der), "suli_set_advection")

    def checkpoint_save(self, path):
        _check(lib().suli_checkpoint_save(os.fsencode(path)), "suli_checkpoint_save")

    def checkpoint_load(self, path):
        _check(lib().suli_checkpoint_load(os.fsencode(path)), "suli_checkpoint_load")

    def restart_write(self, path):
        _check(lib().suli_restart_write(os.fsencode(path)), "suli_restart_write")

    def restart_read(self, path):
        _check(lib().suli_restart_read(os.fsencode(path)), "suli_restart_read")

    def unet_dataset_dims(self):
        """(nx, ny, nz, entry) of the held training set."""
        v = [C.c_int() for _ in range(4)]
        _check(lib().suli_unet_dataset_dims(*[C.byref(x) for x in v]), "suli_unet_dataset_dims")
        return tuple(x.value for x in v)

    def unet_dataset_get(self, idx):
        """Sample idx: (r, rho, faces) with r, rho as [k, j, i] arrays and
        faces = (cx, cy, cz), each with n+1 faces along its own axis, or None
        for samples of the whole grid."""
        nx, ny, nz, entry = self.unet_dataset_dims()
        n = nx * ny * nz
        r = np.empty(n, dtype=np.float32)
        rho = np.empty(n, dtype=np.float32)
        nfx, nfy, nfz = (nx + 1) * ny * nz, nx * (ny + 1) * nz, nx * ny * (nz + 1)
        f = np.empty(nfx + nfy + nfz if entry > 0 else 1, dtype=np.float32)
        P = C.POINTER
        _check(lib().suli_unet_dataset_get(idx, r.ctypes.data_as(P(C.c_float)), rho.ctypes.data_as(P(C.c_float)),
                                           f.ctypes.data_as(P(C.c_float))), "suli_unet_dataset_get")
        faces = None
        if entry > 0:
            faces = (f[:nfx].reshape(nz, ny, nx + 1), f[nfx:nfx + nfy].reshape(nz, ny + 1, nx),
                     f[nfx + nfy:].reshape(nz + 1, ny, nx))
        return r.reshape(nz, ny, nx), rho.reshape(nz, ny, nx), faces

    def unet_sample_loss_grad(self, idx):
        """(loss, gradient, z) of held sample idx through its training operator."""
        nx, ny, nz, _ = self.unet_dataset_dims()
        loss = C.c_double()
        g = np.empty(self.unet_nparams(), dtype=np.float32)
        z = np.empty(nx * ny * nz, dtype=np.float32)
        P = C.POINTER
        _check(lib().suli_unet_sample_loss_grad(idx, C.byref(loss), g.ctypes.data_as(P(C.c_float)),
                                                z.ctypes.data_as(P(C.c_float))), "suli_unet_sample_loss_grad")
        return loss.value, g, z.reshape(nz, ny, nx)


def write_xdmf(path, nx, ny, nz, first, last, t_plot=1, dt_frame=1.0):
    """The ParaView index of the records campos_<first>..campos_<last>."""
    _check(lib().suli_write_xdmf(os.fsencode(path), nx, ny, nz, first, last, t_plot, dt_frame), "suli_write_xdmf")


def load_state(solver: Solver, state, names=None):
    """Upload an oracle State (reference layout) into the GPU solver."""
    names = names or ["u", "v", "w", "prd1", "rho", "ls_nu", "ls", "mod_ls", "kurv", "hs", "hsx", "hsy",
                      "hsz", "nut", "u0", "v0", "w0"]
    for n in names:
        solver.upload(n, state.arr(n))
    if state.scal.get("vol_ini", 0) > 0:
        solver.set_vol_ini(state.scal["vol_ini"])


def obstacle_columns(obst_t, nx, ny, nz, dx, dy, dz):
    """ku (nx+3, ny+2), kv (nx+2, ny+3), kw (nx+2, ny+2) of the reference's
    obstacle geometry obst_t (obstaculo, 3_contorno.f90:494-760)."""
    shapes = [(nx + 3, ny + 2), (nx + 2, ny + 3), (nx + 2, ny + 2)]
    out = [np.zeros(int(np.prod(s)), dtype=np.int32) for s in shapes]
    _check(lib().suli_obstacle_columns(obst_t, nx, ny, nz, dx, dy, dz, *[a.ctypes.data for a in out]),
           "suli_obstacle_columns")
    return [a.reshape(s, order="F") for a, s in zip(out, shapes)]


WEIGHTS_DIR = os.path.join(HERE, "weights")


def shipped_learned(weights_dir=WEIGHTS_DIR):
    """The learned preconditioners shipped in weights/: one dict per
    <name>.json that has its <name>.npy beside it, with the network depth
    (levels), the multigrid level it solves (entry, 0 = the whole grid) and
    whether it is the default one; the default comes first."""
    import json
    out = []
    for f in sorted(os.listdir(weights_dir)) if os.path.isdir(weights_dir) else []:
        if not f.endswith(".json"):
            continue
        npy = os.path.join(weights_dir, f[:-5] + ".npy")
        if not os.path.exists(npy):
            continue
        meta = json.load(open(os.path.join(weights_dir, f)))
        out.append({"name": f[:-5], "levels": int(meta["levels"]), "entry": int(meta.get("entry", 0)),
                    "default": bool(meta.get("default", False)), "path": npy})
    out.sort(key=lambda m: (not m["default"], m["entry"], m["name"]))
    return out


def use_learned(solver, item):
    """Load a shipped learned preconditioner (an item of shipped_learned)."""
    solver.unet_configure(item["levels"])
    solver.unet_set_entry(item["entry"])
    solver.unet_set_weights(np.load(item["path"]))
