"""Shared helpers of the parity tests: oracle states and GPU solvers built
from the same inputs, and array comparison with stated tolerances."""
import numpy as np

import oracle as O

CELL = ["rho", "ls_nu", "ls", "mod_ls", "kurv", "hs", "hsx", "hsy", "hsz", "ddlsdx", "ddlsdy", "ddlsdz"]
VEL = ["u", "v", "w"]


def rel_err(a, b):
    """max |a-b| / max(|b|, tiny), over the whole array."""
    a = np.asarray(a)
    b = np.asarray(b)
    sc = max(float(np.max(np.abs(b))) if b.size else 0.0, 1e-300)
    return float(np.max(np.abs(a - b))) / sc if b.size else 0.0


def consistent_state(p, seed=1, **kw):
    """Oracle state with derived level-set fields, eddy viscosity and ghost
    planes consistent, as they are after contorno(1) in the driver loop."""
    s = O.random_state(p, seed=seed, **kw)
    s.run("visco")
    s.run("contorno", 1)
    return s


def gpu_from_state(state, **kw):
    import pysuli
    sv = pysuli.Solver.from_oracle_params(state.p, **kw)
    pysuli.load_state(sv, state)
    return sv
