"""suli_est_div_noalloc (the allocation-free variant of suli_est_div) called
through the C ABI on device arrays of the caller, against suli_est_div on the
same velocities: two calls in a row, the second on a smaller grid (the
shared scratch buffer then holds the first call's partials beyond the
second's)."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
import pysuli
from helpers import consistent_state

pytestmark = pytest.mark.gpu


def _hip():
    """The HIP runtime libsuli.so is linked against (already loaded with it)."""
    pysuli.lib()
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipDeviceSynchronize.argtypes = []
    for f in (h.hipMalloc, h.hipMemcpy, h.hipFree, h.hipDeviceSynchronize):
        f.restype = C.c_int
    return h


H2D, D2H = 1, 2


class _Dev:
    """A device copy of a host array (test-side allocation)."""

    def __init__(self, h, a):
        self.h = h
        a = np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel(order="F"))
        self.n = a.size
        self.p = C.c_void_p()
        assert h.hipMalloc(C.byref(self.p), max(8 * self.n, 8)) == 0
        assert h.hipMemcpy(self.p, a.ctypes.data, 8 * self.n, H2D) == 0

    def get(self):
        out = np.empty(self.n)
        assert self.h.hipMemcpy(out.ctypes.data, self.p, 8 * self.n, D2H) == 0
        return out

    def free(self):
        self.h.hipFree(self.p)


def _noalloc(h, st):
    fn = pysuli.lib().suli_est_div_noalloc
    p = st.p
    bufs = [_Dev(h, st.u), _Dev(h, st.v), _Dev(h, st.w), _Dev(h, np.full(1, -1.0))]
    try:
        rc = fn(bufs[0].p, bufs[1].p, bufs[2].p, p.nx, p.ny, p.nz, p.dx, p.dy, p.dz, bufs[3].p, None)
        pysuli._check(rc, "suli_est_div_noalloc")
        assert h.hipDeviceSynchronize() == 0
        return float(bufs[3].get()[0])
    finally:
        for b in bufs:
            b.free()


def test_noalloc_matches_est_div():
    h = _hip()
    grids = [O.Params(nx=40, ny=24, nz=20, dx=1 / 40, dy=1 / 48, dz=1 / 40),
             O.Params(nx=12, ny=8, nz=6, dx=1 / 12, dy=1 / 12, dz=1 / 12)]
    for q, p in enumerate(grids):
        print(f"est_div_noalloc call {q + 1}: grid {p.nx}x{p.ny}x{p.nz}", flush=True)
        st = consistent_state(p, seed=40 + q)
        st.u[3, 2, 2] += 5.0 * (q + 1)  # a distinct maximum per call
        sv = pysuli.Solver.from_oracle_params(p, precision=pysuli.PREC_FP64)
        try:
            pysuli.load_state(sv, st)
            want = sv.est_div()
        finally:
            sv.close()
        got = _noalloc(h, st)
        print(f"  suli_est_div {want!r}  suli_est_div_noalloc {got!r}  oracle {st.est_div()!r}")
        assert got == want
        assert abs(got - st.est_div()) <= 1e-13 * st.est_div()


def test_noalloc_rejects_bad_grids():
    L = pysuli.lib()
    fn = L.suli_est_div_noalloc
    one = C.c_void_p(8)  # never dereferenced: the grid check comes first
    for nx, ny, nz in ((0, 4, 4), (4, 1 << 16, 1 << 15), (1 << 20, 1 << 10, 1 << 10)):
        assert fn(one, one, one, nx, ny, nz, 1.0, 1.0, 1.0, one, None) != 0
        buf = C.create_string_buffer(512)
        L.suli_last_error(buf, 512)
        assert b"outside the supported range" in buf.value, buf.value
