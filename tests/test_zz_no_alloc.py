"""fv3d_mlp_grad_noalloc (no reference counterpart): the gradient of the
closure's loss with caller-owned device buffers and the library's static
scratch, against fv3d_mlp_grad (host buffers, allocated per call) on the
same samples: bit for bit, on two calls in a row, the second with fewer
samples (so stale partials of the first would show), and the size limit.

The caller's device buffers come from the HIP runtime the library itself
links (libamdhip64.so.7), so one runtime serves the whole process."""
import ctypes as C

import numpy as np
import pytest

from fv3d import cases, closure
from fv3d.solver import GpuSolver

pytestmark = pytest.mark.gpu


class Hip:
    """hipMalloc / hipMemcpy / hipFree of the runtime libfv3d links."""

    H2D, D2H = 1, 2

    def __init__(self):
        self.L = C.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipDeviceSynchronize.argtypes = []
        self.live = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), nbytes) == 0
        self.live.append(p)
        return p

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        assert self.L.hipMemcpy(p, a.ctypes.data, a.nbytes, self.H2D) == 0
        return p

    def get(self, p, n, dtype=np.float32):
        out = np.empty(n, dtype=dtype)
        assert self.L.hipMemcpy(out.ctypes.data, p, out.nbytes, self.D2H) == 0
        return out

    def free(self):
        for p in self.live:
            self.L.hipFree(p)
        self.live = []


def test_grad_noalloc_matches_allocating_entry_point():
    case = cases.bump_channel(ni=6, nj=6, nk=8, seed=3, max_iter=2)
    g = GpuSolver(case)
    hip = Hip()
    rng = np.random.default_rng(11)
    g.mlp_set(closure.init_params(5), None)
    nan = np.full(closure.N_PARAMS, np.nan, dtype=np.float32)
    for n in (9000, 1234):
        print(f"no-alloc gradient, {n} samples", flush=True)
        x = rng.normal(size=(n, closure.N_IN)).astype(np.float32)
        t = rng.uniform(0.0, 3.0, n).astype(np.float32)
        want_g, want_l = g.mlp_grad(x, t)
        xd, td = hip.put(x), hip.put(t)
        gd, ld = hip.put(nan), hip.put(nan[:1])
        rc = g.L.fv3d_mlp_grad_noalloc(g.h, xd, td, n, gd, ld)
        assert rc == 0, g.L.fv3d_last_error(g.h).decode()
        g.sync()
        got_g, got_l = hip.get(gd, closure.N_PARAMS), float(hip.get(ld, 1)[0])
        np.testing.assert_array_equal(got_g, want_g)
        assert got_l == float(want_l)
    # beyond the scratch: rejected, nothing launched
    big, tb = hip.alloc(16385 * closure.N_IN * 4), hip.alloc(16385 * 4)
    rc = g.L.fv3d_mlp_grad_noalloc(g.h, big, tb, 16385, gd, ld)
    assert rc < 0 and "at most 16384" in g.L.fv3d_last_error(g.h).decode()
    hip.free()
    g.close()
