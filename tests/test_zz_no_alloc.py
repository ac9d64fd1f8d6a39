"""fv3d_mlp_grad_noalloc (no reference counterpart): the gradient of the
closure's loss with caller-owned device buffers and the library's static
scratch, against fv3d_mlp_grad (host buffers, allocated per call) on the
same samples: bit for bit, on two calls in a row, the second with fewer
samples (so stale partials of the first would show), and the size limit."""
import ctypes as C

import numpy as np
import pytest

from fv3d import cases, closure
from fv3d.solver import GpuSolver

pytestmark = pytest.mark.gpu


def test_grad_noalloc_matches_allocating_entry_point():
    import torch
    case = cases.bump_channel(ni=6, nj=6, nk=8, seed=3, max_iter=2)
    g = GpuSolver(case)
    rng = np.random.default_rng(11)
    g.mlp_set(closure.init_params(5), None)
    dev = torch.device("cuda", 0)
    for n in (9000, 1234):
        print(f"no-alloc gradient, {n} samples", flush=True)
        x = rng.normal(size=(n, closure.N_IN)).astype(np.float32)
        t = rng.uniform(0.0, 3.0, n).astype(np.float32)
        want_g, want_l = g.mlp_grad(x, t)
        xd = torch.from_numpy(x).to(dev)
        td = torch.from_numpy(t).to(dev)
        gd = torch.full((closure.N_PARAMS,), float("nan"), dtype=torch.float32, device=dev)
        ld = torch.full((1,), float("nan"), dtype=torch.float32, device=dev)
        torch.cuda.synchronize()
        rc = g.L.fv3d_mlp_grad_noalloc(g.h, C.c_void_p(xd.data_ptr()), C.c_void_p(td.data_ptr()), n,
                                       C.c_void_p(gd.data_ptr()), C.c_void_p(ld.data_ptr()))
        assert rc == 0, g.L.fv3d_last_error(g.h).decode()
        g.sync()
        got_g, got_l = gd.cpu().numpy(), float(ld.cpu().numpy()[0])
        np.testing.assert_array_equal(got_g, want_g)
        assert got_l == float(want_l)
    # beyond the scratch: rejected, nothing launched
    big = torch.zeros((16385, closure.N_IN), dtype=torch.float32, device=dev)
    tb = torch.zeros(16385, dtype=torch.float32, device=dev)
    rc = g.L.fv3d_mlp_grad_noalloc(g.h, C.c_void_p(big.data_ptr()), C.c_void_p(tb.data_ptr()), 16385,
                                   C.c_void_p(gd.data_ptr()), C.c_void_p(ld.data_ptr()))
    assert rc < 0 and "at most 16384" in g.L.fv3d_last_error(g.h).decode()
    g.close()
