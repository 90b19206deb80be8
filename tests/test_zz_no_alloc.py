"""ssam_jacobi19_noalloc: the allocation-free Jacobi entry (residual partials in
the library's fixed device array, the norm written to a device double) must
give exactly ssam_jacobi19's sweeps and residual, call after call, also when the
second call is smaller than the first."""
import ctypes

import numpy as np
import pytest
import torch

from oracle import oracle
from ssam_amd import _lib as L

pytestmark = pytest.mark.gpu


def _vp(t):
    return ctypes.c_void_p(t.data_ptr() if t is not None else 0)


def _both(x, steps, dtype):
    nz, ny, nx = x.shape
    src = torch.from_numpy(x).cuda()
    work = torch.empty_like(src) if steps > 1 else None
    d1 = torch.full_like(src, float("nan"))
    r1 = ctypes.c_double(0.0)
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    code = L.F64 if dtype == np.float64 else L.F32
    L.check(L.lib().ssam_jacobi19(code, _vp(src), _vp(d1), nx, ny, nz, steps, _vp(work),
                                  ctypes.byref(r1), st), "ssam_jacobi19")
    d2 = torch.full_like(src, float("nan"))
    r2 = torch.full((1,), float("nan"), dtype=torch.float64, device="cuda")
    L.check(L.lib().ssam_jacobi19_noalloc(code, _vp(src), _vp(d2), nx, ny, nz, steps, _vp(work),
                                          _vp(r2), st), "ssam_jacobi19_noalloc")
    torch.cuda.synchronize()
    return d1.cpu().numpy(), r1.value, d2.cpu().numpy(), float(r2.item())


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_noalloc_matches_allocating_entry_twice(dtype):
    # a larger call, then a smaller one: the second reuses the array with fewer tiles
    for i, (shape, steps) in enumerate([((70, 96, 130), 5), ((29, 33, 40), 2)]):
        x = oracle.uniform(500 + i, int(np.prod(shape)), 0.0, 1.0, dtype).reshape(shape)
        d1, r1, d2, r2 = _both(x, steps, dtype)
        assert np.array_equal(d1, d2), f"call {i}: sweeps differ"
        assert r1 == r2, f"call {i}: residual {r1} vs {r2}"
        # and against the oracle: the sweeps (values in [0, 1), weights summing to
        # 1, so |error| <= steps * tol) and the FP64 residual of the last input
        cur, rsq = x, 0.0
        for _ in range(steps):
            cur, rsq = oracle.jacobi19(cur)
        tol = (1e-5 if dtype == np.float32 else 1e-12) * steps
        assert np.max(np.abs(d2.astype(np.float64) - cur)) <= tol, f"call {i}: sweeps vs oracle"
        rel = 1e-4 if dtype == np.float32 else 1e-10
        assert abs(r2 - np.sqrt(rsq)) <= rel * np.sqrt(rsq), (i, r2, np.sqrt(rsq))


def test_noalloc_graph_replay():
    # the allocation-free call records into a graph (no host round trip): every
    # replay matches the eager call
    x = oracle.uniform(77, 40 * 48 * 52, 0.0, 1.0, np.float32).reshape(40, 48, 52)
    d1, r1, _, _ = _both(x, 3, np.float32)
    nz, ny, nx = x.shape
    src = torch.from_numpy(x).cuda()
    work = torch.empty_like(src)
    dst = torch.empty_like(src)
    res = torch.zeros(1, dtype=torch.float64, device="cuda")
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        L.check(L.lib().ssam_jacobi19_noalloc(L.F32, _vp(src), _vp(dst), nx, ny, nz, 3, _vp(work),
                                              _vp(res), st), "ssam_jacobi19_noalloc (capture)")
    for rep in range(3):
        dst.fill_(float("nan"))
        res.fill_(float("nan"))
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(dst.cpu().numpy(), d1), f"replay {rep}"
        assert float(res.item()) == r1, f"replay {rep}"
