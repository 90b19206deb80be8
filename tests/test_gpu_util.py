"""Device utilities through the C ABI: the seeded generator equals the oracle's
(bit for bit), and the stream-copy bandwidth kernel copies exactly."""
import numpy as np
import pytest
import torch

from oracle import oracle
from ssam_amd import api

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype,np_dt", [(torch.float32, np.float32), (torch.float64, np.float64)])
@pytest.mark.parametrize("n,offset", [(1, 0), (1000, 5), (1 << 20, 12345)])
def test_fill_uniform_matches_oracle(dtype, np_dt, n, offset):
    t = torch.empty(n, dtype=dtype, device="cuda")
    api.fill_uniform(t, 42, -1.0, 1.0, offset)
    want = oracle.uniform(42, n, -1.0, 1.0, np_dt, offset)
    assert np.array_equal(t.cpu().numpy(), want)


@pytest.mark.parametrize("dtype", [torch.bfloat16, torch.float16])
def test_fill_uniform_reduced_storage_rounds_nearest(dtype):
    n = 4096
    t = torch.empty(n, dtype=dtype, device="cuda")
    api.fill_uniform(t, 9, 0.0, 1.0)
    want = torch.from_numpy(oracle.uniform(9, n, 0.0, 1.0, np.float32)).to(dtype)
    assert torch.equal(t.cpu(), want)


def test_stream_copy_exact():
    a = torch.empty(3 * (1 << 20) + 4, dtype=torch.float32, device="cuda")
    api.fill_uniform(a, 1)
    b = torch.zeros_like(a)
    api.stream_copy(a, b)
    torch.cuda.synchronize()
    assert torch.equal(a, b)
