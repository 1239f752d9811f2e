"""hlbl_qed_kernel_noalloc (the QED kernel at points with its scratch in the library's
fixed device array, no allocation) against hlbl_qed_kernel (per-call scratch), through
the C ABI: two calls in a row, the second with fewer points, each bitwise equal to the
allocating entry point; sizes beyond the array's maximum are refused."""
import ctypes

import numpy as np
import pytest

import hlbl_amd as H

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

MAX_POINTS = 2048


@pytest.fixture(scope="module")
def qed():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    q = H.QedKernel(0)
    yield q
    q.close()


def points(rng, n):
    x = rng.uniform(-3, 3, (n, 4))
    y = rng.uniform(-3, 3, (n, 4))
    x[0] = 0
    y[1] = 0
    return x, y


def test_noalloc_matches_allocating_entry(qed):
    rng = np.random.default_rng(42)
    for n in (1500, 300):
        print("noalloc call with", n, "points", flush=True)
        x, y = points(rng, n)
        dx, dy = torch.tensor(x, device="cuda"), torch.tensor(y, device="cuda")
        want = torch.zeros((n, 384), dtype=torch.float64, device="cuda")
        got = torch.full((n, 384), np.nan, dtype=torch.float64, device="cuda")
        qed.kernel(dx.data_ptr(), dy.data_ptr(), n, want.data_ptr())
        qed.kernel_noalloc(dx.data_ptr(), dy.data_ptr(), n, got.data_ptr())
        torch.cuda.synchronize()
        assert torch.equal(got, want), n
        assert torch.abs(want).max().item() > 0


def test_noalloc_size_limits(qed):
    rng = np.random.default_rng(3)
    x, y = points(rng, MAX_POINTS)
    dx, dy = torch.tensor(x, device="cuda"), torch.tensor(y, device="cuda")
    out = torch.zeros((MAX_POINTS + 1, 384), dtype=torch.float64, device="cuda")
    qed.kernel_noalloc(dx.data_ptr(), dy.data_ptr(), MAX_POINTS, out.data_ptr())  # the maximum
    with pytest.raises(H.HlblError, match="at most 2048"):
        qed.kernel_noalloc(dx.data_ptr(), dy.data_ptr(), MAX_POINTS + 1, out.data_ptr())
    qed.kernel_noalloc(dx.data_ptr(), dy.data_ptr(), 0, out.data_ptr())  # nothing to do
    torch.cuda.synchronize()
