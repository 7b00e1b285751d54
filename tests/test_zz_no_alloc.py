"""exmi_irfft_noalloc (the allocation-free variant of exmi_irfft) against
exmi_irfft through the C ABI: two calls in a row, the second with a smaller
input, in 1-D, 2-D and 3-D; a request larger than the scratch array fails
before any work is enqueued; numpy's irfftn as the independent check."""

import numpy as np
import pytest

from exponax_amd import _lib

pytestmark = pytest.mark.gpu


def _spectra(D, N, count, seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((count,) + (N,) * D)
    return np.fft.rfftn(x, axes=tuple(range(1, D + 1))).astype(np.complex64)


def _irfft_both(D, N, xh_np, device):
    import torch

    lib = _lib.lib()
    count = xh_np.shape[0]
    xh = torch.tensor(xh_np, device=device)
    a = torch.full((count,) + (N,) * D, float("nan"), dtype=torch.float32, device=device)
    b = torch.full_like(a, float("nan"))
    st = _lib.stream_ptr()
    _lib.check(lib.exmi_irfft(D, N, count, xh.data_ptr(), a.data_ptr(), st), "exmi_irfft")
    _lib.check(lib.exmi_irfft_noalloc(D, N, count, xh.data_ptr(), b.data_ptr(), st), "exmi_irfft_noalloc")
    torch.cuda.synchronize()
    return a.cpu().numpy(), b.cpu().numpy()


@pytest.mark.parametrize("D,big,small", [(1, (1024, 64), (256, 8)), (2, (256, 20), (64, 3)),
                                         (3, (64, 6), (32, 2))])
def test_irfft_noalloc_matches_irfft(device, D, big, small):
    print(f"[case] irfft_noalloc D={D} {big} then {small}", flush=True)
    for i, (N, count) in enumerate((big, small)):
        xh = _spectra(D, N, count, seed=300 + 10 * D + i)
        ref, got = _irfft_both(D, N, xh, device)
        np.testing.assert_array_equal(got, ref)
        want = np.fft.irfftn(xh.astype(np.complex128), s=(N,) * D, axes=tuple(range(1, D + 1)))
        err = np.linalg.norm(got - want) / np.linalg.norm(want)
        assert err <= 1e-6, err


@pytest.mark.parametrize("D,N", [(2, 512), (3, 128), (3, 256)])
def test_irfft_noalloc_rejects_oversized(device, D, N):
    """One field more than the scratch holds (none at 256^3) fails with
    EXMI_ERR_INVALID and a message naming the limit."""
    import torch

    lib = _lib.lib()
    per = 8 * N ** (D - 1) * (N // 2 + 1)
    count = _lib.EXMI_NOALLOC_SCRATCH_BYTES // per + 1
    out = torch.zeros(1, device=device)
    rc = lib.exmi_irfft_noalloc(D, N, count, out.data_ptr(), out.data_ptr(), _lib.stream_ptr())
    assert rc == _lib.EXMI_ERR_INVALID
    msg = lib.exmi_last_error().decode()
    assert "scratch" in msg and str(_lib.EXMI_NOALLOC_SCRATCH_BYTES) in msg
