"""ytf_spectral_apply_noalloc (no counterpart in the reference: the spectral
apply for callers that cannot allocate) against ytf_spectral_apply, through the
C-ABI: bit-identical results on two calls in a row, the second on a smaller
grid (the shared device array keeps the first call's contents), the smaller one
also against the oracle's FFT route; grids beyond its fixed array are refused."""
import ctypes as ct

import numpy as np
import pytest

from helpers import relerr

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


def _c(t):
    return ct.c_void_p(t.data_ptr())


def _inputs(nx, ny, seed):
    g = torch.Generator(device="cpu").manual_seed(seed)
    ghat = torch.randn((9, nx * ny), dtype=torch.complex128, generator=g)
    J = torch.randn((3, ny, nx), dtype=torch.complex128, generator=g)  # (nx, ny, 3) column-major
    return ghat.cuda(), J.cuda()


def test_noalloc_spectral_apply_matches_the_allocating_entry_point():
    from yellowthinfilms_amd import _lib
    L = _lib.lib()
    ctx = _lib.Context.get(0)
    ctx.bind_stream()
    for nx, ny, seed in [(512, 512, 1), (128, 64, 2)]:
        print(f"case: noalloc spectral apply {nx}x{ny}", flush=True)
        ghat, J = _inputs(nx, ny, seed)
        E1 = torch.empty_like(J)
        E2 = torch.full_like(J, complex("nan"))
        work = torch.empty((3 * nx * ny,), dtype=torch.complex128, device="cuda")
        scale = 0.37
        _lib.check(L.ytf_spectral_apply(ctx.handle, nx, ny, _c(ghat), _c(J), _c(E1), _c(work), scale), ctx.handle)
        _lib.check(L.ytf_spectral_apply_noalloc(ctx.handle, nx, ny, _c(ghat), _c(J), _c(E2), scale), ctx.handle)
        torch.cuda.synchronize()
        assert torch.equal(E1, E2), float((E1 - E2).abs().max())
    # the smaller grid against the definition E_c = scale bfft2(sum_q g_cq fft2(J_q))
    gn = ghat.cpu().numpy().reshape(3, 3, nx, ny)  # [q][c][x][y]: entry c + 3q, node y + ny x
    Jn = J.cpu().numpy().transpose(0, 2, 1)       # [q][x][y]
    S = np.einsum("qcxy,qxy->cxy", gn, np.fft.fft2(Jn, axes=(1, 2)))
    ref = scale * nx * ny * np.fft.ifft2(S, axes=(1, 2))
    got = E2.cpu().numpy().transpose(0, 2, 1)
    assert relerr(got, ref) < 1e-12, relerr(got, ref)


def test_noalloc_spectral_apply_refuses_what_its_array_cannot_hold():
    from yellowthinfilms_amd import _lib
    L = _lib.lib()
    ctx = _lib.Context.get(0)
    ctx.bind_stream()
    ghat, J = _inputs(8, 8, 3)
    E = torch.empty_like(J)
    for nx, ny, why in [(1024, 512, b"exceeds"), (96, 64, b"unsupported"), (4, 8, b"unsupported")]:
        rc = L.ytf_spectral_apply_noalloc(ctx.handle, nx, ny, _c(ghat), _c(J), _c(E), 1.0)
        assert rc == _lib.YTF_EINVAL and why in L.ytf_last_error(ctx.handle), (nx, ny)
    assert L.ytf_spectral_apply_noalloc(ctx.handle, 8, 8, _c(ghat), _c(J), _c(J), 1.0) == _lib.YTF_EINVAL
