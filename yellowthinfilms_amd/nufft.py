"""2D type-1 nonuniform FFT on the GPU (ytf_nufft2d1), host side.

The transform of the 2D lateral sum of closed_form_G with lateral=:nufft
(FINUFFT type 1, iflag +1; mod_ClosedFormField.jl:758-780):
    f[t, k2, k1] = sum_j c[t, j] exp(+i (k1 x_j + k2 y_j)),
    k1 = -floor(N1/2) .. N1 - floor(N1/2) - 1, likewise k2,
i.e. Julia's F (N1, N2, ntrans) column-major = a (ntrans, N2, N1) tensor.
"""
import ctypes as ct

from . import _lib


def _torch():
    import torch
    return torch


def nufft2d1(x, y, c, N1, N2, eps=1e-9, device=0, out=None):
    """x, y: (M,) real; c: (M,) or (ntrans, M) complex.  Returns (ntrans, N2, N1)
    complex128 on the device (ntrans = 1 for 1-D c)."""
    torch = _torch()
    dev = torch.device("cuda", device)
    x = torch.as_tensor(x, dtype=torch.float64, device=dev).contiguous()
    y = torch.as_tensor(y, dtype=torch.float64, device=dev).contiguous()
    c = torch.as_tensor(c, dtype=torch.complex128, device=dev)
    if c.dim() == 1:
        c = c[None]
    c = c.contiguous()
    M = x.shape[0]
    if y.shape != (M,) or c.shape[1] != M:
        raise ValueError("nufft2d1: x, y and the strengths must have one entry per point")
    nt = c.shape[0]
    f = out if out is not None else torch.empty((nt, N2, N1), dtype=torch.complex128, device=dev)
    ctx = _lib.Context.get(device)
    ctx.bind_stream()
    rc = _lib.lib().ytf_nufft2d1(ctx.handle, M, ct.c_void_p(x.data_ptr()), ct.c_void_p(y.data_ptr()), nt,
                                 ct.c_void_p(c.data_ptr()), int(N1), int(N2), float(eps), ct.c_void_p(f.data_ptr()))
    _lib.check(rc, ctx.handle)
    return f
