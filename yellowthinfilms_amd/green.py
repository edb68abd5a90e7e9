"""Green k-node cache and the uniform-grid Green operator, on the GPU.

Host-side mirror of the reference's interface for the hot path:
  * green_nodes / build_cavity_transfer -- the batched per-node job of
    build_cavity_transfer_1d/_2d (mod_CavityTransfer.jl:196-258, 307-398);
  * uniform_kxky_grid -- mod_UniformSource.jl:89-130;
  * SpectralGreenOperator -- the spectral dyadic g~(kx, ky) on the uniform
    grid (the per-node strength of _strengths_row!, mod_ClosedFormField.jl:
    616-659, without its quadrature weight) and its application to a source
    sheet through the transform pair of mod_UniformSource.jl:229-248.

All device work goes through libytf.so (include/ytf.h); torch only provides
device memory and the stream.
"""
import ctypes as ct
import math

import numpy as np

from . import _lib
from .materials import stack_arrays


def _torch():
    import torch
    return torch


_LAYER_DTYPE = None


def _layer_dtype():
    """numpy record type with ytf_layer's field offsets (from the ctypes layout)."""
    global _LAYER_DTYPE
    if _LAYER_DTYPE is None:
        L = _lib.YtfLayer
        _LAYER_DTYPE = np.dtype({"names": ["eps", "mu", "d", "eps_kind", "mu_kind"],
                                 "formats": [(np.float64, 18), (np.float64, 18), np.float64, np.int32, np.int32],
                                 "offsets": [L.eps.offset, L.mu.offset, L.d.offset, L.eps_kind.offset,
                                             L.mu_kind.offset],
                                 "itemsize": ct.sizeof(L)})
    return _LAYER_DTYPE


def layer_structs(stack):
    """ytf_layer array for a stack (list of FlatOpticalLayer / FlatLayer):
    eps and mu column-major as interleaved (re, im) pairs, built as one numpy
    record array and copied into the ctypes array."""
    eps, mu, d, ek, mk = stack_arrays(stack)
    n = len(d)
    rec = np.zeros(n, dtype=_layer_dtype())
    # column-major 3x3 per layer, then (re, im) interleaved: complex128 viewed as float64 pairs
    rec["eps"] = np.ascontiguousarray(np.transpose(eps, (0, 2, 1)).reshape(n, 9), np.complex128).view(np.float64)
    rec["mu"] = np.ascontiguousarray(np.transpose(mu, (0, 2, 1)).reshape(n, 9), np.complex128).view(np.float64)
    rec["d"] = d
    rec["eps_kind"] = ek
    rec["mu_kind"] = mk
    return (_lib.YtfLayer * n).from_buffer_copy(rec.tobytes())


class NodeCache:
    """Device-resident per-node data of one (source layer, obs layer) pair, as
    the kernels write it: structure of arrays, node index fastest.

      Qsrc  (4, n)   Qf1 Qf2 Qb1 Qb2 of the source layer
      sf,sb (6, n)   2x3 seeds, column-major (mode, moment); sb = -sib_raw
      T     (16, n)  T^o, 4x4 column-major
      basis (16, n)  Exf Eyf Ezf Exb Eyb Ezb KZf KZb (2 modes each) of obs layer
      flags (n,)     uint8: 1 valid, 2 Tvalid, 4 singular source projection, 8 mode failure
    """

    def __init__(self, n, device):
        torch = _torch()
        z = lambda e: torch.zeros((e, n), dtype=torch.complex128, device=device)
        self.Qsrc, self.sf, self.sb, self.T, self.basis = z(4), z(6), z(6), z(16), z(16)
        self.flags = torch.zeros(n, dtype=torch.uint8, device=device)

    def struct(self):
        return _lib.YtfNodeCache(self.Qsrc.data_ptr(), self.sf.data_ptr(), self.sb.data_ptr(), self.T.data_ptr(),
                                 self.basis.data_ptr(), self.flags.data_ptr())

    @property
    def valid(self):
        return (self.flags & _lib.NODE_VALID) != 0

    @property
    def Tvalid(self):
        return (self.flags & _lib.NODE_TVALID) != 0

    def to_numpy(self):
        """Per-node arrays in matrix form: Q (n,4), sf/sb (n,2,3), T (n,4,4), basis (n,8,2)."""
        f = lambda t: t.cpu().numpy()
        n = self.flags.shape[0]
        return dict(Q=f(self.Qsrc).T, sf=f(self.sf).T.reshape(n, 3, 2).transpose(0, 2, 1),
                    sb=f(self.sb).T.reshape(n, 3, 2).transpose(0, 2, 1),
                    T=f(self.T).T.reshape(n, 4, 4).transpose(0, 2, 1), basis=f(self.basis).T.reshape(n, 8, 2),
                    valid=f(self.valid), Tvalid=f(self.Tvalid), flags=f(self.flags))


def green_nodes(stack, omega, kx, ky, source_layer, obs_layer, zsrc=(), zobs=(), cache=True, ghat=True,
                device=0, ghat_out=None, terms="all", dz_box=0.0, ez_local=True):
    """Batched Green k-node solve on the GPU (ytf_green_nodes).

    kx, ky: node wavevectors (torch float64 on the device, or array-likes);
    source_layer, obs_layer: 1-based; zsrc/zobs: depth pairs for the spectral
    dyadic; terms: "all", "direct", "scattered", "toeplitz" or "hankel"
    (closed_form_G's term classes); dz_box > 0: box sources of that thickness
    centred at zsrc (ytf_green_nodes_box), with the local Ez inside the box's
    slab if ez_local.  Returns (NodeCache or None, ghat (npairs, 9, n)
    complex128 or None).
    """
    torch = _torch()
    dev = torch.device("cuda", device)
    kx = torch.as_tensor(kx, dtype=torch.float64, device=dev).contiguous()
    ky = torch.as_tensor(ky, dtype=torch.float64, device=dev).contiguous()
    if kx.shape != ky.shape or kx.dim() != 1:
        raise ValueError("green_nodes: kx and ky must be 1-D of equal length")
    n = kx.shape[0]
    zs = np.ascontiguousarray(np.atleast_1d(np.asarray(zsrc, dtype=np.float64)))
    zo = np.ascontiguousarray(np.atleast_1d(np.asarray(zobs, dtype=np.float64)))
    if zs.shape != zo.shape:
        raise ValueError("green_nodes: zsrc and zobs must pair up")
    npairs = zs.shape[0]
    nc = NodeCache(n, dev) if cache else None
    g = None
    if ghat and npairs:
        g = ghat_out if ghat_out is not None else torch.empty((npairs, 9, n), dtype=torch.complex128, device=dev)
    ctx = _lib.Context.get(device)
    ctx.bind_stream()
    layers = layer_structs(stack)
    cs = nc.struct() if nc is not None else None
    rc = _lib.lib().ytf_green_nodes_box(
        ctx.handle, layers, len(stack), float(omega), ct.c_void_p(kx.data_ptr()), ct.c_void_p(ky.data_ptr()), n,
        int(source_layer), int(obs_layer), npairs, zs.ctypes.data_as(ct.POINTER(ct.c_double)),
        zo.ctypes.data_as(ct.POINTER(ct.c_double)), ct.c_void_p(g.data_ptr()) if g is not None else None,
        ct.byref(cs) if cs is not None else None, _terms(terms), float(dz_box), 1 if ez_local else 0)
    _lib.check(rc, ctx.handle)
    return nc, g


def _terms(terms):
    try:
        return _lib.TERMS[terms]
    except KeyError:
        raise ValueError(f"closed_form_G: terms must be one of {sorted(_lib.TERMS)} (got {terms!r})") from None


def build_cavity_transfer(stack, omega, source_layer, kx, ky, obs_layers=None, device=0):
    """Per-node cache for every observation layer (one launch per obs layer),
    the GPU counterpart of build_cavity_transfer_1d/_2d's node loop
    (mod_CavityTransfer.jl:224-251, 380-392).  Returns {obs_layer: NodeCache}."""
    N = len(stack)
    if not (1 <= source_layer < N):
        raise ValueError(f"build_cavity_transfer: source_layer={source_layer} must be in 1:{N - 1} "
                         "(the last, semi-infinite layer cannot host a source).")
    obs = sorted(set(range(1, N + 1) if obs_layers is None else obs_layers))
    if not all(1 <= o <= N for o in obs):
        raise ValueError(f"build_cavity_transfer: obs_layers out of range 1:{N}")
    return {o: green_nodes(stack, omega, kx, ky, source_layer, o, ghat=False, device=device)[0] for o in obs}


# ---------------------------------------------------------------------------
# Uniform DFT grid -- mod_UniformSource.jl:89-130
# ---------------------------------------------------------------------------
def fft_wavenumbers(N, dk):
    """FFT-native wavenumber order: 0, dk, ..., then the negative half
    (_fft_wavenumbers, mod_UniformSource.jl:127-130)."""
    half = N // 2
    return np.array([dk * (i - 1 if i <= half else i - 1 - N) for i in range(1, N + 1)], dtype=np.float64)


def uniform_kxky_grid(L, dx, Ly=None, dy=None):
    """uniform_kxky_grid(; L, dx, Ly=L, dy=dx) -- mod_UniformSource.jl:89-116."""
    Ly = L if Ly is None else Ly
    dy = dx if dy is None else dy
    L, Ly, dx, dy = float(L), float(Ly), float(dx), float(dy)
    if not (L > 0 and Ly > 0 and dx > 0 and dy > 0):
        raise ValueError("uniform_kxky_grid: L, dx, Ly, dy must all be positive.")
    Nx = math.ceil(2 * L / dx)
    Ny = math.ceil(2 * Ly / dy)
    if Nx < 2 or Ny < 2:
        raise ValueError(f"uniform_kxky_grid: derived grid too small (Nx={Nx}, Ny={Ny})")
    dx = 2 * L / Nx
    dy = 2 * Ly / Ny
    dkx, dky = math.pi / L, math.pi / Ly
    return dict(kx=fft_wavenumbers(Nx, dkx), ky=fft_wavenumbers(Ny, dky), Nx=Nx, Ny=Ny, dkx=dkx, dky=dky,
                kx_max=math.pi / dx, ky_max=math.pi / dy, midx=1 + Nx // 2, midy=1 + Ny // 2, L=L, Ly=Ly, dx=dx,
                dy=dy)


class SpectralGreenOperator:
    """g~(kx, ky; zsrc, zobs) on every node of a uniform grid, and its
    application E = F^-1{ g~ . F{J} } to a source density sheet J.

    Node order is x-major (node = y + Ny*x), the layout the fused y-pass of
    the apply reads.  J and E are the reference's column-major (Nx, Ny, 3)
    complex arrays; as torch tensors that is a contiguous (3, Ny, Nx).
    """

    def __init__(self, stack, omega, grid, source_layer, zsrc, obs_layer, zobs, device=0):
        torch = _torch()
        self.grid = grid
        self.device = device
        nx, ny = grid["Nx"], grid["Ny"]
        dev = torch.device("cuda", device)
        kx = torch.as_tensor(grid["kx"], dtype=torch.float64, device=dev)
        ky = torch.as_tensor(grid["ky"], dtype=torch.float64, device=dev)
        self.kx_nodes = kx.repeat_interleave(ny)  # node = y + ny*x
        self.ky_nodes = ky.repeat(nx)
        self.stack, self.omega = stack, omega
        self.source_layer, self.obs_layer = source_layer, obs_layer
        self.zsrc, self.zobs = float(zsrc), float(zobs)
        self.ghat = None
        self.cache = None
        self.work = torch.empty((3, nx, ny), dtype=torch.complex128, device=dev)
        # _forward_transform: dx dy ramp fft;  _inverse_transform: dkx dky nx ny/(2pi)^2 ifft(. conj(ramp))
        self.scale = (grid["dx"] * grid["dy"]) * (grid["dkx"] * grid["dky"] * nx * ny / (2 * math.pi) ** 2) / (nx * ny)

    def build(self, with_cache=False):
        """Solve every grid node (one Green k-node each) and store g~."""
        self.cache, g = green_nodes(self.stack, self.omega, self.kx_nodes, self.ky_nodes, self.source_layer,
                                    self.obs_layer, [self.zsrc], [self.zobs], cache=with_cache, ghat=True,
                                    device=self.device, ghat_out=self.ghat)
        self.ghat = g
        return self

    def apply(self, J, out=None):
        torch = _torch()
        nx, ny = self.grid["Nx"], self.grid["Ny"]
        if self.ghat is None:
            raise RuntimeError("SpectralGreenOperator.apply: call build() first")
        J = torch.as_tensor(J, dtype=torch.complex128, device=self.ghat.device)
        if tuple(J.shape) != (3, ny, nx):
            raise ValueError(f"apply: J must be (3, Ny, Nx) = {(3, ny, nx)} (column-major (Nx, Ny, 3)), "
                             f"got {tuple(J.shape)}")
        J = J.contiguous()
        E = out if out is not None else torch.empty_like(J)
        ctx = _lib.Context.get(self.device)
        ctx.bind_stream()
        rc = _lib.lib().ytf_spectral_apply(ctx.handle, nx, ny, ct.c_void_p(self.ghat.data_ptr()),
                                           ct.c_void_p(J.data_ptr()), ct.c_void_p(E.data_ptr()),
                                           ct.c_void_p(self.work.data_ptr()), self.scale)
        _lib.check(rc, ctx.handle)
        return E
