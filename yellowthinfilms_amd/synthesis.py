"""1D (azimuthally symmetric) closed-form Green's function on the GPU.

Host-side mirror of the reference's 1D route:
  * CavityTransfer1D -- the node set (kx_k, ky = 0) and quadrature weights of
    a CavityTransferCache1D (mod_CavityTransfer.jl:160-258) for one stack and
    source layer; the per-node solve runs in ytf_green_nodes;
  * closed_form_G -- closed_form_G(cache, obs_layer, zsrc_local, z_obs, xs,
    ys) (mod_ClosedFormField.jl:435-487): per z row, the node's spectral
    dyadic, then the J0/J1/J2 contraction at the distinct pixel radii and the
    rotation to each pixel (ytf_closed_form_G_1d);
  * radial_synthesis -- the contraction alone at given radii
    (ytf_radial_synthesis), the response R(rho) of a pixel on the local x axis.
"""
import ctypes as ct

import numpy as np

from . import _lib
from .green import green_nodes


def _torch():
    import torch
    return torch


def _dptr(t):
    return ct.c_void_p(t.data_ptr())


def _mode(mode):
    try:
        return _lib.SYNTH_MODES[mode]
    except KeyError:
        raise ValueError(f"unknown synthesis precision {mode!r}; one of {sorted(_lib.SYNTH_MODES)}") from None


class CavityTransfer1D:
    """Nodes kx_k (ky = 0) with Sommerfeld weights wt_k for one stack and
    source layer; the spectral dyadics are solved per (obs layer, z rows)."""

    def __init__(self, stack, omega, kx_nodes, weights, source_layer, device=0):
        torch = _torch()
        N = len(stack)
        if not (1 <= source_layer < N):
            raise ValueError(f"CavityTransfer1D: source_layer={source_layer} must be in 1:{N - 1} "
                             "(the last, semi-infinite layer cannot host a source).")
        dev = torch.device("cuda", device)
        self.stack, self.omega, self.source_layer, self.device = stack, float(omega), int(source_layer), device
        self.kx = torch.as_tensor(np.asarray(kx_nodes, np.float64), device=dev).contiguous()
        self.wt = torch.as_tensor(np.asarray(weights, np.float64), device=dev).contiguous()
        if self.kx.shape != self.wt.shape or self.kx.dim() != 1:
            raise ValueError("CavityTransfer1D: kx_nodes and weights must be 1-D of equal length")
        self.ky = torch.zeros_like(self.kx)

    @property
    def n_nodes(self):
        return self.kx.shape[0]

    def ghat(self, obs_layer, zsrc_local, z_obs, dz_box=0.0):
        """g~(kx_k, 0; zsrc, z_obs[i]) for every node: (nz, 9, n) complex128.
        dz_box > 0: a box source of that thickness.  The 1D path has no
        interior-slab term: an observation plane inside the box's own slab
        raises, as direct_amplitude's branch does (mod_ClosedFormField.jl:208-214)."""
        z_obs = np.atleast_1d(np.asarray(z_obs, np.float64))
        if dz_box > 0 and obs_layer == self.source_layer:
            d = self.stack[self.source_layer - 1].d
            zpos = d - float(zsrc_local) if self.source_layer == 1 else float(zsrc_local)
            for z in z_obs:
                zeta = z - zpos
                if abs(zeta) < dz_box / 2:
                    raise ValueError(f"direct_amplitude: z_obs is inside the box's own z-slab (|ζ|={abs(zeta)} < "
                                     f"Δz/2={dz_box / 2}) -- that needs the interior-slab term, not the exterior "
                                     "seed.")
        _, g = green_nodes(self.stack, self.omega, self.kx, self.ky, self.source_layer, obs_layer,
                           np.full(len(z_obs), float(zsrc_local)), z_obs, cache=False, ghat=True,
                           device=self.device, dz_box=dz_box)
        return g


def radial_synthesis(cache, ghat, rho, mode="fp64"):
    """R[p, c + 3q, r] of a pixel at (rho_r, 0) for every depth pair p of
    `ghat` ((np, 9, n) from CavityTransfer1D.ghat).  Returns (np, 9, nr)."""
    torch = _torch()
    rho = torch.as_tensor(np.asarray(rho, np.float64), device=cache.kx.device).contiguous()
    npairs = ghat.shape[0]
    R = torch.empty((npairs, 9, rho.shape[0]), dtype=torch.complex128, device=cache.kx.device)
    ctx = _lib.Context.get(cache.device)
    ctx.bind_stream()
    rc = _lib.lib().ytf_radial_synthesis(ctx.handle, cache.n_nodes, _dptr(cache.kx), _dptr(cache.wt), npairs,
                                         _dptr(ghat), rho.shape[0], _dptr(rho), _dptr(R), _mode(mode))
    _lib.check(rc, ctx.handle)
    return R


def closed_form_G(cache, obs_layer, zsrc_local, z_obs, xs, ys, mode="fp64", dz_box=0.0):
    """G(x, y, z, c, q) of a unit moment q at (0, 0, zsrc_local) in the source
    layer, at pixels (xs[i], ys[j]) and depths z_obs in obs_layer.

    Returns a torch complex128 tensor of shape (3, 3, nz, ny, nx), i.e. the
    reference's column-major (nx, ny, nz, 3, 3) array; `.permute(4, 3, 2, 1, 0)`
    gives (x, y, z, c, q) indexing."""
    torch = _torch()
    xs = np.ascontiguousarray(np.atleast_1d(np.asarray(xs, np.float64)))
    ys = np.ascontiguousarray(np.atleast_1d(np.asarray(ys, np.float64)))
    g = cache.ghat(obs_layer, zsrc_local, z_obs, dz_box=dz_box)
    nz = g.shape[0]
    G = torch.empty((3, 3, nz, len(ys), len(xs)), dtype=torch.complex128, device=cache.kx.device)
    ctx = _lib.Context.get(cache.device)
    ctx.bind_stream()
    dp = ct.POINTER(ct.c_double)
    rc = _lib.lib().ytf_closed_form_G_1d(ctx.handle, cache.n_nodes, _dptr(cache.kx), _dptr(cache.wt), nz, _dptr(g),
                                         xs.ctypes.data_as(dp), len(xs), ys.ctypes.data_as(dp), len(ys), _dptr(G),
                                         _mode(mode))
    _lib.check(rc, ctx.handle)
    return G
