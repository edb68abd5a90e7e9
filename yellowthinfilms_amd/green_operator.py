"""2D closed-form Green's function and the padded, z-structured Green
operator on the GPU.

Host-side mirror of the reference's 2D route:
  * CavityTransfer2D -- the (k_rho, phi) nodes and weights of a
    CavityTransferCache2D (build_cavity_transfer_2d, mod_CavityTransfer.jl:
    307-398); the per-node solve runs inside every call (ytf_green_nodes);
  * closed_form_G_2d -- closed_form_G(cache::CavityTransferCache2D, ...)
    (mod_ClosedFormField.jl:711-782), lateral=:direct or :nufft;
  * grid_offsets, padded_offsets, Sheet, point_current_sheet,
    build_closed_form_green, apply_ghat, apply_closed_form,
    precompute_ghats, closed_form_field -- the multi-source lateral assembly
    of mod_ClosedFormAssembly.jl (G on the 2Nx x 2Ny padded offsets, FFT'd
    once, exact linear convolution);
  * Dipole, z_skeletons, cache_skeletons, z_skeletons_grouped, build_layer_operator,
    build_layer_operator_grouped, apply_operator, dipole_moments,
    operator_summary -- mod_GreenOperator.jl.

Array conventions: the reference's column-major Julia arrays are contiguous
torch tensors with the axes reversed: P (Nx, Ny, Ns, 3) is a (3, Ns, Ny, Nx)
tensor, E (Nx, Ny, No, 3) a (3, No, Ny, Nx) tensor, G (Nx, Ny, Nz, 3, 3) a
(3, 3, Nz, Ny, Nx) tensor.  All device work goes through libytf.so.
"""
import ctypes as ct
import math
import sys
import time
import warnings
from dataclasses import dataclass

import numpy as np

from . import _lib
from .green import green_nodes, layer_structs, _terms


def _torch():
    import torch
    return torch


def _dptr(t):
    return ct.c_void_p(t.data_ptr())


_DP = ct.POINTER(ct.c_double)


def _lateral(lateral):
    try:
        return _lib.LATERAL[lateral]
    except KeyError:
        raise ValueError("closed_form_G: lateral must be :direct or :nufft") from None


def sinc_phys(u):
    """sin(u)/u with sinc_phys(0) = 1 (mod_ClosedFormField.jl:515)."""
    return np.sinc(np.asarray(u, dtype=np.float64) / np.pi)


def box_lateral_factor(kx, ky, dx_box, dy_box):
    """The box footprint's 2D Fourier transform, S(0, 0) = 1
    (mod_ClosedFormField.jl:517-524)."""
    return sinc_phys(np.asarray(kx) * dx_box / 2) * sinc_phys(np.asarray(ky) * dy_box / 2)


class CavityTransfer2D:
    """The (k_rho, phi) nodes of a CavityTransferCache2D for one stack and
    source layer: kx = k_rho cos(phi), ky = k_rho sin(phi), flattened with
    k_rho fastest (_node_kxky, mod_ClosedFormField.jl:526-528), and the node
    weights W = k_rho w_rho w_phi / 4 pi^2 (:723-724).

    rotation ("auto", True, False; build_cavity_transfer_2d, mod_CavityTransfer.jl:296-306):
    on an azimuthally symmetric stack every k_rho is solved once, at phi = 0,
    and the builds turn it to every phi (mod_CavityTransfer.jl:361-378); "auto"
    takes it iff the stack is symmetric, True on another stack is an error,
    False solves every node."""

    def __init__(self, stack, omega, source_layer, krho_nodes, krho_weights, phi_nodes, phi_weights,
                 obs_layers=None, device=0, rotation="auto"):
        from .kspace import _is_azimuthally_symmetric
        kr = np.asarray(krho_nodes, np.float64)
        wr = np.asarray(krho_weights, np.float64)
        ph = np.asarray(phi_nodes, np.float64)
        wp = np.asarray(phi_weights, np.float64)
        if kr.shape != wr.shape or ph.shape != wp.shape:
            raise ValueError("build_cavity_transfer_2d: kρ nodes/weights length mismatch" if kr.shape != wr.shape
                             else "build_cavity_transfer_2d: φ nodes/weights length mismatch")
        sym = _is_azimuthally_symmetric(stack)
        use_rot = sym if (isinstance(rotation, str) and rotation == "auto") else bool(rotation)
        if use_rot and not sym:
            raise ValueError("build_cavity_transfer_2d: rotation=true requires an azimuthally symmetric stack")
        K, Ph = np.meshgrid(kr, ph, indexing="ij")
        kx = (K * np.cos(Ph)).ravel(order="F")
        ky = (K * np.sin(Ph)).ravel(order="F")
        W = (np.outer(kr * wr, wp) / (4 * np.pi ** 2)).ravel(order="F")
        self._init(stack, omega, source_layer, kx, ky, W, obs_layers, device)
        self.krho_nodes, self.krho_weights, self.phi_nodes, self.phi_weights = kr, wr, ph, wp
        self.rotation = bool(use_rot and len(kr) > 0 and len(ph) > 0)

    @classmethod
    def from_nodes(cls, stack, omega, source_layer, kx, ky, W, obs_layers=None, device=0):
        """Any node set with its (complex) weights, e.g. the concatenated
        nodes of several (k_rho, phi) groups."""
        self = cls.__new__(cls)
        self._init(stack, omega, source_layer, np.asarray(kx, np.float64), np.asarray(ky, np.float64),
                   np.asarray(W), obs_layers, device)
        self.rotation = False
        return self

    def rotation_fields(self):
        """(rot_n_phi, rot_n_krho, rot_krho, rot_phi) of ytf_operator_options."""
        if not getattr(self, "rotation", False):
            return 0, 0, None, None
        kr = np.ascontiguousarray(self.krho_nodes)
        ph = np.ascontiguousarray(self.phi_nodes)
        return len(ph), len(kr), kr.ctypes.data_as(_DP), ph.ctypes.data_as(_DP)

    def _init(self, stack, omega, source_layer, kx, ky, W, obs_layers, device):
        torch = _torch()
        N = len(stack)
        if not (1 <= source_layer < N):
            raise ValueError(f"build_cavity_transfer_2d: source_layer={source_layer} must be in 1:{N - 1} "
                             "(the last, semi-infinite layer cannot host a source).")
        if kx.shape != ky.shape or kx.shape != W.shape or kx.ndim != 1:
            raise ValueError("CavityTransfer2D: kx, ky and W must be 1-D of equal length")
        dev = torch.device("cuda", device)
        self.stack, self.omega, self.source_layer, self.device = stack, float(omega), int(source_layer), device
        self.obs_layers = sorted(set(range(1, N + 1) if obs_layers is None else obs_layers))
        self.kx_host, self.ky_host, self.W_host = kx, ky, W.astype(np.complex128)
        self.kx = torch.as_tensor(kx, device=dev).contiguous()
        self.ky = torch.as_tensor(ky, device=dev).contiguous()
        self.W = torch.as_tensor(self.W_host, device=dev).contiguous()

    @property
    def n_nodes(self):
        return self.kx.shape[0]

    def weights(self, box_lateral=None):
        """Node weights, times box_lateral_factor for a box footprint."""
        if box_lateral is None:
            return self.W
        torch = _torch()
        S = box_lateral_factor(self.kx_host, self.ky_host, float(box_lateral[0]), float(box_lateral[1]))
        return torch.as_tensor(self.W_host * S, device=self.W.device).contiguous()

    def check_obs(self, obs_layer, who="amplitude_at"):
        if obs_layer not in self.obs_layers:
            raise ValueError(f"{who}: obs_layer={obs_layer} not in this cache's obs_layers {self.obs_layers}")


def closed_form_G_2d(cache, obs_layer, zsrc_local, z_obs, xs, ys, box_lateral=None, dz_box=0.0, ez_local=True,
                     lateral="direct", nufft_eps=1e-9, terms="all"):
    """G(x, y, z, c, q) of a unit moment q at (0, 0, zsrc_local) in the source
    layer, at pixels (xs[i], ys[j]) and depths z_obs in obs_layer
    (closed_form_G for a CavityTransferCache2D, mod_ClosedFormField.jl:711-782).
    A box source is box_lateral = (dx, dy) (footprint factor on the node
    weights) and dz_box (the axial factor, the interior slab and, if
    ez_local, the local Ez).  Returns a (3, 3, nz, ny, nx) complex128 tensor
    (the column-major (nx, ny, nz, 3, 3) array)."""
    torch = _torch()
    cache.check_obs(obs_layer)
    lat = _lateral(lateral)
    xs = np.ascontiguousarray(np.atleast_1d(np.asarray(xs, np.float64)))
    ys = np.ascontiguousarray(np.atleast_1d(np.asarray(ys, np.float64)))
    z_obs = np.atleast_1d(np.asarray(z_obs, np.float64))
    _, g = green_nodes(cache.stack, cache.omega, cache.kx, cache.ky, cache.source_layer, obs_layer,
                       np.full(len(z_obs), float(zsrc_local)), z_obs, cache=False, ghat=True, device=cache.device,
                       terms=terms, dz_box=dz_box, ez_local=ez_local)
    W = cache.weights(box_lateral)
    G = torch.empty((3, 3, len(z_obs), len(ys), len(xs)), dtype=torch.complex128, device=cache.kx.device)
    ctx = _lib.Context.get(cache.device)
    ctx.bind_stream()
    rc = _lib.lib().ytf_closed_form_G_2d(ctx.handle, cache.n_nodes, _dptr(cache.kx), _dptr(cache.ky), _dptr(W),
                                         len(z_obs), _dptr(g), xs.ctypes.data_as(_DP), len(xs),
                                         ys.ctypes.data_as(_DP), len(ys), lat, float(nufft_eps), _dptr(G))
    _lib.check(rc, ctx.handle)
    return G


# ---------------------------------------------------------------------------
# Sources -- mod_Sources.jl, mod_UniformSource.jl
# ---------------------------------------------------------------------------
@dataclass
class Dipole:
    """A point dipole of moment p at (x, y) and local depth z of `layer`
    (mod_Sources.jl:25-35)."""
    p: tuple
    omega: float
    x: float
    z: float
    layer: int
    y: float = 0.0


def _moment_vec(p):
    if isinstance(p, str):
        return {"x": np.array([1, 0, 0], complex), "y": np.array([0, 1, 0], complex),
                "z": np.array([0, 0, 1], complex)}[p]
    return np.asarray(p, dtype=np.complex128).reshape(3)


def nearest_node_index(grid, x, y):
    """1-based pixel (i, j) of (x, y), clamped into the grid with a warning
    (_nearest_node_index, mod_UniformSource.jl:169-178; round half to even)."""
    i = int(np.rint(x / grid["dx"])) + grid["midx"]
    j = int(np.rint(y / grid["dy"])) + grid["midy"]
    if not (1 <= i <= grid["Nx"]) or not (1 <= j <= grid["Ny"]):
        warnings.warn(f"point_current_sheet: position (x={x}, y={y}) snaps to pixel ({i},{j}), outside this grid's "
                      f"Nx={grid['Nx']}/Ny={grid['Ny']} -- clamping into range. Enlarge L/Ly if this point should be "
                      "included.")
        i = min(max(i, 1), grid["Nx"])
        j = min(max(j, 1), grid["Ny"])
    return i, j


@dataclass
class Sheet:
    """A current-density sheet J (moment per unit area) at local depth z
    (mod_Sources.jl Sheet): J is a (3, Ny, Nx) array (column-major (Nx, Ny, 3))."""
    J: np.ndarray
    z: float
    weight: float = 1.0


def point_current_sheet(grid, z, points, weight=1.0):
    """Sheet of point moments (x, y, p) snapped to pixels, each written as
    p / (dx dy) (point_current_sheet, mod_UniformSource.jl:212-219)."""
    J = np.zeros((3, grid["Ny"], grid["Nx"]), np.complex128)
    for (x, y, p) in points:
        i, j = nearest_node_index(grid, x, y)
        J[:, j - 1, i - 1] += _moment_vec(p) / (grid["dx"] * grid["dy"])
    return Sheet(J, float(z), float(weight))


def grid_offsets(grid):
    """Pixel offsets x_i = (i - midx) dx, y_j = (j - midy) dy
    (mod_ClosedFormAssembly.jl:60-62)."""
    return (np.array([(i - grid["midx"]) * grid["dx"] for i in range(1, grid["Nx"] + 1)]),
            np.array([(j - grid["midy"]) * grid["dy"] for j in range(1, grid["Ny"] + 1)]))


def padded_offsets(grid):
    """The 2Nx x 2Ny centred lateral offsets G is sampled on, xs[Nx] = 0
    (0-based; mod_ClosedFormAssembly.jl:75-79)."""
    Px, Py = 2 * grid["Nx"], 2 * grid["Ny"]
    return (np.array([(i - Px // 2) * grid["dx"] for i in range(Px)]),
            np.array([(j - Py // 2) * grid["dy"] for j in range(Py)]))


# ---------------------------------------------------------------------------
# The operator -- mod_GreenOperator.jl
# ---------------------------------------------------------------------------
class LayerGreenOperator:
    """Precomputed operator of one source layer (LayerGreenOperator,
    mod_GreenOperator.jl:105-113): the device handle plus its description.
    `blocks` lists one dict per observation layer (ascending), as
    ytf_operator_block_info reports it."""

    def __init__(self, handle, cache, grid, z_src, kwargs, device):
        self.handle = handle
        self.source_layer = cache.source_layer
        self.grid = grid
        self.z_src = np.asarray(z_src, np.float64)
        self.flip = cache.source_layer == 1
        self.kwargs = kwargs
        self.device = device
        self.skeletons = None
        self.refresh()

    def refresh(self):
        """Re-read the block descriptions (after ytf_operator_accumulate the
        build times are the sums over the groups)."""
        L = _lib.lib()
        ctx = _lib.Context.get(self.device)
        handle = self.handle
        self.blocks = []
        for b in range(L.ytf_operator_num_blocks(handle)):
            info = _lib.YtfBlockInfo()
            _lib.check(L.ytf_operator_block_info(ctx.handle, handle, b, ct.byref(info)), ctx.handle)
            self.blocks.append(dict(obs_layer=info.obs_layer, kind=_lib.BLOCK_KINDS[info.kind], n_obs=info.n_obs,
                                    n_src_skel=info.n_src_skel, n_obs_skel=info.n_obs_skel,
                                    n_toeplitz=info.n_toeplitz, n_hankel=info.n_hankel, n_dense=info.n_dense,
                                    bytes=info.bytes, build_s=info.build_ms / 1e3,
                                    check_residual=None if info.check_residual < 0 else info.check_residual))
        self.timings = {b["obs_layer"]: b["build_s"] for b in self.blocks}

    def __del__(self):
        h = getattr(self, "handle", None)
        if h:
            try:
                _lib.lib().ytf_operator_destroy(_lib.Context.get(self.device).handle, h)
            except Exception:
                pass
            self.handle = None

    def skeleton(self, block, side):
        """(idx, P, err) of one side ("source" or "observation") of a block:
        skeleton planes (0-based), the (N, k) interpolation matrix and its
        largest error over the node terms (ZSkeleton, mod_GreenOperator.jl:202-208)."""
        L = _lib.lib()
        ctx = _lib.Context.get(self.device)
        sd = {"source": 0, "observation": 1}[side]
        k = ct.c_int32()
        _lib.check(L.ytf_operator_skeleton(ctx.handle, self.handle, block, sd, ct.byref(k), None, None, None),
                   ctx.handle)
        N = len(self.z_src) if sd == 0 else self.blocks[block]["n_obs"]
        idx = (ct.c_int32 * max(k.value, 1))()
        P = np.zeros(2 * N * k.value, np.float64)
        err = ct.c_double()
        _lib.check(L.ytf_operator_skeleton(ctx.handle, self.handle, block, sd, ct.byref(k), idx,
                                           P.ctypes.data_as(_DP), ct.byref(err)), ctx.handle)
        return np.array(idx[:k.value]), P.view(np.complex128).reshape((N, k.value), order="F"), err.value

    def ghat_plane(self, block, part, slot):
        """One stored Ĝ plane (1/(Px Py) folded in) as a (3, 3, Py, Px) tensor
        (the reference's (Px, Py, 3, 3)); part "toeplitz", "hankel" or "dense"."""
        torch = _torch()
        g = self.grid
        out = torch.empty((3, 3, 2 * g["Ny"], 2 * g["Nx"]), dtype=torch.complex128,
                          device=torch.device("cuda", self.device))
        ctx = _lib.Context.get(self.device)
        ctx.bind_stream()
        p = {"toeplitz": 0, "hankel": 1, "dense": 2}[part]
        _lib.check(_lib.lib().ytf_operator_ghat(ctx.handle, self.handle, block, p, slot, _dptr(out)), ctx.handle)
        return out


def _grid_struct(grid):
    return _lib.YtfGrid(int(grid["Nx"]), int(grid["Ny"]), float(grid["dx"]), float(grid["dy"]))


def _obs_arrays(z_obs_by_layer):
    obs = list(z_obs_by_layer)
    n_obs = (ct.c_int32 * max(len(obs), 1))(*[len(np.atleast_1d(z_obs_by_layer[o])) for o in obs])
    layers_c = (ct.c_int32 * max(len(obs), 1))(*obs)
    zo = np.ascontiguousarray(np.concatenate([np.atleast_1d(np.asarray(z_obs_by_layer[o], np.float64))
                                              for o in obs]) if obs else np.zeros(1))
    return obs, n_obs, layers_c, zo


class ZSkeletons:
    """The result of z_skeletons (mod_GreenOperator.jl:335-359): `src` and
    `obs[o]` are ZSkeleton records (idx 0-based and increasing, P the (N, k)
    interpolation matrix, err, pnorm, whole); the device handle goes to
    build_layer_operator(..., skeletons=).  Source planes are in increasing
    cavity-frame depth, as the operator stores them."""

    def __init__(self, handle, source_layer, obs_layers, device):
        self.handle, self.source_layer, self.device = handle, source_layer, device
        self.src = self._side(0)
        self.obs = {o: self._side(o) for o in obs_layers}

    def _side(self, layer):
        L = _lib.lib()
        ctx = _lib.Context.get(self.device)
        N, k = ct.c_int32(), ct.c_int32()
        _lib.check(L.ytf_skeletons_get(ctx.handle, self.handle, layer, ct.byref(N), ct.byref(k), None, None, None),
                   ctx.handle)
        idx = (ct.c_int32 * max(k.value, 1))()
        P = np.zeros(2 * N.value * k.value, np.float64)
        err = ct.c_double()
        _lib.check(L.ytf_skeletons_get(ctx.handle, self.handle, layer, ct.byref(N), ct.byref(k), idx,
                                       P.ctypes.data_as(_DP), ct.byref(err)), ctx.handle)
        P = P.view(np.complex128).reshape((N.value, k.value), order="F")
        whole = k.value == N.value
        return ZSkeleton(np.array(idx[:k.value]), P, err.value,
                         float(np.linalg.norm(P, 2)) if P.size else 1.0, whole)

    def __del__(self):
        h = getattr(self, "handle", None)
        if h:
            try:
                _lib.lib().ytf_skeletons_destroy(_lib.Context.get(self.device).handle, h)
            except Exception:
                pass
            self.handle = None


@dataclass
class ZSkeleton:
    """ZSkeleton (mod_GreenOperator.jl:202-208)."""
    idx: np.ndarray
    P: np.ndarray
    err: float
    pnorm: float
    whole: bool


def _z_skeletons(stack, omega, source_layer, kx, ky, z_src, z_obs_by_layer, tol, device):
    torch = _torch()
    if not tol > 0:
        raise ValueError("z_skeletons: tol must be > 0")
    dev = torch.device("cuda", device)
    kx = torch.as_tensor(np.asarray(kx, np.float64), device=dev).contiguous()
    ky = torch.as_tensor(np.asarray(ky, np.float64), device=dev).contiguous()
    zs = np.ascontiguousarray(np.atleast_1d(np.asarray(z_src, np.float64)))
    obs, n_obs, layers_c, zo = _obs_arrays(z_obs_by_layer)
    ctx = _lib.Context.get(device)
    ctx.bind_stream()
    h = ct.c_void_p()
    _lib.check(_lib.lib().ytf_z_skeletons(ctx.handle, layer_structs(stack), len(stack), float(omega), source_layer,
                                          kx.shape[0], _dptr(kx), _dptr(ky), zs.ctypes.data_as(_DP), len(zs),
                                          len(obs), layers_c, n_obs, zo.ctypes.data_as(_DP), float(tol),
                                          ct.byref(h)), ctx.handle)
    return ZSkeletons(h, source_layer, [o for o in sorted(obs) if len(np.atleast_1d(z_obs_by_layer[o]))], device)


def cache_skeletons(cache, z_src, z_obs_by_layer, tol=1e-6):
    """The z skeletons of a cache's source side and of every observation
    layer of z_obs_by_layer, from the mode wavenumbers at the cache's nodes
    (_cache_skeletons, mod_GreenOperator.jl:361-363): what build_layer_operator
    builds for itself.  z_src in the caller's order (Dipole.z).  Under the
    rotation fill the wavenumbers do not depend on phi, and each k_rho counts
    once (at phi = 0), as in the build."""
    if getattr(cache, "rotation", False):
        kx, ky = cache.krho_nodes, np.zeros_like(cache.krho_nodes)
    else:
        kx, ky = cache.kx_host, cache.ky_host
    return _z_skeletons(cache.stack, cache.omega, cache.source_layer, kx, ky, z_src, z_obs_by_layer, tol,
                        cache.device)


def _q4(Q):
    """(Qf, Qb) as 2 x n arrays (or n-vectors of 2-vectors) -> (4, n) complex."""
    Qf, Qb = (np.asarray(q, np.complex128) for q in Q)
    Qf = Qf if Qf.ndim == 2 and Qf.shape[0] == 2 else Qf.reshape(-1, 2).T
    Qb = Qb if Qb.ndim == 2 and Qb.shape[0] == 2 else Qb.reshape(-1, 2).T
    return np.ascontiguousarray(np.concatenate([Qf, Qb]))


def z_skeletons(Qsrc, Qobs, stack, s, zs_int, zobs, layers, tol, device=0):
    """z_skeletons(Qsrc, Qobs, stack, s, zs_int, zobs, layers; tol)
    (mod_GreenOperator.jl:335-345): Qsrc = (Qf, Qb) of source layer s and
    Qobs[o] = (KZf, KZb) for o in layers, on every quadrature node; zs_int the
    source planes in increasing cavity-frame depth.  Returns a ZSkeletons."""
    torch = _torch()
    dev = torch.device("cuda", device)
    q_src = torch.as_tensor(_q4(Qsrc), device=dev).contiguous()
    n = q_src.shape[1]
    layers = [int(o) for o in layers]
    q_obs = [torch.as_tensor(_q4(Qobs[o]), device=dev).contiguous() for o in layers]
    if any(q.shape[1] != n for q in q_obs):
        raise ValueError("z_skeletons: every Q list must hold the same nodes")
    zs = np.ascontiguousarray(np.atleast_1d(np.asarray(zs_int, np.float64)))
    _, n_obs, layers_c, zo = _obs_arrays({o: zobs[o] for o in layers})
    qp = (ct.c_void_p * max(len(q_obs), 1))(*[q.data_ptr() for q in q_obs])
    ctx = _lib.Context.get(device)
    ctx.bind_stream()
    h = ct.c_void_p()
    _lib.check(_lib.lib().ytf_z_skeletons_q(ctx.handle, layer_structs(stack), len(stack), int(s), n, _dptr(q_src),
                                            zs.ctypes.data_as(_DP), len(zs), len(layers), layers_c, qp, n_obs,
                                            zo.ctypes.data_as(_DP), float(tol), ct.byref(h)), ctx.handle)
    return ZSkeletons(h, int(s), sorted(layers), device)


def _group_arrays(g):
    if isinstance(g, dict):
        return g["krho"], g["wrho"], g["phi"], g["wphi"]
    return g


def z_skeletons_grouped(stack, omega, source_layer, groups, z_src, z_obs_by_layer, tol=1e-6, device=0):
    """z_skeletons_grouped (mod_GreenOperator.jl:372-382): one skeleton per
    side for all groups, from each layer's modes at every (k_rho, phi) node of
    every group (phi = 0 only, each k_rho once, on an azimuthally symmetric stack)."""
    from .kspace import _is_azimuthally_symmetric
    sym = _is_azimuthally_symmetric(stack)
    kx, ky = [], []
    for g in groups:
        kr, _, ph, _ = (np.asarray(a, np.float64) for a in _group_arrays(g))
        ph = np.zeros(1) if sym else ph
        K, P = np.meshgrid(kr, ph, indexing="ij")
        kx.append((K * np.cos(P)).ravel(order="F"))
        ky.append((K * np.sin(P)).ravel(order="F"))
    kx, ky = np.concatenate(kx), np.concatenate(ky)
    if sym:
        kx, i = np.unique(kx, return_index=True)
        ky = ky[i]
    return _z_skeletons(stack, omega, source_layer, kx, ky, z_src, z_obs_by_layer, tol, device)


def _build(cache, grid, z_src, z_obs_by_layer, structure, check, check_tol, skeleton, skeleton_tol, lateral,
           nufft_eps, uniform_z, box_lateral, dz_box, ez_local=True, skeletons=None):
    if structure not in _lib.STRUCTURE:
        raise ValueError("build_layer_operator: structure must be :auto, :structured or :dense")
    for o in z_obs_by_layer:
        cache.check_obs(o)
    zs = np.ascontiguousarray(np.atleast_1d(np.asarray(z_src, np.float64)))
    obs, n_obs, layers_c, zo = _obs_arrays(z_obs_by_layer)
    opt = _lib.YtfOperatorOptions(_lib.STRUCTURE[structure], int(bool(check)), float(check_tol),
                                  int(bool(skeleton)), float(skeleton_tol), _lateral(lateral), float(nufft_eps),
                                  int(bool(uniform_z)), float(dz_box), int(bool(ez_local)),
                                  skeletons.handle if skeletons is not None else None, *cache.rotation_fields())
    W = cache.weights(box_lateral)
    ctx = _lib.Context.get(cache.device)
    ctx.bind_stream()
    h = ct.c_void_p()
    warn = ct.c_uint32(0)
    rc = _lib.lib().ytf_operator_build(ctx.handle, layer_structs(cache.stack), len(cache.stack), cache.omega,
                                       cache.source_layer, cache.n_nodes, _dptr(cache.kx), _dptr(cache.ky), _dptr(W),
                                       ct.byref(_grid_struct(grid)), zs.ctypes.data_as(_DP), len(zs), len(obs),
                                       layers_c, n_obs, zo.ctypes.data_as(_DP), ct.byref(opt), ct.byref(warn),
                                       ct.byref(h))
    _lib.check(rc, ctx.handle)
    kwargs = dict(structure=structure, check=check, check_tol=check_tol, skeleton=skeleton,
                  skeleton_tol=skeleton_tol, lateral=lateral, nufft_eps=nufft_eps, box_lateral=box_lateral,
                  dz_box=dz_box, ez_local=ez_local)
    op = LayerGreenOperator(h, cache, grid, zs, kwargs, cache.device)
    op.skeletons = skeletons
    if warn.value & _lib.WARN_STRUCTURE_CHECK:
        for b in op.blocks:
            if b["obs_layer"] == cache.source_layer and b["kind"] != "structured":
                warnings.warn(f"build_layer_operator: Toeplitz+Hankel check failed on layer {b['obs_layer']} "
                              f"(residual {b['check_residual']}) -- using the dense form")
    return op


def build_layer_operator(cache, grid, z_src, z_obs_by_layer, structure="auto", check=True, check_tol=1e-8,
                         skeleton=True, skeleton_tol=1e-6, skeletons=None, lateral="direct", nufft_eps=1e-9,
                         box_lateral=None, dz_box=0.0, ez_local=True):
    """build_layer_operator (mod_GreenOperator.jl:430-480): the Ĝ blocks of
    `cache`'s source layer onto every layer of z_obs_by_layer (uniform z grids),
    Toeplitz+Hankel where the media allow, dense (or skeleton) elsewhere.
    `skeletons`: precomputed z_skeletons / z_skeletons_grouped shared by the
    blocks without z structure (by default built from this cache's nodes)."""
    return _build(cache, grid, z_src, z_obs_by_layer, structure, check, check_tol, skeleton, skeleton_tol, lateral,
                  nufft_eps, True, box_lateral, dz_box, ez_local, skeletons)


def build_layer_operator_grouped(stack, omega, source_layer, groups, grid, z_src, z_obs_by_layer, device=0,
                                 node_budget=900_000, **kwargs):
    """build_layer_operator_grouped (mod_GreenOperator.jl:497-531): the
    operator of several (k_rho, w_rho, phi, w_phi) groups, one cache and build
    per group, a group of more than `node_budget` k_rho x phi nodes split into
    phi chunks, the chunk operators summed block by block on the device
    (ytf_operator_accumulate).  The z skeletons are built once on the nodes of
    all groups (z_skeletons_grouped) and shared; only the first chunk runs the
    structure check.  Returns (op, t_cache, t_build)."""
    obs = sorted(z_obs_by_layer)
    op = None
    t_cache = t_build = 0.0
    sk = kwargs.pop("skeletons", None)
    if sk is None and kwargs.get("skeleton", True) and kwargs.get("structure", "auto") != "dense":
        t0 = time.perf_counter()
        sk = z_skeletons_grouped(stack, omega, source_layer, groups, z_src, z_obs_by_layer,
                                 tol=kwargs.get("skeleton_tol", 1e-6), device=device)
        t_build += time.perf_counter() - t0
    check = kwargs.pop("check", True)
    L = _lib.lib()
    for g in groups:
        kr, wr, ph, wp = (np.asarray(a, np.float64) for a in _group_arrays(g))
        chunk = int(np.clip(node_budget // max(len(kr), 1), 1, max(len(ph), 1)))
        for r0 in range(0, len(ph), chunk):
            r = slice(r0, r0 + chunk)
            t0 = time.perf_counter()
            c = CavityTransfer2D(stack, omega, source_layer, kr, wr, ph[r], wp[r], obs_layers=obs, device=device)
            t_cache += time.perf_counter() - t0
            t0 = time.perf_counter()
            opc = build_layer_operator(c, grid, z_src, z_obs_by_layer, skeletons=sk,
                                       check=(op is None and check), **kwargs)
            if op is None:
                op = opc
            else:
                ctx = _lib.Context.get(device)
                ctx.bind_stream()
                _lib.check(L.ytf_operator_accumulate(ctx.handle, op.handle, opc.handle), ctx.handle)
                del opc
            t_build += time.perf_counter() - t0
    if op is not None:
        op.refresh()
    return op, t_cache, t_build


def apply_operator(op, P):
    """apply_operator (mod_GreenOperator.jl:674-712).  P: (3, Ns, Ny, Nx)
    dipole moments (column-major (Nx, Ny, Ns, 3)).  A torch tensor stays on the
    device and gives {obs_layer: (3, No, Ny, Nx) complex128 tensor}; a numpy
    array is the reference's host array and gives host arrays, staged through
    the operator's own device buffers (ytf_operator_apply_host)."""
    g = op.grid
    Ns = len(op.z_src)
    shape = (3, Ns, g["Ny"], g["Nx"])
    if isinstance(P, np.ndarray):
        if P.shape != shape:
            raise ValueError(f"apply_operator: P must be (Nx, Ny, Ns, 3) = {(g['Nx'], g['Ny'], Ns, 3)}, got "
                             f"{tuple(reversed(P.shape))}")
        Ph = np.ascontiguousarray(P, dtype=np.complex128)
        outs = [np.empty((3, b["n_obs"], g["Ny"], g["Nx"]), np.complex128) for b in op.blocks]
        ptrs = (ct.c_void_p * max(len(outs), 1))(*[E.ctypes.data for E in outs])
        ctx = _lib.Context.get(op.device)
        ctx.bind_stream()
        _lib.check(_lib.lib().ytf_operator_apply_host(ctx.handle, op.handle, Ph.ctypes.data, ptrs), ctx.handle)
        return {b["obs_layer"]: E for b, E in zip(op.blocks, outs)}
    torch = _torch()
    dev = torch.device("cuda", op.device)
    P = torch.as_tensor(P, dtype=torch.complex128, device=dev)
    if tuple(P.shape) != shape:
        raise ValueError(f"apply_operator: P must be (Nx, Ny, Ns, 3) = {(g['Nx'], g['Ny'], Ns, 3)}, got "
                         f"{tuple(reversed(P.shape))}")
    P = P.contiguous()
    outs = [torch.empty((3, b["n_obs"], g["Ny"], g["Nx"]), dtype=torch.complex128, device=dev) for b in op.blocks]
    ptrs = (ct.c_void_p * max(len(outs), 1))(*[E.data_ptr() for E in outs])
    ctx = _lib.Context.get(op.device)
    ctx.bind_stream()
    _lib.check(_lib.lib().ytf_operator_apply(ctx.handle, op.handle, _dptr(P), ptrs), ctx.handle)
    return {b["obs_layer"]: E for b, E in zip(op.blocks, outs)}


def dipole_moments(op, dips):
    """Moments array for apply_operator: each dipole snapped to its pixel and
    to its source plane (dipole_moments, mod_GreenOperator.jl:720-732).
    Returns a (3, Ns, Ny, Nx) complex128 numpy array."""
    g = op.grid
    P = np.zeros((3, len(op.z_src), g["Ny"], g["Nx"]), np.complex128)
    tol = 1e-9 * max(1.0, float(np.max(np.abs(op.z_src))))
    for d in dips:
        if d.layer != op.source_layer:
            raise ValueError(f"dipole_moments: dipole in layer {d.layer}, operator source layer {op.source_layer}")
        k = next((k for k, z in enumerate(op.z_src) if abs(z - d.z) <= tol), None)
        if k is None:
            raise ValueError(f"dipole_moments: dipole depth {d.z} is not one of the operator's source planes")
        i, j = nearest_node_index(g, d.x, d.y)
        P[:, k, j - 1, i - 1] += _moment_vec(d.p)
    return P


def operator_summary(op, file=sys.stdout):
    """Memory and build time per block (operator_summary, mod_GreenOperator.jl:735-747)."""
    tot = 0
    for b in op.blocks:
        tot += b["bytes"]
        sk = ""
        if b["kind"] in ("skeleton", "toeplitz_skeleton"):
            sk = f"  skeleton {b['n_src_skel']} × {b['n_obs_skel']} of {len(op.z_src)} × {b['n_obs']} planes"
        print(f"  source layer {op.source_layer} → obs layer {b['obs_layer']}: {b['kind']:<17s} {b['n_obs']:3d} obs "
              f"planes  {b['bytes'] / 2 ** 20:8.1f} MB  built in {b['build_s']:.2f} s{sk}", file=file)
    print(f"  total {tot / 2 ** 20:.1f} MB", file=file)


# ---------------------------------------------------------------------------
# Multi-source lateral assembly -- mod_ClosedFormAssembly.jl
# ---------------------------------------------------------------------------
class ClosedFormGreen:
    """Ĝ of one (source layer, source depth, obs layer) at depths z_obs on the
    padded grid (ClosedFormGreen, mod_ClosedFormAssembly.jl:51-58), held as a
    one-source-plane dense operator."""

    def __init__(self, op, obs_layer, zsrc_local, grid, z_obs):
        self.op, self.source_layer, self.obs_layer = op, op.source_layer, obs_layer
        self.zsrc_local, self.grid, self.z_obs = float(zsrc_local), grid, np.asarray(z_obs, np.float64)

    @property
    def Ghat(self):
        """(3, 3, Nz, Py, Px) tensor = the reference's (Px, Py, Nz, 3, 3) with
        dx dy / (Px Py) folded in."""
        torch = _torch()
        planes = [self.op.ghat_plane(0, "dense", j) for j in range(len(self.z_obs))]
        return torch.stack(planes, dim=2) * (self.grid["dx"] * self.grid["dy"])


def build_closed_form_green(cache, obs_layer, zsrc_local, grid, z_obs, lateral="direct", nufft_eps=1e-9,
                            box_lateral=None, dz_box=0.0):
    """build_closed_form_green (mod_ClosedFormAssembly.jl:110-124)."""
    op = _build(cache, grid, [zsrc_local], {obs_layer: list(np.atleast_1d(z_obs))}, "dense", False, 1e-8, False,
                1e-6, lateral, nufft_eps, False, box_lateral, dz_box)
    return ClosedFormGreen(op, obs_layer, zsrc_local, grid, np.atleast_1d(z_obs))


def _check_sheet(grid, sheet, who):
    J = np.asarray(sheet.J) if not hasattr(sheet.J, "shape") else sheet.J
    if tuple(J.shape) != (3, grid["Ny"], grid["Nx"]):
        raise ValueError(f"{who}: sheet.J has shape {tuple(reversed(J.shape))}, but this grid is (Nx,Ny,3) = "
                         f"{(grid['Nx'], grid['Ny'], 3)} -- build the sheet against this SAME grid (e.g. "
                         "`point_current_sheet(grid, ...)`).")


def _sheet_moments(grid, sheets):
    torch = _torch()
    J = sum(torch.as_tensor(s.J, dtype=torch.complex128) * s.weight for s in sheets)
    return J * (grid["dx"] * grid["dy"])


def apply_ghat(g, sheet):
    """apply_ghat (mod_ClosedFormAssembly.jl:191-196): field (3, Nz, Ny, Nx)
    of one sheet at g's source depth."""
    _check_sheet(g.grid, sheet, "apply_ghat")
    P = _sheet_moments(g.grid, [sheet])[:, None]
    return apply_operator(g.op, P)[g.obs_layer]


def apply_closed_form(cache, obs_layer, grid, z_obs, sheets, **kw):
    """apply_closed_form (mod_ClosedFormAssembly.jl:206-219): every sheet of
    the source layer, grouped by depth, one operator over the distinct depths."""
    torch = _torch()
    for s in sheets:
        _check_sheet(grid, s, "apply_closed_form")
    depths = sorted({float(s.z) for s in sheets})
    op = _build(cache, grid, depths, {obs_layer: list(np.atleast_1d(z_obs))}, "dense", False, 1e-8, False, 1e-6,
                kw.get("lateral", "direct"), kw.get("nufft_eps", 1e-9), False, kw.get("box_lateral"),
                kw.get("dz_box", 0.0))
    P = torch.stack([_sheet_moments(grid, [s for s in sheets if float(s.z) == z]) for z in depths], dim=1)
    return apply_operator(op, P)[obs_layer]


def precompute_ghats(ghat_cache, caches, grid, z_src_by_layer, z_obs_by_layer, **kw):
    """precompute_ghats! (mod_ClosedFormAssembly.jl:134-143): one dense
    operator per source layer over its declared depths and every obs layer,
    keyed (source_layer, depths, obs layers) as closed_form_field looks them up."""
    for l, zss in z_src_by_layer.items():
        depths = tuple(sorted({float(z) for z in zss}))
        key = (l, depths, tuple(sorted(o for o, zo in z_obs_by_layer.items() if len(zo))))
        if key not in ghat_cache:
            ghat_cache[key] = _build(caches[l], grid, list(depths), {o: zo for o, zo in z_obs_by_layer.items()
                                                                     if len(zo)},
                                     "dense", False, 1e-8, False, 1e-6, kw.get("lateral", "direct"),
                                     kw.get("nufft_eps", 1e-9), False, kw.get("box_lateral"), kw.get("dz_box", 0.0))
    return ghat_cache


def closed_form_field(caches, grid, z_obs_by_layer, dips, ghat_cache=None, **kw):
    """closed_form_field (mod_ClosedFormAssembly.jl:242-270): multi-dipole,
    multi-layer field on `grid`.  Dipoles are grouped by source layer and
    exact depth; per source layer one dense operator over its depths; fields
    of all source layers are summed per obs layer.  Returns
    {obs_layer: (3, Nz, Ny, Nx) tensor}."""
    torch = _torch()
    if not dips:
        raise ValueError("closed_form_field: need at least one source")
    by_layer = {}
    for d in dips:
        if d.layer not in caches:
            raise ValueError(f"closed_form_field: no CavityTransferCache for source layer {d.layer}")
        by_layer.setdefault(d.layer, []).append(d)
    obs = {o: zo for o, zo in z_obs_by_layer.items() if len(np.atleast_1d(zo))}
    out = {}
    for l, ds in by_layer.items():
        depths = tuple(sorted({float(d.z) for d in ds}))
        key = (l, depths, tuple(sorted(obs)))
        if ghat_cache is not None and key in ghat_cache:
            op = ghat_cache[key]
        else:
            op = _build(caches[l], grid, list(depths), obs, "dense", False, 1e-8, False, 1e-6,
                        kw.get("lateral", "direct"), kw.get("nufft_eps", 1e-9), False, kw.get("box_lateral"),
                        kw.get("dz_box", 0.0))
            if ghat_cache is not None:
                ghat_cache[key] = op
        # one sheet per depth: moments snapped to pixels (point_current_sheet's
        # density times dx dy)
        E = apply_operator(op, torch.as_tensor(dipole_moments(op, ds), device=torch.device("cuda", op.device)))
        for o, e in E.items():
            out[o] = out[o] + e if o in out else e
    dev = torch.device("cuda", next(iter(caches.values())).device)
    for o, zo in z_obs_by_layer.items():
        if o not in out:
            out[o] = torch.zeros((3, len(np.atleast_1d(zo)), grid["Ny"], grid["Nx"]), dtype=torch.complex128,
                                 device=dev)
    return out
