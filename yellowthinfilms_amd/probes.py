"""Probe integrands of the adaptive k_rho grids, evaluated for many k_rho at once.

The reference's box grid (`box_kρ_grid`, mod_ClosedFormGrid.jl:155-255)
converges its k_rho axis on probe integrands: the box Green's tensor at a few
probe points, each a function of k_rho alone, evaluated through the same 2D
synthesis the field uses.  For every visited k_rho it builds a one-node
`CavityTransferCache2D` (that k_rho with weight 1, every phi of the grid) and
evaluates `closed_form_G` at each probe (`probe_val`, `vals_at`, :204-218),
about 1.3 ms per k_rho.

Here a whole batch of k_rho is one job: `ytf_polar_nodes` lays out the nodes
and weights of every (k_rho, phi), one `ytf_green_nodes` launch per
observation layer solves them with that layer's depth pairs, and
`ytf_segment_lateral_sum` forms every (k_rho, probe) value at once (one
segment = the phi nodes of one k_rho).
"""
import ctypes as ct
import math

import numpy as np

from . import _lib
from .green import green_nodes


def _torch():
    import torch
    return torch


def thinnest_finite_layer(stack):
    """Minimum thickness among the layers 1 .. N-1 (thinnest_finite_layer,
    mod_KxGrid.jl:153-157)."""
    return math.inf if len(stack) <= 1 else min(float(l.d) for l in stack[:-1])


def box_probes(stack, source_layer, box, zbar=None, lateral_dir=0.0, rho_far=0.0, include_interior=True):
    """The probes of box_kρ_grid (mod_ClosedFormGrid.jl:163-197): a list of
    (obs_layer, z_obs, x, y, h, (zbar, tag)) with z_obs in the observation
    layer's frame, (x, y) the lateral offset and h the distance from the box's
    near face.  For each box-centre depth zbar (local, caller's frame; default
    the layer centre and both near-face positions that fit):
      self voxel and lateral neighbour (inside the slab, h = 0; with
      include_interior), vertical neighbours +-dz in the layer (h = dz/2),
      the layers above and below just across each interface, and with
      rho_far > 0 a far lateral point just outside the slab and one inside it."""
    N = len(stack)
    if not 1 <= source_layer < N:
        raise ValueError(f"box_kρ_grid: source_layer={source_layer} must be in 1:{N - 1}")
    dx, dy, dz = (float(v) for v in box)
    d = float(stack[source_layer - 1].d)
    if dz > d:
        raise ValueError(f"box_kρ_grid: Δz_box={dz} does not fit in layer {source_layer} (d={d})")
    if zbar is None:
        zb = []
        for z in (d / 2, dz / 2 + 1e-3 * d, d - dz / 2 - 1e-3 * d):
            if z not in zb:
                zb.append(z)
    else:
        zb = [float(z) for z in np.atleast_1d(zbar)]
    zb = [z for z in zb if dz / 2 - 1e-12 <= z <= d - dz / 2 + 1e-12]
    if not zb:
        raise ValueError("box_kρ_grid: no box-centre depth fits in the layer")
    zpos_of = (lambda z: d - z) if source_layer == 1 else (lambda z: z)
    ct_, st_ = math.cos(lateral_dir), math.sin(lateral_dir)
    probes = []
    for zbr in zb:
        zp = zpos_of(zbr)
        if include_interior:
            probes.append((source_layer, zp, 0.0, 0.0, 0.0, (zbr, "self voxel")))
            probes.append((source_layer, zp, dx * ct_, dy * st_, 0.0, (zbr, "lateral neighbour")))
        if zp + dz <= d:
            probes.append((source_layer, zp + dz, 0.0, 0.0, dz / 2, (zbr, "vertical neighbour +")))
        if zp - dz >= 0:
            probes.append((source_layer, zp - dz, 0.0, 0.0, dz / 2, (zbr, "vertical neighbour −")))
        delta = 0.5 * min(dz, thinnest_finite_layer(stack))
        if source_layer > 1:
            dup = float(stack[source_layer - 2].d)
            probes.append((source_layer - 1, dup - delta, 0.0, 0.0, zp - dz / 2 + delta,
                           (zbr, "layer above, near interface")))
        probes.append((source_layer + 1, delta, 0.0, 0.0, (d - zp - dz / 2) + delta, (zbr, "layer below, near interface")))
        if rho_far > 0:
            zf = zp + dz / 2 + delta if zp + dz / 2 + delta <= d else zp - dz / 2 - delta
            if zf >= 0:
                probes.append((source_layer, zf, rho_far * ct_, rho_far * st_, delta, (zbr, "far lateral")))
            if include_interior:
                probes.append((source_layer, zp, rho_far * ct_, rho_far * st_, 0.0, (zbr, "far lateral, inside the slab")))
    return probes


def probe_values(stack, omega, source_layer, krho, phis, wphis, probes, box, device=0):
    """The probe integrands at every k_rho of `krho`: (n_krho, n_probes, 3, 3)
    complex, entry [i, m, c, q] = closed_form_G of the one-node cache (krho[i]
    with weight 1, the phi grid (phis, wphis)) at probe m, source depth zbar,
    observation depth z_obs and offset (x, y), for a box source of footprint
    (dx, dy) and thickness dz (the box factors, the interior slab and the local
    Ez) -- `vals_at` (mod_ClosedFormGrid.jl:204-218) for all k_rho at once."""
    torch = _torch()
    dev = torch.device("cuda", device)
    kr = np.ascontiguousarray(np.atleast_1d(np.asarray(krho, np.float64)))
    ph = np.asarray(phis, np.float64)
    wp = np.asarray(wphis, np.float64)
    nk, L = kr.size, ph.size
    out = np.zeros((nk, len(probes), 3, 3), np.complex128)
    if nk == 0 or not probes:
        return out
    dx, dy, dz = (float(v) for v in box)
    # node k = i L + j: k_rho i, phi j (one segment per k_rho); W = k_rho w_phi / 4 pi^2 x the box footprint
    kr_d = torch.as_tensor(kr, device=dev)
    ph_d = torch.as_tensor(np.ascontiguousarray(ph), device=dev)
    wp_d = torch.as_tensor(np.ascontiguousarray(wp), device=dev)
    kx_d = torch.empty(nk * L, dtype=torch.float64, device=dev)
    ky_d = torch.empty(nk * L, dtype=torch.float64, device=dev)
    W_d = torch.empty(nk * L, dtype=torch.complex128, device=dev)
    ctx = _lib.Context.get(device)
    ctx.bind_stream()
    _lib.check(_lib.lib().ytf_polar_nodes(ctx.handle, nk, ct.c_void_p(kr_d.data_ptr()), L, ct.c_void_p(ph_d.data_ptr()),
                                          ct.c_void_p(wp_d.data_ptr()), dx, dy, ct.c_void_p(kx_d.data_ptr()),
                                          ct.c_void_p(ky_d.data_ptr()), ct.c_void_p(W_d.data_ptr())), ctx.handle)
    # every observation layer's launches first, the values to the host once at the end
    # (one wait instead of one per layer)
    parts = []
    for obs in sorted({p[0] for p in probes}):
        idx = [m for m, p in enumerate(probes) if p[0] == obs]
        zs = [probes[m][5][0] for m in idx]
        zo = [probes[m][1] for m in idx]
        _, g = green_nodes(stack, omega, kx_d, ky_d, source_layer, obs, zs, zo, cache=False, ghat=True,
                           device=device, dz_box=dz, ez_local=True)
        xp = np.ascontiguousarray([probes[m][2] for m in idx], np.float64)
        yp = np.ascontiguousarray([probes[m][3] for m in idx], np.float64)
        vals = torch.empty((nk, len(idx), 9), dtype=torch.complex128, device=dev)
        ctx.bind_stream()
        _lib.check(_lib.lib().ytf_segment_lateral_sum(
            ctx.handle, nk, L, ct.c_void_p(kx_d.data_ptr()), ct.c_void_p(ky_d.data_ptr()),
            ct.c_void_p(W_d.data_ptr()), len(idx), ct.c_void_p(g.data_ptr()), xp.ctypes.data_as(ct.c_void_p),
            yp.ctypes.data_as(ct.c_void_p), ct.c_void_p(vals.data_ptr())), ctx.handle)
        del g
        parts.append((idx, vals))
    host = torch.cat([v.reshape(nk, -1) for _, v in parts], dim=1).cpu().numpy()
    col = 0
    for idx, _ in parts:
        v = host[:, col:col + 9 * len(idx)].reshape(nk, len(idx), 3, 3)  # entry c + 3q -> [q][c]
        out[:, idx] = np.swapaxes(v, 2, 3)
        col += 9 * len(idx)
    return out


def check_depths(stack, source_layer, z_fracs, hmin):
    """_check_depths (mod_ClosedFormGrid.jl:474-481): the source depths
    z_frac d (clamped inside the layer) at least hmin from the layer's real
    interfaces (layer 1: only its bottom one); if none is, hmin itself."""
    d = float(stack[source_layer - 1].d)
    dist = (lambda z: z) if source_layer == 1 else (lambda z: min(z, d - z))
    zs = [min(max(zf * d, 1e-9 * d), d - 1e-9 * d) for zf in z_fracs]
    ok = [z for z in zs if dist(z) >= hmin]
    return ok if ok else [min(hmin, d)]


def point_probes(stack, source_layer, z_fracs, dist, rho_max):
    """_point_probes (mod_ClosedFormGrid.jl:453-472): (obs_layer, z_obs,
    z_src, rho, moment) for each depth z_frac d: in the source layer dist
    above and below the source and both layer edges, dist into each
    neighbouring layer; moments x and z; rho = 0 and rho_max."""
    d = float(stack[source_layer - 1].d)
    eps = 1e-9 * max(d, 1.0)
    probes = []
    for zf in z_fracs:
        z = min(max(zf * d, eps), d - eps)
        zp = d - z if source_layer == 1 else z
        pts = []
        if zp + dist <= d:
            pts.append((source_layer, zp + dist))
        if zp - dist >= 0:
            pts.append((source_layer, zp - dist))
        pts += [(source_layer, eps), (source_layer, d - eps)]
        if source_layer > 1:
            pts.append((source_layer - 1, float(stack[source_layer - 2].d) - dist))
        pts.append((source_layer + 1, min(dist, float(stack[source_layer].d))))
        rhos = [0.0] if rho_max == 0.0 else [0.0, float(rho_max)]
        for obs, zo in pts:
            for mom in ((1.0, 0.0, 0.0), (0.0, 0.0, 1.0)):
                for rho in rhos:
                    probes.append((obs, zo, z, rho, mom))
    return probes


def point_probe_values(stack, omega, source_layer, kx, probes, wt=None, device=0):
    """The point field of every probe, node by node: (n_kx, n_probes, 3)
    complex, entry [k, m] = closed_form_pixel at probe m of the 1D cache of
    node k alone (weight wt[k], default 1) -- the probe integrands of
    point_kx_grid (cache_at, mod_ClosedFormGrid.jl:407-411) for a whole batch
    of k_x.  Summed over k with the rule's weights it is the probes' field on
    that rule (_probe_values_1d, :483-487).  One ytf_green_nodes launch per
    observation layer (its distinct (z_src, z_obs) pairs), then
    ytf_radial_terms at the probes' radii."""
    torch = _torch()
    dev = torch.device("cuda", device)
    kxh = np.ascontiguousarray(np.atleast_1d(np.asarray(kx, np.float64)))
    n = kxh.size
    out = np.zeros((n, len(probes), 3), np.complex128)
    if n == 0 or not probes:
        return out
    wth = np.ones(n) if wt is None else np.ascontiguousarray(np.asarray(wt, np.float64))
    if wth.shape != (n,):
        raise ValueError("point_probe_values: wt must have one weight per k_x")
    kx_d = torch.as_tensor(kxh, device=dev)
    ky_d = torch.zeros(n, dtype=torch.float64, device=dev)
    wt_d = torch.as_tensor(wth, device=dev)
    ctx = _lib.Context.get(device)
    for obs in sorted({p[0] for p in probes}):
        idx = [m for m, p in enumerate(probes) if p[0] == obs]
        pairs = sorted({(probes[m][2], probes[m][1]) for m in idx})
        rhos = sorted({probes[m][3] for m in idx})
        _, g = green_nodes(stack, omega, kx_d, ky_d, source_layer, obs, [p[0] for p in pairs], [p[1] for p in pairs],
                           cache=False, ghat=True, device=device)
        rho_d = torch.as_tensor(np.asarray(rhos, np.float64), device=dev)
        T = torch.empty((n, len(pairs), 9, len(rhos)), dtype=torch.complex128, device=dev)
        ctx.bind_stream()
        _lib.check(_lib.lib().ytf_radial_terms(ctx.handle, n, ct.c_void_p(kx_d.data_ptr()), ct.c_void_p(wt_d.data_ptr()),
                                               len(pairs), ct.c_void_p(g.data_ptr()), len(rhos),
                                               ct.c_void_p(rho_d.data_ptr()), ct.c_void_p(T.data_ptr())), ctx.handle)
        Th = T.cpu().numpy().reshape(n, len(pairs), 3, 3, len(rhos))       # [k, pair, q, c, r] (entry c + 3q)
        for m in idx:
            j = pairs.index((probes[m][2], probes[m][1]))
            r = rhos.index(probes[m][3])
            out[:, m] = np.einsum("kqc,q->kc", Th[:, j, :, :, r], np.asarray(probes[m][4], np.complex128))
    return out
