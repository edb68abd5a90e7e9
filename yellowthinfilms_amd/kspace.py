"""Real-axis pole search of the k_rho integrand (host side of
mod_KSpaceFeatures.jl): where the field of a source has guided-mode,
plasmon or polariton poles, found before any quadrature is built.

  cavity_response  -- _cavity_response (:109-127) at a batch of nodes, one
                      GPU launch (ytf_cavity_response);
  real_axis_poles  -- real_axis_poles (:151-214): the log-spaced scan, the
                      prominent local maxima, golden-section location of
                      each and its half-width from the half-level crossings.
                      Every evaluation runs on the GPU: the scan of all
                      directions in one launch, then each refinement step of
                      all candidates at once (the reference refines one
                      candidate after another);
  clusters         -- _clusters (:228-241), the intervals a rule resolves;
  kspace_features  -- (:316-379) light lines, cut-off, poles, clusters, N_phi;
  fixed_krho_rule  -- fixed_kρ_rule (:407-433), the graded Gauss-Legendre rule
                      built from the features;
  adaptive_seeds   -- (:448-460) the seeds of an adaptive grid;
  material_class, choose_engine -- (:54-82).
"""
import ctypes as ct
import math

import numpy as np

from . import _lib
from .green import layer_structs
from .materials import stack_arrays


def _torch():
    import torch
    return torch


def cavity_response(stack, omega, source_layer, kx, ky, device=0):
    """_cavity_response at nodes (kx, ky) (kx = k_rho cos(phi), ky = k_rho
    sin(phi)); returns a float64 device tensor."""
    torch = _torch()
    dev = torch.device("cuda", device)
    kx = torch.as_tensor(kx, dtype=torch.float64, device=dev).contiguous().reshape(-1)
    ky = torch.as_tensor(ky, dtype=torch.float64, device=dev).contiguous().reshape(-1)
    if kx.shape != ky.shape:
        raise ValueError("cavity_response: kx and ky must have the same length")
    out = torch.empty_like(kx)
    ctx = _lib.Context.get(device)
    ctx.bind_stream()
    _lib.check(_lib.lib().ytf_cavity_response(ctx.handle, layer_structs(stack), len(stack), float(omega),
                                              int(source_layer), ct.c_void_p(kx.data_ptr()),
                                              ct.c_void_p(ky.data_ptr()), kx.shape[0], ct.c_void_p(out.data_ptr())),
               ctx.handle)
    return out


# Response evaluations of the last real_axis_poles call (scan, refinement).
search_stats = {"scan_evaluations": 0, "refine_evaluations": 0}


def _pole_candidates(R, ks, nphi, W, log_prom, log_floor_rel, log_floor_abs, device):
    """ytf_pole_candidates on the scan R (device, nphi x n), sorted by (phi
    index, i): arrays m, i and val (n_cand, 5) = (base, prom, k of the left
    valley, k of the right valley, y[i])."""
    torch = _torch()
    dev = torch.device("cuda", device)
    n = len(ks)
    ks_d = torch.as_tensor(ks, dtype=torch.float64, device=dev)
    ctx = _lib.Context.get(device)
    ctx.bind_stream()
    cap = nphi * (n // 2 + 1)  # local maxima are at least two samples apart
    while True:
        key = torch.empty(cap, dtype=torch.int64, device=dev)
        val = torch.empty((cap, 5), dtype=torch.float64, device=dev)
        cnt = ct.c_int64()
        _lib.check(_lib.lib().ytf_pole_candidates(
            ctx.handle, int(nphi), int(n), ct.c_void_p(R.data_ptr()), ct.c_void_p(ks_d.data_ptr()), int(W),
            float(log_prom), float(log_floor_rel), float(log_floor_abs), int(cap), ct.c_void_p(key.data_ptr()),
            ct.c_void_p(val.data_ptr()), ct.byref(cnt)), ctx.handle)
        if cnt.value <= cap:
            break
        cap = int(cnt.value)
    nc = int(cnt.value)
    key = key[:nc].cpu().numpy()
    val = val[:nc].cpu().numpy()
    o = np.argsort(key, kind="stable")
    return key[o] // n, key[o] % n, val[o]


def real_axis_poles(stack, omega, source_layer, phis=(0.0,), k_lo=None, k_hi=None, scan_step=1e-3,
                    prominence=1.05, exclude=(), floor_rel=1e-6, floor_abs=1e-10, device=0):
    """real_axis_poles (mod_KSpaceFeatures.jl:151-214).  Returns a list of
    dicts (q, Gamma, phi, prom):
      1. the response on k = k_lo e^{j scan_step} up to k_hi, every phi;
      2. local maxima standing `prominence` x above the higher of the two
         valleys around them (searched over a factor 4 in k) and above
         max(floor_rel x the largest response along that phi, floor_abs);
      3. q by golden-section search between the maximum's neighbours (60
         steps at most, to 1e-12 relative), Gamma as the nearer of the two
         crossings of b + (peak - b)/sqrt(2) (b the valley level), found by
         bisection towards each valley.
    Peaks within 1e-6 q of a value in `exclude` (the light lines) are dropped.
    Step 1 is one response launch; steps 2 and 3 run on the device
    (ytf_pole_candidates, ytf_pole_refine: every refinement evaluation inside
    the kernel)."""
    if k_lo is None or k_hi is None or not (0 < k_lo < k_hi):
        raise ValueError("real_axis_poles: need 0 < k_lo < k_hi")
    L = math.log(k_hi / k_lo)
    nstep = int(math.floor(L / scan_step + 1e-10))
    ks = k_lo * np.exp(np.arange(nstep + 1) * scan_step)   # Julia's 0.0:step:L
    n = len(ks)
    W = int(math.ceil(math.log(4.0) / scan_step))
    phis = [float(p) for p in phis]
    ph_all = np.asarray(phis, np.float64)[:, None]
    R = cavity_response(stack, omega, source_layer, (ks[None, :] * np.cos(ph_all)).ravel(),
                        (ks[None, :] * np.sin(ph_all)).ravel(), device=device)
    search_stats.update(scan_evaluations=len(phis) * n, refine_evaluations=0)
    # candidates (phi index, i; base, prom, left valley k, right valley k, y[i]),
    # found on the device (ytf_pole_candidates), in (phi, i) order
    cm, ci, cval = _pole_candidates(R, ks, len(phis), W, math.log(prominence), math.log(floor_rel),
                                    math.log(floor_abs), device)
    if cm.size == 0:
        return []
    ph = np.asarray(phis, np.float64)[cm]
    # golden-section location and half-level crossings, on the device (ytf_pole_refine)
    cv = np.stack([ph, ks[ci - 1], ks[ci + 1], cval[:, 2], cval[:, 3], cval[:, 0], cval[:, 4]], axis=1)
    out = _pole_refine(stack, omega, source_layer, cv, device)
    q, xl, xr = out[:, 0], out[:, 1], out[:, 2]
    keep = np.ones(q.size, bool)
    for Lx in exclude:
        keep &= ~(np.abs(q - Lx) <= 1e-6 * q)
    G = np.maximum(np.minimum(q - xl, xr - q), 1e-12 * q)
    ql, Gl, pl, rl = q.tolist(), G.tolist(), ph.tolist(), cval[:, 1].tolist()
    return [{"q": ql[j], "Gamma": Gl[j], "phi": pl[j], "prom": rl[j]} for j in np.nonzero(keep)[0].tolist()]


def _pole_refine(stack, omega, source_layer, cv, device):
    """ytf_pole_refine on the candidates cv (n, 7): (n, 3) = (q, left crossing,
    right crossing); the number of responses taken goes to `search_stats`."""
    torch = _torch()
    dev = torch.device("cuda", device)
    c = torch.as_tensor(np.ascontiguousarray(cv), device=dev)
    out = torch.empty((len(cv), 3), dtype=torch.float64, device=dev)
    ctx = _lib.Context.get(device)
    ctx.bind_stream()
    ne = ct.c_int64()
    _lib.check(_lib.lib().ytf_pole_refine(ctx.handle, layer_structs(stack), len(stack), float(omega),
                                          int(source_layer), len(cv), ct.c_void_p(c.data_ptr()),
                                          ct.c_void_p(out.data_ptr()), ct.byref(ne)), ctx.handle)
    search_stats["refine_evaluations"] = int(ne.value)
    return out.cpu().numpy()


def clusters(poles, merge_rel):
    """_clusters (mod_KSpaceFeatures.jl:228-241): one [q, q] interval per pole
    plus, for two poles adjacent in q within merge_rel q (or 4 widths) of each
    other, the segment between them at the smaller width."""
    ps = sorted(poles, key=lambda p: p["q"])
    out = []
    for i, p in enumerate(ps):
        e = (p["q"], p["q"], p["Gamma"])
        if e not in out:
            out.append(e)
        if i > 0:
            r = ps[i - 1]
            if p["q"] - r["q"] <= max(merge_rel * p["q"], 4 * min(p["Gamma"], r["Gamma"])) and p["q"] > r["q"]:
                e = (r["q"], p["q"], min(p["Gamma"], r["Gamma"]))
                if e not in out:
                    out.append(e)
    return [dict(a=a, b=b, Gamma=G) for a, b, G in out]


def material_class(layer):
    """material_class (mod_KSpaceFeatures.jl:54-62): "gyrotropic" for a
    non-symmetric eps or mu, else from the signs of the eigenvalues of Re eps
    and Re mu: "dielectric", "metallic" (all of eps negative) or "hyperbolic"."""
    eps, mu, _, _, _ = stack_arrays([layer])
    e, m = eps[0], mu[0]
    if not (np.array_equal(e, e.T) and np.array_equal(m, m.T)):
        return "gyrotropic"
    ee = np.linalg.eigvalsh(e.real)
    em = np.linalg.eigvalsh(m.real)
    if np.all(ee > 0) and np.all(em > 0):
        return "dielectric"
    if np.all(ee < 0) and np.all(em > 0):
        return "metallic"
    return "hyperbolic"


def choose_engine(stack, engine="auto"):
    """choose_engine (mod_KSpaceFeatures.jl:71-75): "adaptive" when every layer
    is dielectric, "fixed" as soon as one is metallic, hyperbolic or gyrotropic."""
    if engine not in ("auto", "fixed", "adaptive"):
        raise ValueError(f"engine must be :auto, :fixed or :adaptive (got :{engine})")
    if engine != "auto":
        return engine
    return "adaptive" if all(material_class(l) == "dielectric" for l in stack) else "fixed"


def _is_azimuthally_symmetric(stack):
    """_is_azimuthally_symmetric_stack (mod_ModeSolver.jl:427-434): every layer
    takes the closed-form branch (isotropic / uniaxial eps about z, isotropic mu)."""
    from .materials import EPS_ISOTROPIC, EPS_UNIAXIAL, IsotropicMaterial
    return all(l.eps.kind in (EPS_ISOTROPIC, EPS_UNIAXIAL) and isinstance(l.mu, IsotropicMaterial) for l in stack)


def _pole_track_eta(poles, phi_scan):
    """_pole_track_η (mod_KSpaceFeatures.jl:257-275): the smallest phi-width
    Gamma dphi / |dq| of the phi-peaks that poles moving with phi put in the 2D
    integrand (inf with fewer than two directions or no track)."""
    ph = sorted(set(float(p) for p in phi_scan))
    if len(ph) < 2:
        return math.inf
    pairs = [(ph[i], ph[i + 1], ph[i + 1] - ph[i]) for i in range(len(ph) - 1)]
    step = ph[1] - ph[0]
    if abs(ph[-1] - ph[0] + step - 2 * math.pi) < 1e-9:
        pairs.append((ph[-1], ph[0], step))
    eta = math.inf
    for (f1, f2, dphi) in pairs:
        for p in poles:
            if p["phi"] != f1:
                continue
            for r in poles:
                if r["phi"] != f2:
                    continue
                dq = abs(p["q"] - r["q"])
                if 0 < dq <= dphi * max(p["q"], r["q"]):
                    eta = min(eta, min(p["Gamma"], r["Gamma"]) * dphi / dq)
    return eta


def kspace_features(stack, omega, source_layer, h_min, rho_max=0.0, dims=1, n_phi_scan=32, target_decay=25.0,
                    krho_cap=None, krho_cap_rel=1e5, scan=True, scan_step=1e-3, prominence=1.05, c_phi=2.5,
                    n_phi_min=64, Kcut=None, phi_scan=None, rtol=1e-6, c_eta=1.5, device=0):
    """kspace_features (mod_KSpaceFeatures.jl:316-379): light lines (all layers,
    every scanned direction; the semi-infinite layers' ones as singular lines),
    kmax, the cut-off Kcut (target_decay / h_min within [2 kmax, cap]), the
    real-axis poles of the source layer (real_axis_poles on the GPU) and their
    clusters, and in 2D the phi count N_phi = max(n_phi_min, c_phi Kcut rho_max,
    c_eta log(10 / rtol) / eta_min) rounded up to even."""
    from . import quadrature as Q
    if dims not in (1, 2):
        raise ValueError("kspace_features: dims must be 1 or 2")
    N = len(stack)
    w = float(omega)
    classes = [material_class(l) for l in stack]
    sym = _is_azimuthally_symmetric(stack)
    eps_mu = [(np.asarray(l.eps.eps), np.asarray(l.mu.eps)) for l in stack]
    diag = all(np.count_nonzero(e - np.diag(np.diag(e))) == 0 and np.count_nonzero(m - np.diag(np.diag(m))) == 0
               for e, m in eps_mu)
    if dims == 1 or sym:
        phis, merge_rel = [0.0], 0.0
    elif phi_scan is not None:
        phis = sorted(float(p) for p in phi_scan)
        merge_rel = min(np.diff(phis)) if len(phis) > 1 else 0.0
    elif diag:
        phis = list(np.linspace(0.0, math.pi / 2, max(2, n_phi_scan // 4 + 1)))
        merge_rel = (math.pi / 2) / max(1, n_phi_scan // 4)
    else:
        phis = [2 * math.pi * p / n_phi_scan for p in range(n_phi_scan)]
        merge_rel = 2 * math.pi / n_phi_scan
    phi_ll = [0.0] if (dims == 1 or sym) else sorted(set(float(np.mod(p, math.pi)) for p in phis))
    ll, sing = [], []
    for j, lay in enumerate(stack):
        for p in phi_ll:
            L = Q.layer_light_lines(lay, w, p)
            ll += L
            if j == 0 or j == N - 1:
                sing += L
    ll = Q.dedup(sorted(ll))
    sing = Q.dedup(sorted(sing))
    kmax = ll[-1] if ll else w
    K_geom = Q.decay_cutoff(h_min, target_decay)
    K_cap = krho_cap_rel * kmax if krho_cap is None else float(krho_cap)
    capped = False
    if Kcut is None:
        Kcut = Q.tail_bound(K_geom, kmax, krho_cap=K_cap, kx_ratio_min=2.0)
        capped = Kcut < K_geom
    else:
        Kcut = float(Kcut)
    poles = []
    if scan:
        k_lo = 0.5 * (ll[0] if ll else w)
        poles = real_axis_poles(stack, w, source_layer, phis=phis, k_lo=k_lo, k_hi=Kcut, scan_step=scan_step,
                                prominence=prominence, exclude=ll, device=device)
    cl = clusters(poles, merge_rel)
    n_lat = int(math.ceil(c_phi * Kcut * rho_max))
    eta = _pole_track_eta(poles, phis)
    n_pole = int(math.ceil(c_eta * math.log(10 / rtol) / eta)) if math.isfinite(eta) else 0
    n_phi = 2 * -(-max(n_phi_min, n_lat, n_pole) // 2) if dims == 2 else 0
    return dict(classes=classes, light_lines=ll, singular_lines=sing, kmax=kmax, Kcut=Kcut, h_min=float(h_min),
                rho_max=float(rho_max), poles=poles, clusters=cl, N_phi=n_phi, N_phi_lateral=n_lat,
                N_phi_poles=n_pole, eta_min=eta, phi_scan=phis, dims=dims, K_geom=K_geom, K_cap=K_cap, capped=capped)


def fixed_krho_rule(feat, order=8, rtol=1e-6, alpha=0.5, grade=12, r_rel=0.05, c_h=4.0, c_rho=4.0,
                    pole_safety=0.5):
    """fixed_kρ_rule (mod_KSpaceFeatures.jl:407-433): composite Gauss-Legendre
    (`order` nodes per panel) on [0, Kcut], breaks at every light line, panels
    graded to the smallest of the background width min(c_h / h_min,
    c_rho / rho_max, max(r_rel k, kmax / 40)), max(L 2^-grade, alpha |k - L|)
    at every singular light line L (whose touching panels are integrated in
    u, k = L -+ u^2), and max(w_Gamma, alpha dist(k, [a, b])) at every pole
    cluster, w_Gamma = pole_safety 2 Gamma / eps_B (the Bernstein-ellipse half
    width that keeps the panel's error below rtol).  Returns (nodes, weights, panels)."""
    from . import quadrature as Q
    if r_rel > alpha:
        raise ValueError("fixed_kρ_rule: r_rel must be ≤ α")
    rhoB = (1 / float(rtol)) ** (1 / (2 * order))
    epsB = (rhoB - 1 / rhoB) / 2
    w_flat = min(c_h / feat["h_min"], c_rho / feat["rho_max"] if feat["rho_max"] > 0 else math.inf)
    w_floor = feat["kmax"] / 40
    sing = np.asarray(feat["singular_lines"], np.float64)
    ca = np.asarray([c["a"] for c in feat["clusters"]], np.float64)
    cb = np.asarray([c["b"] for c in feat["clusters"]], np.float64)
    wG = pole_safety * 2 * np.asarray([c["Gamma"] for c in feat["clusters"]], np.float64) / epsB

    def width(k):
        wk = min(w_flat, max(r_rel * k, w_floor))
        if sing.size:
            wk = min(wk, float(np.min(np.maximum(sing * 2.0 ** -grade, alpha * np.abs(k - sing)))))
        if ca.size:
            dist = np.where(k < ca, ca - k, np.where(k > cb, k - cb, 0.0))
            wk = min(wk, float(np.min(np.maximum(wG, alpha * dist))))
        return wk

    breaks = [0.0] + [L for L in feat["light_lines"] if L < feat["Kcut"]] + [feat["Kcut"]]
    panels = Q.graded_panels(breaks, width, alpha=alpha)
    nodes, wts = Q.gl_rule(panels, order=order, singular=feat["singular_lines"])
    return nodes, wts, panels


def adaptive_seeds(feat, rtol=1e-6, n_gamma=4.0, grade=None):
    """adaptive_seeds (mod_KSpaceFeatures.jl:448-460): starting breakpoints for an
    adaptive k_rho grid: light lines, L(1 -+ 2^-j) grading onto the singular
    ones, each cluster's [a - n Gamma, a, b, b + n Gamma], and x2 seeds from
    2 kmax to Kcut."""
    if grade is None:
        grade = min(50, int(math.ceil(2 * math.log2(10 / float(rtol)))))
    s = list(feat["light_lines"])
    for L in feat["singular_lines"]:
        for j in range(1, grade + 1):
            s += [L * (1 - 2.0 ** -j), L * (1 + 2.0 ** -j)]
    for c in feat["clusters"]:
        s += [c["a"] - n_gamma * c["Gamma"], c["a"], c["b"], c["b"] + n_gamma * c["Gamma"]]
    k = 2 * feat["kmax"]
    while k < feat["Kcut"]:
        s.append(k)
        k *= 2
    return sorted(set(x for x in s if 0 < x < feat["Kcut"]))
