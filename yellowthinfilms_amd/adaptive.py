"""The adaptive k_rho grids: h-adaptive Gauss–Kronrod refinement of probe
integrands with every refinement step's new k_rho evaluated as one batch.

`adaptive_kx_grid` follows mod_KxGrid.jl:272-305.  Each probe is refined on its
own by QuadGK.jl's algorithm (the reference calls QuadGK 2.11.3's
`quadgk_segbuf`, a third-party package absent from the reference tree; its
published algorithm, restated here):
  - the seed partition's segments are evaluated with the order-7 Gauss–Kronrod
    rule (15 points; the 7-point Gauss rule is embedded), each segment carrying
    its Kronrod integral I_s and error E_s = |I_s^Kronrod - I_s^Gauss| (2-norm
    over the probe's entries);
  - while max(atol, rtol |I|) < E (I, E the sums over the segments) and fewer
    than maxevals points were taken, the segment of largest E is split at its
    midpoint and both halves evaluated (I and E updated as I - I_s + I_1 + I_2);
  - a probe whose E stays above the tolerance is "stalled".
The probes advance in lock-step: every step pops each active probe's worst
segment and evaluates all their new points in one call of the batched
integrand (`probes.probe_values` on the device for the box grid), memoized per
k_rho as the reference memoizes its probe values.  Each probe sees exactly the
sequence of splits of its own sequential run.  The union of the probes'
accepted panel boundaries then carries the Kronrod rule (`_panels_to_nodes`,
:208-231): one node/weight rule converged for every probe.

`box_krho_grid` is box_kρ_grid (mod_ClosedFormGrid.jl:155-255) on it, and
`box_krho_grid_grouped` (:307-333) runs it on each group of `phi_groups`
(:264-284), the phi-dependent k_rho grid that `build_layer_operator_grouped`
takes.
"""
import heapq
import math

import numpy as np

from .quadrature import dedup, layer_light_lines, decay_cutoff, tail_bound, build_kphi_grid

# Gauss–Kronrod G7-K15 on [-1, 1]: Kronrod nodes x_0 > x_1 > ... > x_7 = 0, the
# Gauss nodes are x_1, x_3, x_5, x_7.
_XK = np.array([0.991455371120812639206854697526329, 0.949107912342758524526189684047851,
                0.864864423359769072789712788640926, 0.741531185599394439863864773280788,
                0.586087235467691130294144845693013, 0.405845151377397166906606412076961,
                0.207784955007898467600689403773245, 0.0])
_WK = np.array([0.022935322010529224963732008058970, 0.063092092629978553290700663189204,
                0.104790010322250183839876322541518, 0.140653259715525918745189590510238,
                0.169004726639267902826583426598550, 0.190350578064785409913256402421014,
                0.204432940075298892414161999234649, 0.209482141084727828012999174891714])
_WG = np.array([0.129484966168869693270611432679082, 0.279705391489276667901467771423780,
                0.381830050505118944950369775488975, 0.417959183673469387755102040816327])  # at x_1, x_3, x_5, x_7


def kronrod_reference_rule(kr_order=7):
    """The full Kronrod rule on [-1, 1]: abscissas xi and weights (15 of each
    for order 7), Sum w h(xi) ~ int_{-1}^{1} h (_kronrod_reference_rule,
    mod_KxGrid.jl:178-197)."""
    if kr_order != 7:
        raise ValueError("adaptive_kx_grid: kr_order 7 (the G7-K15 rule) is the one supported")
    xi = np.concatenate([[0.0], _XK[:-1], -_XK[:-1]])
    w = np.concatenate([[_WK[-1]], _WK[:-1], _WK[:-1]])
    return xi, w


def _points_many(a, b):
    """The 15 points of each segment [a_i, b_i] (S, 15) in QuadGK's evaluation
    form a + (1 -+ x) s: [centre, a + (1 + x_i) s (i < 7), a + (1 - x_i) s (i < 7)]."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    s = 0.5 * (b - a)
    x = _XK[:-1]
    return np.concatenate([(a + s)[:, None], a[:, None] + (1 + x) * s[:, None], a[:, None] + (1 - x) * s[:, None]],
                          axis=1)


def _points(a, b):
    return _points_many([a], [b])[0]


def _evalrule_many(a, b, f):
    """(I (S, D), E (S,)) of the segments [a_i, b_i] from their 15 values each,
    f (S, 15, D) at _points_many(a, b): Kronrod integral and 2-norm of its
    difference from the embedded Gauss one."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    s = 0.5 * (b - a)
    pair = f[:, 1:8] + f[:, 8:15]
    Ik = (f[:, 0] * _WK[-1] + (pair * _WK[None, :-1, None]).sum(axis=1)) * s[:, None]
    Ig = (f[:, 0] * _WG[-1] + (pair[:, [1, 3, 5]] * _WG[None, :-1, None]).sum(axis=1)) * s[:, None]
    d = Ik - Ig
    E = np.sqrt((d.real ** 2 + d.imag ** 2).sum(axis=1)) if np.iscomplexobj(d) else np.sqrt((d * d).sum(axis=1))
    bad = ~np.isfinite(E)
    if bad.any():
        j = int(np.argmax(bad))
        raise FloatingPointError(f"adaptive_kx_grid: integrand produced {E[j]} in the interval ({a[j]}, {b[j]})")
    return Ik, E


def _evalrule(a, b, f):
    """(I, E) of segment [a, b] from the 15 values f (15, D) at _points(a, b)."""
    I, E = _evalrule_many([a], [b], np.asarray(f)[None])
    return I[0], float(E[0])


def _norm(v):
    return math.sqrt(float(np.vdot(v, v).real))


class _Memo:
    """Probe values per k_rho (the reference's `memo`), filled in batches and
    held in one growing array (rows in evaluation order)."""

    def __init__(self, F, n_probes):
        self.F, self.P = F, n_probes
        self.row = {}
        self.store = None
        self.n = 0

    def fill(self, ks):
        new = sorted({k for k in np.asarray(ks, np.float64).ravel().tolist() if k not in self.row})
        if new:
            v = np.asarray(self.F(np.asarray(new, np.float64))).reshape(len(new), self.P, -1)
            if self.store is None:
                self.store = np.empty((max(1024, 2 * len(new)),) + v.shape[1:], v.dtype)
            if self.store.dtype != np.result_type(self.store, v):
                self.store = self.store.astype(np.result_type(self.store, v))
            if self.n + len(new) > self.store.shape[0]:
                grown = np.empty((2 * (self.n + len(new)),) + self.store.shape[1:], self.store.dtype)
                grown[:self.n] = self.store[:self.n]
                self.store = grown
            self.store[self.n:self.n + len(new)] = v
            for j, k in enumerate(new):
                self.row[k] = self.n + j
            self.n += len(new)
        return len(new)

    def has(self, ks):
        return all(k in self.row for k in np.asarray(ks, np.float64).ravel().tolist())

    def rows(self, ks):
        ks = np.asarray(ks, np.float64)
        return np.fromiter((self.row[k] for k in ks.ravel().tolist()), np.int64, count=ks.size).reshape(ks.shape)


def adaptive_kx_grid(F, n_probes, kbreaks, kcut, rtol=1e-6, atol=0.0, kr_order=7, maxevals=10_000, speculate=True):
    """adaptive_kx_grid (mod_KxGrid.jl:272-305) with batched probes.
    F(k) -> (len(k), n_probes, ...) values of every probe integrand at the
    k_rho values k.  Returns (nodes, weights, info) with info: n_panels,
    n_nodes, per_probe ({E, stalled} per probe), any_stalled,
    min_panel_width, kcut, panels, n_krho_evaluated, krho_visited (the
    distinct k_rho evaluated, sorted), n_krho_used (those the rule
    evaluations read), n_batches (refinement steps + 1), n_calls (calls of
    F).  speculate: each call of F also takes the points of every split a
    probe can make next (see the loop), which halves the calls and leaves
    the refinement unchanged."""
    if n_probes < 1:
        raise ValueError("adaptive_kx_grid: fs cannot be empty")
    kc = float(kcut)
    if not kc > 0:
        raise ValueError(f"adaptive_kx_grid: kcut must be positive (got {kcut})")
    xi, wt = kronrod_reference_rule(kr_order)
    P = n_probes
    pts = sorted(set([0.0, kc] + [float(k) for k in kbreaks if 0 < float(k) < kc]))
    A, B = np.array(pts[:-1]), np.array(pts[1:])
    S = A.size
    X = _points_many(A, B)
    memo = _Memo(F, P)
    memo.fill(X)
    n_batches = n_calls = 1
    used = set(X.ravel().tolist())
    V = memo.store[memo.rows(X)]                                  # (S, 15, P, D)
    D = V.shape[-1]
    I0, E0 = _evalrule_many(np.tile(A, P), np.tile(B, P), V.transpose(2, 0, 1, 3).reshape(P * S, 15, D))
    I0, E0 = I0.reshape(P, S, D), E0.reshape(P, S)
    runs = []
    for p in range(P):
        segs = [(pts[j], pts[j + 1], I0[p, j], float(E0[p, j])) for j in range(S)]
        heap = [(-sg[3], j, sg) for j, sg in enumerate(segs)]
        heapq.heapify(heap)
        # QuadGK returns the seed evaluation unrefined when it already took maxevals points
        runs.append({"heap": heap, "I": I0[p].sum(axis=0), "E": float(E0[p].sum()), "n": 15 * S, "seq": S,
                     "active": 15 * S < maxevals})
    while True:
        todo = []
        for p, r in enumerate(runs):
            if not r["active"]:
                continue
            tol = max(atol, rtol * _norm(r["I"]))
            if not tol < r["E"] or r["n"] > maxevals:
                r["active"] = False
                continue
            _, _, sg = heapq.heappop(r["heap"])
            todo.append((p, r, sg, (sg[0] + sg[1]) / 2))
        if not todo:
            break
        a2 = np.array([(t[2][0], t[3]) for t in todo]).ravel()        # left and right halves, interleaved
        b2 = np.array([(t[3], t[2][1]) for t in todo]).ravel()
        X = _points_many(a2, b2)
        if not memo.has(X):
            # Speculation: a probe's next split is one of the two new halves or its heap top, so
            # their halves' points go into the same batch (memo entries only: the splits made
            # are unchanged, and every other step then needs no device call).
            sa, sb = [], []
            for _, r, sg, mid in todo:
                for a, b in ((sg[0], mid), (mid, sg[1])):
                    m = (a + b) / 2
                    sa += [a, m]
                    sb += [m, b]
                if r["heap"]:
                    t = r["heap"][0][2]
                    m = (t[0] + t[1]) / 2
                    sa += [t[0], m]
                    sb += [m, t[1]]
            memo.fill(np.concatenate([X.ravel(), _points_many(sa, sb).ravel()]) if speculate else X)
            n_calls += 1
        used.update(X.ravel().tolist())
        n_batches += 1
        pidx = np.repeat(np.array([t[0] for t in todo]), 2)
        I2, E2 = _evalrule_many(a2, b2, memo.store[memo.rows(X), pidx[:, None]])
        for j, (p, r, sg, mid) in enumerate(todo):
            h1 = (sg[0], mid, I2[2 * j], float(E2[2 * j]))
            h2 = (mid, sg[1], I2[2 * j + 1], float(E2[2 * j + 1]))
            r["I"] = (r["I"] - sg[2]) + h1[2] + h2[2]
            r["E"] = (r["E"] - sg[3]) + h1[3] + h2[3]
            r["n"] += 30
            for t in (h1, h2):
                heapq.heappush(r["heap"], (-t[3], r["seq"], t))
                r["seq"] += 1
    breaks = set(pts)
    per_probe = []
    for r in runs:
        segs = [h[2] for h in r["heap"]]
        I = sum(s[2] for s in segs)  # re-summed over the final segments
        E = sum(s[3] for s in segs)
        per_probe.append({"E": E, "stalled": bool(E > max(atol, rtol * _norm(I)))})
        for s in segs:
            breaks.add(s[0])
            breaks.add(s[1])
    bp = sorted(breaks)
    panels = list(zip(bp[:-1], bp[1:]))
    a = np.array([p[0] for p in panels])
    b = np.array([p[1] for p in panels])
    h, c = 0.5 * (b - a), 0.5 * (a + b)
    nodes = (c[:, None] + h[:, None] * xi[None, :]).ravel()
    weights = (h[:, None] * wt[None, :]).ravel()
    info = {"n_panels": len(panels), "n_nodes": nodes.size, "per_probe": per_probe,
            "any_stalled": any(q["stalled"] for q in per_probe),
            "min_panel_width": float((b - a).min()) if panels else math.inf, "kcut": kc, "panels": panels,
            "n_krho_evaluated": len(memo.row), "krho_visited": np.array(sorted(memo.row)), "n_batches": n_batches,
            "n_krho_used": len(used), "n_calls": n_calls}
    return nodes, weights, info


def branch_points(stack, omega, phis=(0.0, math.pi / 2)):
    """branch_points (mod_KxGrid.jl:74-80): the de-duplicated light lines of
    every layer at every direction in phis."""
    ks = []
    for lay in stack:
        for phi in phis:
            ks += layer_light_lines(lay, omega, phi)
    return dedup(sorted(ks))


def box_tail_cutoff(delta, rtol):
    """box_tail_cutoff (mod_ClosedFormGrid.jl:97-103): the smallest K with
    log(K delta)/(K delta) <= rtol, by the fixed-point iteration
    K = max(log(K delta), 1)/(rtol delta) from K = 10/(rtol delta)."""
    K = 10.0 / (rtol * delta)
    for _ in range(50):
        K = max(math.log(K * delta), 1.0) / (rtol * delta)
    return K


def is_lossless(stack):
    """is_lossless (mod_ClosedFormGrid.jl:111-113): no imaginary part in any
    eps (or mu) entry of any layer."""
    for lay in stack:
        for m in (lay.eps.eps, lay.mu.eps):
            if np.any(np.imag(np.asarray(m, np.complex128)) != 0):
                return False
    return True


def box_krho_grid(stack, omega, source_layer, box, zbar=None, n_phi=64, phi=None, lateral_dir=0.0, rho_far=0.0,
                  rtol=1e-4, atol=0.0, kr_order=7, maxevals=10_000, krho_cap=2000.0, target_decay=25.0,
                  include_interior=True, on_stall="error", scan_poles=True, n_phi_scan=32, device=0):
    """box_kρ_grid (mod_ClosedFormGrid.jl:155-255): the k_rho axis of a 2D
    (k_rho, phi) grid for a box source, converged on the box probes
    (`probes.box_probes`) with every refinement step's k_rho evaluated as one
    batch on the device (`probes.probe_values`).  Returns (krho, w_krho, phi,
    w_phi, info); info adds the probes, h_min, tail_bound, the number of
    distinct k_rho solves and the features to adaptive_kx_grid's."""
    from .kspace import kspace_features, adaptive_seeds
    from .probes import box_probes, probe_values, thinnest_finite_layer
    if on_stall not in ("error", "warn"):
        raise ValueError(f"box_kρ_grid: on_stall={on_stall} must be :error or :warn")
    om = float(omega)
    if phi is None:
        phis, wph = build_kphi_grid(n_phi)
    else:
        phis, wph = np.asarray(phi[0], np.float64), np.asarray(phi[1], np.float64)
    probes = box_probes(stack, source_layer, box, zbar=zbar, lateral_dir=lateral_dir, rho_far=rho_far,
                        include_interior=include_interior)
    dx, dy, dz = (float(v) for v in box)
    hs = [p[4] for p in probes]
    h_min = max(min([h for h in hs if h > 0], default=dz / 2), 1e-6 * thinnest_finite_layer(stack))
    kmax = max(branch_points(stack, om, phis=sorted(set(np.mod(phis, math.pi).tolist()))) + [om])
    Kbox = box_tail_cutoff(min(dx, dy, dz), rtol)
    Kcut = tail_bound(max(min(decay_cutoff(h, target_decay), Kbox) if h > 0 else Kbox for h in hs), kmax,
                      krho_cap=krho_cap)
    rho_probe = max(math.hypot(p[2], p[3]) for p in probes)
    feat = kspace_features(stack, om, source_layer, h_min, rho_max=rho_probe, dims=2, Kcut=Kcut,
                           phi_scan=(phis if len(phis) <= n_phi_scan else None), n_phi_scan=n_phi_scan,
                           scan=scan_poles, device=device)
    grade = min(30, min(50, int(math.ceil(2 * math.log2(10 / float(rtol))))))
    seeds = adaptive_seeds(feat, rtol=rtol, grade=grade)

    def F(ks):
        return probe_values(stack, om, source_layer, ks, phis, wph, probes, box, device=device).reshape(len(ks),
                                                                                                           len(probes), 9)

    kr, wr, info = adaptive_kx_grid(F, len(probes), seeds, Kcut, rtol=float(rtol), atol=float(atol),
                                    kr_order=kr_order, maxevals=maxevals)
    info.update({"probes": [(p[0], p[1], p[2], p[3], p[4], p[5][1], p[5][0]) for p in probes], "h_min": h_min,
                 "tail_bound": Kcut, "n_krho_solves": info["n_krho_used"], "features": feat})
    stalled = [i for i, q in enumerate(info["per_probe"]) if q["stalled"]]
    if on_stall == "error" and stalled:
        lst = "\n".join(f"  layer {probes[i][0]}, z̄ = {probes[i][5][0]}, {probes[i][5][1]} (error estimate "
                        f"{info['per_probe'][i]['E']:.2g})" for i in stalled)
        raise RuntimeError(f"box_kρ_grid: {len(stalled)}/{len(probes)} probes' finite bands did not reach rtol={rtol} "
                           f"within maxevals={maxevals}:\n{lst}\n"
                           + ("The stack is lossless: real-axis guided-mode poles cannot be integrated; add a small "
                              "loss. " if is_lossless(stack) else "Likely near-real poles (low loss) or an "
                              "under-resolved probe. ") + "Pass on_stall=:warn to get the unconverged grid anyway.")
    return kr, wr, phis, wph, info


def phi_groups(stack, n_phi, classes_per_group=1):
    """phi_groups (mod_ClosedFormGrid.jl:264-284): a partition of the uniform
    phi grid's indices 0 .. n_phi-1 into groups that can share a k_rho grid.
    Azimuthally symmetric stack: one group.  Diagonal eps and mu in every
    layer: the mirror classes {phi, -phi, pi - phi, pi + phi} (the orbit of
    index p under p -> -p and, for even n_phi, p -> n_phi/2 - p), in order of
    their smallest phi.  Otherwise one phi per group.  classes_per_group
    merges that many neighbouring classes into one group."""
    from .kspace import _is_azimuthally_symmetric
    if _is_azimuthally_symmetric(stack):
        return [list(range(n_phi))]

    def merge(cl):
        return [sorted(i for c in cl[r:r + classes_per_group] for i in c) for r in range(0, len(cl), classes_per_group)]

    def offdiag_zero(m):
        m = np.asarray(m, np.complex128)
        return not np.any(m - np.diag(np.diag(m)))
    if not all(offdiag_zero(l.eps.eps) and offdiag_zero(l.mu.eps) for l in stack):
        return merge([[p] for p in range(n_phi)])
    seen = [False] * n_phi
    groups = []
    for p in range(n_phi):
        if seen[p]:
            continue
        orbit = {p, (-p) % n_phi}
        if n_phi % 2 == 0:
            orbit |= {(n_phi // 2 - q) % n_phi for q in list(orbit)}
        g = sorted(orbit)
        for q in g:
            seen[q] = True
        groups.append(g)
    return merge(groups)


def box_krho_grid_grouped(stack, omega, source_layer, box, n_phi=64, classes_per_group=1, **kwargs):
    """box_kρ_grid_grouped (mod_ClosedFormGrid.jl:307-333): box_krho_grid on
    each group of phi_groups(stack, n_phi) with the group's own phi nodes and
    weights, the lateral probes along the group's first phi (along x for a
    single group).  Returns (groups, info): groups[g] = {krho, wrho, phi, wphi}
    (together they partition the phi grid, so a quantity linear in the
    quadrature sum is the sum over the groups); info: n_groups, n_nodes
    (sum of k_rho x phi), nkrho per group, any_stalled, per_group.  The groups
    run one after another (each one's refinement batches are device work)."""
    phis, wph = build_kphi_grid(n_phi)
    idxs = phi_groups(stack, n_phi, classes_per_group=classes_per_group)
    groups, infos = [], []
    kwargs = dict(kwargs)
    kwargs.pop("lateral_dir", None)
    for idx in idxs:
        kr, wr, pg, wg, info = box_krho_grid(stack, omega, source_layer, box, phi=(phis[idx], wph[idx]),
                                             lateral_dir=(0.0 if len(idx) == n_phi else float(phis[idx[0]])),
                                             **kwargs)
        groups.append({"krho": kr, "wrho": wr, "phi": pg, "wphi": wg})
        infos.append(info)
    info = {"n_groups": len(groups), "n_nodes": sum(g["krho"].size * g["phi"].size for g in groups),
            "nkrho": [g["krho"].size for g in groups], "any_stalled": any(i["any_stalled"] for i in infos),
            "per_group": infos}
    return groups, info


def kronrod_rule(panels, kr_order=7):
    """kronrod_rule (mod_KxGrid.jl:313-318): the G7-K15 Kronrod rule on each
    panel (the rule adaptive_kx_grid returns, rebuilt from its panels)."""
    xi, wt = kronrod_reference_rule(kr_order)
    a = np.array([float(p[0]) for p in panels])
    b = np.array([float(p[1]) for p in panels])
    h, c = 0.5 * (b - a), 0.5 * (a + b)
    return (c[:, None] + h[:, None] * xi[None, :]).ravel(), (h[:, None] * wt[None, :]).ravel()


def _rel_change(v, v2):
    """_rel_change (mod_ClosedFormGrid.jl:489-494): per probe |v - v2| /
    max(|v2|, 1e-6 x the largest |v2|), and its maximum."""
    n2 = np.linalg.norm(v2, axis=1)
    scale = float(n2.max()) if n2.size else 0.0
    rel = np.linalg.norm(v - v2, axis=1) / np.maximum(n2, 1e-6 * scale)
    return {"rel_change": float(rel.max()) if rel.size else 0.0, "per_probe": rel}


def point_kx_grid(stack, omega, source_layer, rho_max=0.0, pz=None, z_fracs=(0.1, 0.3, 0.5, 0.7, 0.9), engine="auto",
                  rtol=1e-6, atol=0.0, kr_order=7, maxevals=10_000, krho_cap=None, krho_cap_rel=1e5,
                  target_decay=25.0, check=False, scan_poles=True, rule=None, h_min=None, device=0):
    """point_kx_grid (mod_ClosedFormGrid.jl:381-424): the layer-level radial
    rule of a point source in source_layer (the 1D path), reused for every
    depth, moment and lateral offset.  The band is cut at target_decay /
    (pz/2) (pz default thinnest_finite_layer / 20; h_min overrides pz/2).
    Engine (choose_engine): "fixed" -> fixed_krho_rule on the features, no
    probes; "adaptive" -> adaptive_kx_grid on the point probes
    (probes.point_probes), every refinement step's new k_x one device batch
    (probes.point_probe_values).  check: the probes on the rule and on the
    rule with every panel halved, info["check"] = _rel_change of the two.
    Returns (kx, wt, info)."""
    from .kspace import kspace_features, adaptive_seeds, fixed_krho_rule, choose_engine
    from .probes import thinnest_finite_layer, check_depths, point_probes, point_probe_values
    from .quadrature import gl_rule, halve_panels
    N = len(stack)
    if not 1 <= source_layer < N:
        raise ValueError(f"point_kx_grid: source_layer={source_layer} must be in 1:{N - 1}")
    om = float(omega)
    pzf = thinnest_finite_layer(stack) / 20 if pz is None else float(pz)
    dist = pzf / 2 if h_min is None else float(h_min)
    d = float(stack[source_layer - 1].d)
    zf_probe = list(z_fracs) if h_min is None else [z / d for z in check_depths(stack, source_layer, z_fracs, dist)]
    probes = point_probes(stack, source_layer, zf_probe, dist, float(rho_max))
    feat = kspace_features(stack, om, source_layer, dist, rho_max=rho_max, dims=1, target_decay=target_decay,
                           krho_cap=krho_cap, krho_cap_rel=krho_cap_rel, scan=scan_poles, device=device)
    eng = choose_engine(stack, engine)
    rule = dict(rule or {})
    if eng == "fixed":
        kx, wt, panels = fixed_krho_rule(feat, rtol=rtol, **rule)
        info = {"engine": "fixed", "n_nodes": kx.size, "n_panels": len(panels), "panels": panels, "any_stalled": False}
        refined = lambda: gl_rule(halve_panels(panels), order=rule.get("order", 8), singular=feat["singular_lines"])
    else:
        def F(ks):
            return point_probe_values(stack, om, source_layer, ks, probes, device=device)
        kx, wt, info = adaptive_kx_grid(F, len(probes), adaptive_seeds(feat, rtol=rtol), feat["Kcut"], rtol=float(rtol),
                                        atol=float(atol), kr_order=kr_order, maxevals=maxevals)
        info.update({"engine": "adaptive", "n_kx_solves": info["n_krho_used"]})
        refined = lambda: kronrod_rule(halve_panels(info["panels"]), kr_order=kr_order)
    chk = None
    if check:
        v = point_probe_values(stack, om, source_layer, kx, probes, wt=wt, device=device).sum(axis=0)
        k2, w2 = refined()
        v2 = point_probe_values(stack, om, source_layer, k2, probes, wt=w2, device=device).sum(axis=0)
        chk = _rel_change(v, v2)
    info.update({"n_probes": len(probes), "tail_bound": feat["Kcut"], "z_fracs": list(z_fracs),
                 "rho_max": rho_max, "features": feat, "check": chk})
    return kx, wt, info


def chunked_2d(f, stack, omega, source_layer, krho, wrho, phi, wphi, obs_layers=None, max_nodes=1_500_000, device=0):
    """chunked_2d (mod_ClosedFormGrid.jl:457-467): the sum over consecutive
    chunks of the k_rho rule (at most max_nodes k_rho x phi nodes each) of
    f(cache) on each chunk's 2D cache -- exact for anything linear in the
    quadrature sum, with one chunk's nodes resident at a time."""
    from .green_operator import CavityTransfer2D
    kr = np.asarray(krho, np.float64)
    wr = np.asarray(wrho, np.float64)
    step = max(1, max_nodes // max(1, len(phi)))
    acc = None
    for i in range(0, kr.size, step):
        c = CavityTransfer2D(stack, omega, source_layer, kr[i:i + step], wr[i:i + step], phi, wphi,
                             obs_layers=obs_layers, device=device)
        v = f(c)
        acc = v if acc is None else acc + v
        del c
    return acc


def point_krho_grid_2d(stack, omega, source_layer, rho_max, pz=None, z_fracs=(0.1, 0.5, 0.9), n_phi=None, rtol=1e-6,
                       krho_cap=None, krho_cap_rel=1e5, target_decay=25.0, c_phi=2.5, n_phi_scan=32, check=False,
                       scan_poles=True, rule=None, h_min=None, terms="all", max_nodes=1_500_000, device=0):
    """point_kρ_grid_2d (mod_ClosedFormGrid.jl:536-585): the (k_rho, phi)
    tensor grid of a point source in any stack: fixed_krho_rule on the 2D
    features (light lines at every scanned phi, the poles along their tracks,
    cut-off target_decay / h_min) and a uniform phi rule of n_phi nodes
    (default the features' N_phi).  check: closed_form_G at probe points
    (per depth: +-dist in the layer, dist into each neighbour; the source
    plane too for terms="scattered"; lateral offsets 0 and, with rho_max > 0,
    (+-rho_max, 0), (0, rho_max)) on the rule, on the rule with every panel
    halved (info["check"]["krho"]) and with 2 n_phi nodes (the mean of the
    rule and the rule shifted by pi / n_phi: info["check"]["phi"]), each
    summed over chunks of at most max_nodes nodes.  Returns (krho, wrho, phi,
    wphi, info)."""
    from .green_operator import closed_form_G_2d
    from .kspace import kspace_features, fixed_krho_rule
    from .probes import thinnest_finite_layer, check_depths
    from .quadrature import gl_rule, halve_panels
    N = len(stack)
    if not 1 <= source_layer < N:
        raise ValueError(f"point_kρ_grid_2d: source_layer={source_layer} must be in 1:{N - 1}")
    om = float(omega)
    pzf = thinnest_finite_layer(stack) / 20 if pz is None else float(pz)
    dist = pzf / 2 if h_min is None else float(h_min)
    feat = kspace_features(stack, om, source_layer, dist, rho_max=rho_max, dims=2, n_phi_scan=n_phi_scan,
                           target_decay=target_decay, krho_cap=krho_cap, krho_cap_rel=krho_cap_rel, scan=scan_poles,
                           c_phi=c_phi, rtol=rtol, device=device)
    rule = dict(rule or {})
    kr, wr, panels = fixed_krho_rule(feat, rtol=rtol, **rule)
    nphi = feat["N_phi"] if n_phi is None else int(n_phi)
    phis, wph = build_kphi_grid(nphi)
    chk = None
    if check:
        d = float(stack[source_layer - 1].d)
        zs = check_depths(stack, source_layer, z_fracs, dist / 2)
        pts = []                                           # (obs, z_obs in the obs layer's frame, z_src local)
        for z in zs:
            zp = d - z if source_layer == 1 else z
            if zp + dist <= d:
                pts.append((source_layer, zp + dist, z))
            if zp - dist >= 0:
                pts.append((source_layer, zp - dist, z))
            if source_layer > 1:
                pts.append((source_layer - 1, float(stack[source_layer - 2].d) - dist, z))
            pts.append((source_layer + 1, min(dist, float(stack[source_layer].d)), z))
        if terms == "scattered":
            for z in zs:
                pts.append((source_layer, d - z if source_layer == 1 else z, z))
        lat = [(0.0, 0.0), (rho_max, 0.0), (-rho_max, 0.0), (0.0, rho_max)] if rho_max > 0 else [(0.0, 0.0)]
        obs_layers = sorted({p[0] for p in pts})

        def f(c):
            return np.stack([closed_form_G_2d(c, o, z, [zo], [x], [y], terms=(terms if o == source_layer else "all"))
                             [:, :, 0, 0, 0].cpu().numpy().T.ravel() for (o, zo, z) in pts for (x, y) in lat])

        def run(kn, kw, phv):
            return chunked_2d(f, stack, om, source_layer, kn, kw, phv, wph, obs_layers=obs_layers,
                              max_nodes=max_nodes, device=device)
        v = run(kr, wr, phis)
        k2, w2 = gl_rule(halve_panels(panels), order=rule.get("order", 8), singular=feat["singular_lines"])
        v2phi = 0.5 * (v + run(kr, wr, phis + math.pi / nphi))
        chk = {"krho": _rel_change(v, run(k2, w2, phis)), "phi": _rel_change(v, v2phi)}
    info = {"engine": "fixed", "n_krho": kr.size, "N_phi": nphi, "n_nodes": kr.size * nphi, "n_panels": len(panels),
            "panels": panels, "tail_bound": feat["Kcut"], "features": feat, "check": chk}
    return kr, wr, phis, wph, info
