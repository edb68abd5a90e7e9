"""Closed forms that pin the oracle's anisotropic and layered-media paths without
going through its own formulation (Berreman matrix, eigen-solve, S-matrix).

Each check_* raises AssertionError when the oracle disagrees.  The same checks
run against the default oracle (tests/test_oracle_closed_forms.py) and against
builds of the oracle with one planted mistake each (tests/test_oracle_mutants.py),
which must fail them.

Conventions (units eps0 = mu0 = c = 1, time e^{-i w t}):
- the Berreman state is psi = (Ex, Ey, Hx, Hy) (mod_ModeSolver.jl:122-193; Ez from
  mod_CavityTransfer.jl:87-130), so every eigenpair (Q, psi) is a plane wave
  exp(i(kx x + ky y + Q z)) that solves the full Maxwell system;
- the reference's z runs from layer 1 down (comparisons/gila_airy/plane_wave_slice.jl:16-18):
  against the physical z-up frame, entries of a dyadic with one z index change sign;
- in layer 1 a source z is its height above the layer's bottom interface (_zpos,
  mod_ClosedFormField.jl:72-73), while an observation z is in the cavity frame, from the
  layer's top reference plane down to the interface at z = d (_obs_phases, :326-335):
  height = d - z;
- g~(kx, ky) is the 2D Fourier density of the field of a unit dipole,
  E(r) = (2 pi)^-2 int g~ exp(i(kx x + ky y)) dkx dky, as the oracle's lateral sum
  weights it (mod_ClosedFormField.jl:723-727).
"""
import json
import os

import numpy as np
from scipy.special import j0, j1

from helpers import (po, relerr, ostack, lay, NM, IsotropicMaterial, UniaxialMaterial, GyrotropicMaterial,
                     FlatLayer, kr_rule, E_free, c1_stack, c2_stack)

W0 = 2 * np.pi
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FLIP = np.diag([1.0, 1.0, -1.0])


def kz_branch(kz2):
    """sqrt with Im >= 0, the outgoing / decaying branch (_kz_branch,
    mod_Dipz_SemiAnalytical.jl:94-97)."""
    kz = np.sqrt(complex(kz2))
    return -kz if kz.imag < 0 else kz


def tilted_uniaxial(eo, ee, theta):
    """diag(eo, eo, ee) with its axis turned by theta about y: eps_xz != 0."""
    c, s = np.cos(theta), np.sin(theta)
    R = np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])
    return R @ np.diag([eo, eo, ee]).astype(complex) @ R.T


def passive_tensor(herm, v, loss):
    """A Hermitian part plus i times a positive-definite loss (loss I + v v^H): a
    passive, non-reciprocal tensor with every entry non-zero."""
    A = np.asarray(herm, dtype=complex)
    assert np.allclose(A, A.conj().T)
    v = np.asarray(v, dtype=complex)
    return A + 1j * (loss * np.eye(3) + 0.1 * np.outer(v, v.conj()))


def gyro_tensor(exx, g, ezz):
    return np.array([[exx, 1j * g, 0], [-1j * g, exx, 0], [0, 0, ezz]], dtype=complex)


# ---------------------------------------------------------------------------
# 1. Every Berreman eigenpair is a bulk plane-wave solution of Maxwell's equations
# ---------------------------------------------------------------------------
def maxwell_residual(eps, mu, omega, kx, ky, Q, psi):
    """|K x E - w mu H| + |K x H + w eps E| for the plane wave (Q, psi), with Ez and
    Hz taken from the z rows of the two curl equations."""
    Ex, Ey, Hx, Hy = psi
    K = np.array([kx, ky, Q])
    Ez = (-(kx * Hy - ky * Hx) / omega - eps[2, 0] * Ex - eps[2, 1] * Ey) / eps[2, 2]
    E = np.array([Ex, Ey, Ez])
    Hz = (np.cross(K, E)[2] / omega - mu[2, 0] * Hx - mu[2, 1] * Hy) / mu[2, 2]
    H = np.array([Hx, Hy, Hz])
    r = np.linalg.norm(np.cross(K, E) - omega * mu @ H) + np.linalg.norm(np.cross(K, H) + omega * eps @ E)
    scale = (np.linalg.norm(K) + omega * (np.linalg.norm(eps) + np.linalg.norm(mu))) * (
        np.linalg.norm(E) + np.linalg.norm(H))
    return r / scale


def check_bulk_maxwell_null_vectors(eps, mu=None, kts=((0.0, 0.0), (3.1, -1.7), (0.4, 9.0), (-25.0, 12.0)),
                                    tol=1e-12):
    """(mod_ModeSolver.jl:122-193, 274-326) The eigenpairs of the Berreman matrix,
    and the sorted forward/backward modes of _modes_layer, are plane waves of the
    bulk medium: (w^2 eps + K x K x) E = 0 and its magnetic partner."""
    eps = np.asarray(eps, dtype=complex)
    mu = np.eye(3, dtype=complex) if mu is None else np.asarray(mu, dtype=complex)
    mk = 0 if np.allclose(mu, mu[0, 0] * np.eye(3)) else 1
    worst = 0.0
    for kx, ky in kts:
        Q, V = po.eigen4(po.berreman(eps, mu, W0, kx, ky))
        for j in range(4):
            worst = max(worst, maxwell_residual(eps, mu, W0, kx, ky, Q[j], V[:, j]))
        _, _, M = po.modes_layer(eps, mu, 3, mk, W0, kx, ky)
        for Vs, Qs in ((M["Vf"], M["Qf"]), (M["Vb"], M["Qb"])):
            for j in range(2):
                worst = max(worst, maxwell_residual(eps, mu, W0, kx, ky, Qs[j], Vs[:, j]))
    assert worst < tol, f"Maxwell residual of the Berreman eigenpairs {worst:.2e}"
    return worst


# ---------------------------------------------------------------------------
# 2. Gyrotropic layer at normal incidence: circular modes of fixed handedness
# ---------------------------------------------------------------------------
def check_gyro_normal_incidence(exx=3 + 0.02j, g=0.3, ezz=2.8 + 0.02j, tol=1e-12):
    """eps = [[exx, i g, 0], [-i g, exx, 0], [0, 0, ezz]] at k_t = 0: the transverse
    wave equation w^2 eps_t E_t = Q^2 E_t gives Q^2 = w^2 (exx + g) with
    (Ex, Ey) ~ (1, -i) and Q^2 = w^2 (exx - g) with (Ex, Ey) ~ (1, +i), for both
    directions.  eps -> eps^T (g -> -g) exchanges the two."""
    eps = gyro_tensor(exx, g, ezz)
    _, _, M = po.modes_layer(eps, np.eye(3), 3, 0, W0, 0.0, 0.0)
    seen = {+1: 0, -1: 0}
    for Vs, Qs in ((M["Vf"], M["Qf"]), (M["Vb"], M["Qb"])):
        for j in range(2):
            q2 = Qs[j] ** 2 / W0 ** 2
            sgn = +1 if abs(q2 - (exx + g)) < abs(q2 - (exx - g)) else -1
            assert abs(q2 - (exx + sgn * g)) < tol * abs(exx), (q2, exx, g)
            Ex, Ey = Vs[0, j], Vs[1, j]
            assert abs(Ey - (-1j * sgn) * Ex) < tol * np.hypot(abs(Ex), abs(Ey)), (sgn, Ex, Ey)
            seen[sgn] += 1
    assert seen == {+1: 2, -1: 2}
    return True


# ---------------------------------------------------------------------------
# 3. The point-source seed: its particular solution is the bulk Maxwell solve
# ---------------------------------------------------------------------------
def check_particular_solution(eps, kts=((0.7, 0.2), (3.1, -1.4), (9.0, 4.0), (25.0, -7.0)), tol=1e-11):
    """Source_projection_Berreman_xyz (mod_SourceProj.jl:34-60) decomposes the jump
    of psi across a point dipole into modes (sf, sib).  For a unit moment spread
    over a slab of unit thickness the interior field has the constant part
    psi_p = -V (iQ)^-1 [sif; sib] plus the local Ez -p_z / eps_zz
    (ez_local_density, mod_ClosedFormField.jl:861-862).  With no z dependence
    (k_z = 0) that field is the bulk solution of
        (i/w)(K K^T - K.K I) E + i w eps E = J,  J = -i w p,  K = (kx, ky, 0),
    H = K x E / w (Gz_ClosedForm_Box_Validation.jl:147-186, [I3]).  The bulk
    solve never sees the Berreman source vector, so the eps_tz term of J_Ht is
    pinned only by this."""
    eps = np.asarray(eps, dtype=complex)
    stack = [lay(IsotropicMaterial(1.0 + 0.001j), 1.0), lay(GyrotropicMaterial(eps), 0.8),
             lay(IsotropicMaterial(1.0 + 0.001j), 1.0)]
    st = ostack(stack)
    worst = 0.0
    for kx, ky in kts:
        c = po.nodes_batch(st, W0, [kx], [ky], 2, [2])
        M = po.stack_smatrix(st, W0, kx, ky)[0][1]
        V = np.hstack([M["Vf"], M["Vb"]])
        Q = np.concatenate([c.Qf[0], c.Qb[0]])
        s = np.vstack([c.sf[0], -c.sb[0]])  # the cache stores sb = -sib (mod_CavityTransfer.jl:243)
        psi = -V @ (s / (1j * Q)[:, None])
        K = np.array([kx, ky, 0.0])
        A = (1j / W0) * (np.outer(K, K) - (K @ K) * np.eye(3)) + 1j * W0 * eps
        Eb = np.linalg.solve(A, -1j * W0 * np.eye(3))
        Hb = np.array([np.cross(K, Eb[:, q]) / W0 for q in range(3)]).T
        Ez = -(eps[2, 0] * psi[0] + eps[2, 1] * psi[1] + (kx * psi[3] - ky * psi[2]) / W0) / eps[2, 2]
        Ez = Ez + np.array([0, 0, -1.0 / eps[2, 2]])
        got = np.vstack([psi[0], psi[1], Ez, psi[2], psi[3]])
        want = np.vstack([Eb, Hb[:2]])
        worst = max(worst, relerr(got, want))
    assert worst < tol, f"seed particular solution vs bulk Maxwell {worst:.2e}"
    return worst


# ---------------------------------------------------------------------------
# 4. Source and observation above a stack of z-uniaxial layers: the dyadic in
#    closed form from the TE and TM Airy reflection coefficients
# ---------------------------------------------------------------------------
def uniaxial_layers(stack):
    """(eps_t, eps_z, d) per layer of a stack of isotropic / z-uniaxial layers."""
    out = []
    for l in stack:
        e = l.eps.eps
        assert np.allclose(e, np.diag(np.diag(e))) and e[0, 0] == e[1, 1]
        assert np.allclose(l.mu.eps, np.eye(3))
        out.append((complex(e[0, 0]), complex(e[2, 2]), float(l.d)))
    return out


def airy_reflection(Z, kz, d):
    """Reflection seen from layer 1 of a planar stack: interfaces
    r = (Z_j - Z_j+1) / (Z_j + Z_j+1) composed bottom-up with the round-trip
    phase exp(2 i kz_j d_j) of every finite layer."""
    n = len(Z)
    G = 0.0 + 0.0j
    for j in range(n - 2, -1, -1):
        r = (Z[j] - Z[j + 1]) / (Z[j] + Z[j + 1])
        if j + 1 < n - 1:
            P = G * np.exp(2j * kz[j + 1] * d[j + 1])
            G = (r + P) / (1 + r * P)
        else:
            G = r
    return G


def rs_rp(layers, omega, kr):
    """TE reflection of E_s and TM reflection of H_s (equivalently of the
    p-amplitude with p = s x K / k for both waves) at the top of a stack of
    z-uniaxial layers.  TE: kz = sqrt(eps_t w^2 - kr^2), Z = kz.  TM:
    kz = sqrt(eps_t (w^2 - kr^2 / eps_z)), Z = kz / eps_t (E_t = kz H / (w eps_t))."""
    d = [l[2] for l in layers]
    kzo = [kz_branch(et * omega ** 2 - kr ** 2) for et, ez, _ in layers]
    kze = [kz_branch(et * (omega ** 2 - kr ** 2 / ez)) for et, ez, _ in layers]
    Rs = airy_reflection(kzo, kzo, d)
    Rp = airy_reflection([k / l[0] for k, l in zip(kze, layers)], kze, d)
    return Rs, Rp


def layer1_dyadic(layers, omega, kx, ky, h, zo, terms="all"):
    """g~ in the reference's frame for a dipole at height h and the observation at
    height zo above the top interface, in an isotropic layer 1 (angular spectrum:
    direct (i / 2 eps1 k1z)(k1^2 I - K K) e^{i k1z |zo-h|}, reflected
    (i k1^2 / 2 eps1 k1z)(Rs s s^T + Rp p+ p-^T) e^{i k1z (zo+h)})."""
    e1 = layers[0][0]
    assert layers[0][1] == e1
    k1 = omega * np.sqrt(e1)
    kr = np.hypot(kx, ky)
    k1z = kz_branch(k1 ** 2 - kr ** 2)
    pref = 1j / (2 * e1 * k1z)
    g = np.zeros((3, 3), complex)
    if terms in ("all", "direct"):
        K = np.array([kx, ky, k1z if zo > h else -k1z])
        g += pref * (k1 ** 2 * np.eye(3) - np.outer(K, K)) * np.exp(1j * k1z * abs(zo - h))
    if terms in ("all", "scattered"):
        Rs, Rp = rs_rp(layers, omega, kr)
        s = np.array([-ky, kx, 0.0]) / kr
        pu = np.cross(s, [kx, ky, k1z]) / k1
        pd = np.cross(s, [kx, ky, -k1z]) / k1
        g += pref * k1 ** 2 * (Rs * np.outer(s, s) + Rp * np.outer(pu, pd)) * np.exp(1j * k1z * (zo + h))
    return FLIP @ g @ FLIP


def check_layer1_dyadic(stack, kts, pairs, terms="all", tol=1e-10):
    """The oracle's per-node dyadic (modes, S-matrix, T^o, seeds, basis and the
    g~ epilogue together) against layer1_dyadic, all nine entries."""
    layers = uniaxial_layers(stack)
    st = ostack(stack)
    kts = np.asarray(kts, float)
    zs = [p[0] for p in pairs]
    zo = [stack[0].d - p[1] for p in pairs]  # heights -> layer-1 cavity frame
    g = po.ghat_batch(st, W0, kts[:, 0], kts[:, 1], 1, 1, zs, zo, terms=terms)
    worst = 0.0
    for n, (kx, ky) in enumerate(kts):
        for p, (h, z) in enumerate(pairs):
            want = layer1_dyadic(layers, W0, kx, ky, h, z, terms)
            worst = max(worst, relerr(g[n, p], want))
    assert worst < tol, f"layer-1 dyadic vs angular-spectrum closed form {worst:.2e}"
    return worst


def check_gyro_normal_reflection(stack=None, pairs=((20 * NM, 40 * NM), (5 * NM, 300 * NM)), tol=1e-10):
    """A stack of isotropic layers and z-gyrotropic films at k_t = 0: the circular
    Jones vectors c+ = (1, -i)/sqrt2 and c- = (1, +i)/sqrt2 are eigenvectors of
    every layer's eps_t (eigenvalues exx +- g), so each sees its own Airy stack
    with n = sqrt(exx +- g) in the films, and the reflected dyadic in layer 1 is
    (i k1 / 2 eps1) [c+ r+ c+^H + c- r- c-^H] e^{i k1 (zo + h)} on the transverse
    block.  Fixes the handedness of the Faraday coupling (g -> -g swaps r+ and r-)."""
    stack = c2_stack() if stack is None else stack
    st = ostack(stack)
    e1 = complex(stack[0].eps.eps[0, 0])
    k1 = W0 * np.sqrt(e1)
    d = [l.d for l in stack]
    r = {}
    for sgn in (+1, -1):
        kz = []
        for l in stack:
            e = l.eps.eps
            kz.append(W0 * np.sqrt(e[0, 0] + sgn * (e[0, 1] / 1j)))  # e[0,1] = i g
        r[sgn] = airy_reflection(kz, kz, d)
    cp = np.array([1, -1j]) / np.sqrt(2)
    cm = np.array([1, 1j]) / np.sqrt(2)
    Rt = r[+1] * np.outer(cp, cp.conj()) + r[-1] * np.outer(cm, cm.conj())
    zs = [p[0] for p in pairs]
    zo = [stack[0].d - p[1] for p in pairs]
    g = po.ghat_batch(st, W0, [0.0], [0.0], 1, 1, zs, zo, terms="scattered")
    worst = 0.0
    for p, (h, z) in enumerate(pairs):
        want = np.zeros((3, 3), complex)
        want[:2, :2] = (1j * k1 / (2 * e1)) * Rt * np.exp(1j * k1 * (z + h))
        worst = max(worst, relerr(g[0, p], want))
    assert worst < tol, f"gyrotropic normal-incidence reflection {worst:.2e}"
    return worst


# ---------------------------------------------------------------------------
# 5. The reference's 3-layer semi-analytical model (z dipole above a film)
# ---------------------------------------------------------------------------
def semi_external_z(layers, omega, kr, wt, h, zo, rho):
    """eval_semi_external (mod_Dipz_SemiAnalytical.jl:371-434, with _A_coeffs
    :104-119) for a z dipole at height h above a film on a substrate and the
    observation at height zo in layer 1: the reflected part of (G_xz, G_zz) at
    lateral offsets rho (x > 0, y = 0), in the reference's frame, on the rule
    (kr, wt).  The film may be z-uniaxial: its TM k_z is sqrt(eps_t (w^2 -
    kr^2 / eps_z)) and eps_t replaces its eps in the TM admittances (the same
    algebra as fresnel, Gz_ClosedForm_Quadrature_Validation.jl:121-125)."""
    (e1, _, _), (et2, ez2, d2), (e3, _, _) = layers
    kr = np.asarray(kr, float)
    k1z = np.array([kz_branch(e1 * omega ** 2 - k ** 2) for k in kr])
    k2z = np.array([kz_branch(et2 * (omega ** 2 - k ** 2 / ez2)) for k in kr])
    k3z = np.array([kz_branch(e3 * omega ** 2 - k ** 2) for k in kr])
    ex2 = np.exp(2j * k2z * d2)
    f1 = et2 * k1z - e1 * k2z
    f2 = et2 * k1z + e1 * k2z
    f3 = e3 * k2z - et2 * k3z
    f4 = e3 * k2z + et2 * k3z
    R = (f1 * f4 + f2 * f3 * ex2) / (f2 * f4 + f1 * f3 * ex2)
    ph = np.exp(1j * k1z * (zo + h))
    rho = np.atleast_1d(rho)
    Gzz = np.array([np.sum(wt * 1j * kr ** 3 / k1z * R * j0(kr * x) * ph) for x in rho]) / (4 * np.pi * e1)
    Gxz = -np.array([np.sum(wt * kr ** 2 * R * j1(kr * x) * ph) for x in rho]) / (4 * np.pi * e1)
    return Gxz, Gzz


C1_HEIGHTS = np.array([20, 30, 40, 60, 80, 120, 160, 240]) * NM
C1_SOURCE = 10 * NM


def c1_rule():
    """A k_rho rule for the C1 stack: panels graded onto the light lines of air,
    the film and glass, out to 80 k0 (exp(-k (h + z')) < 1e-10 at 30 nm)."""
    return kr_rule([W0, W0 * np.sqrt(2.25), W0 * np.sqrt(2.5), W0 * np.sqrt(3.0)], kmax=80 * W0, width=4.0,
                   order=16, levels=30)


def check_semi_analytical_c1(rho=(0.01, 0.05, 0.2, 0.6, 1.3, 2.0), tol=1e-10):
    """C1 (air / 50 nm uniaxial film / glass, z dipole 10 nm above): the oracle's 1D
    closed_form_G (reflected part) equals the semi-analytical model on the same
    rule, at every observation height of the configuration; its direct term is
    the Sommerfeld sum of the free-space field on the same rule (the free-space
    closed form itself is pinned in tests/test_oracle.py).

    Both sums run over the nodes with 2 Im(k1z) d1 <= 700.  Beyond that the
    reference's own formulation drops a node: layer 1 is referenced at its top
    plane d1 = 1 above the film, so the round trip exp(-2 kappa1 d1) leaves the
    double range (the _phase window, mod_ClosedFormField.jl:59-69, and the
    denormal flush, mod_FieldSolver.jl:24-26).  Returns the relative weight of
    those dropped nodes in the model's G as well."""
    stack = c1_stack()
    layers = uniaxial_layers(stack)
    st = ostack(stack)
    kr, wt = c1_rule()
    k1z = np.array([kz_branch(W0 ** 2 - k ** 2) for k in kr])
    keep = 2 * k1z.imag * stack[0].d <= 700.0
    wk = np.where(keep, wt, 0.0)
    c = po.nodes_batch(st, W0, kr, np.zeros_like(kr), 1, [1])
    assert c.valid.all()
    rho = np.asarray(rho, float)
    zcav = stack[0].d - C1_HEIGHTS
    G = po.closed_form_G_1d(c, 1, wk, C1_SOURCE, zcav, rho, [0.0], with_direct=False)
    Gd = po.closed_form_G_1d(c, 1, wk, C1_SOURCE, zcav, rho, [0.0], with_direct=True)
    worst = worst_d = dropped = 0.0
    for iz, zo in enumerate(C1_HEIGHTS):
        Gxz, Gzz = semi_external_z(layers, W0, kr, wk, C1_SOURCE, zo, rho)
        got = np.stack([G[:, 0, iz, 0, 2], G[:, 0, iz, 2, 2]])
        worst = max(worst, relerr(got, np.stack([Gxz, Gzz])))
        ph = np.exp(1j * k1z * (zo - C1_SOURCE))  # observation above the source
        Dzz = np.array([np.sum(wk * 1j * kr ** 3 / k1z * j0(kr * x) * ph) for x in rho]) / (4 * np.pi)
        Dxz = -np.array([np.sum(wk * kr ** 2 * j1(kr * x) * ph) for x in rho]) / (4 * np.pi)
        gotd = np.stack([Gd[:, 0, iz, 0, 2], Gd[:, 0, iz, 2, 2]]) - got
        worst_d = max(worst_d, relerr(gotd, np.stack([Dxz, Dzz])))
        Fxz, Fzz = semi_external_z(layers, W0, kr, wt, C1_SOURCE, zo, rho)
        dropped = max(dropped, relerr(np.stack([Gxz, Gzz]), np.stack([Fxz, Fzz])))
    assert worst < tol, f"C1 reflected G vs the semi-analytical model {worst:.2e}"
    assert worst_d < tol, f"C1 direct G vs its Sommerfeld sum {worst_d:.2e}"
    return worst, worst_d, dropped


# ---------------------------------------------------------------------------
# 6. The reference's logged s/p reflection norms of its base and hyperbolic stacks
# ---------------------------------------------------------------------------
PW_DIRECTION = np.array([0.5, 0.3]) / np.hypot(0.5, 0.3)  # plane_wave_slice.jl:44
PW_FILMS = {  # films bottom -> top as (eps_t, eps_z, d), plane_wave_slice.jl:37-40
    "base": [(2.25, 2.25, 1 / 4), (4.0, 4.0, 3 / 16), (2.0 + 0.1j, 2.0 + 0.1j, 1 / 8)],
    "hyperbolic": [(2.25, 2.25, 1 / 4), (4.0 + 0.1j, -3.2 + 0.3j, 3 / 16), (2.0 + 0.1j, 2.0 + 0.1j, 1 / 8)],
}


def pw_stack(name, general_eigen=False):
    """own_stack (plane_wave_slice.jl:50-52) with layer 1 of zero thickness, so the
    S-matrix is referenced at the stack face (:63-64).  general_eigen gives the
    films as full tensors: the same layers through the general eigen-solve."""
    def mat(et, ez):
        if general_eigen and et != ez:
            return GyrotropicMaterial(np.diag([et, et, ez]))
        return IsotropicMaterial(et) if et == ez else UniaxialMaterial(et, ez)
    films = [lay(mat(complex(et), complex(ez)), d) for et, ez, d in reversed(PW_FILMS[name])]
    return [lay(IsotropicMaterial(1.0), 0.0)] + films + [lay(IsotropicMaterial(1.0), 4 / 64)]


def sp_basis(kt, kz):
    """(s, p) of a wave with in-plane k_t and vertical wavenumber kz in the z-up frame,
    p = s x k normalised by its Hermitian norm (sp_basis, plane_wave_slice.jl:55-59)."""
    k = np.array([kt[0], kt[1], kz], complex)
    nt = np.hypot(*kt)
    s = np.array([0, 1, 0], complex) if nt == 0 else np.array([-kt[1], kt[0], 0], complex) / nt
    p = np.cross(s, k)
    return s, p / np.linalg.norm(p)


def reflection_sp(stack, q, direction=PW_DIRECTION):
    """Physical reflection matrix at the stack face in the (s, p) bases of the
    incident (downward) and reflected waves, from the oracle's S-matrix
    (reflection_sp, plane_wave_slice.jl:63-80)."""
    st = ostack(stack)
    kt = q * np.asarray(direction)
    kzv = np.sqrt(complex(W0 ** 2 - q ** 2))
    M, _, _, _, SR = po.stack_smatrix(st, W0, kt[0], kt[1])
    S21 = SR[0][2:4, 0:2]
    k_own = np.array([kt[0], kt[1], kzv])
    inb, outb = sp_basis(kt, -kzv), sp_basis(kt, kzv)
    B = np.stack([o[:2] for o in outb], axis=1)
    R = np.zeros((2, 2), complex)
    for j, e in enumerate(inb):
        E = FLIP @ e
        Hm = np.cross(k_own, E) / W0
        a = np.linalg.lstsq(M[0]["Vf"], np.array([E[0], E[1], Hm[0], Hm[1]]), rcond=None)[0]
        psr = M[0]["Vb"] @ (S21 @ a)
        R[:, j] = np.linalg.solve(B, psr[:2])
    return R


def pw_logged():
    with open(os.path.join(GOLDEN, "plane_wave_rsp.json")) as f:
        return json.load(f)["cases"]


def _sigfig_tol(text):
    """Half a unit in the last printed digit of a logged value."""
    t = text.strip()
    dec = len(t.split(".")[1]) if "." in t else 0
    return 0.5 * 10.0 ** (-dec)


def check_plane_wave_logs(tol_tmm=1e-12):
    """For each logged (stack, incidence): the largest singular value of the
    oracle's s/p reflection matrix equals the logged ||r_sp|| to its printed
    digits; the matrix is diag(Rs, Rp) of the TE/TM Airy recursion to 1e-12;
    and the films given as full tensors (general eigen-solve) give the same."""
    worst = 0.0
    for case in pw_logged():
        q = case["kt_over_k0"] * W0
        R = reflection_sp(pw_stack(case["stack"]), q)
        sig = np.linalg.svd(R, compute_uv=False)[0]
        assert abs(sig - float(case["r_sp"])) <= _sigfig_tol(case["r_sp"]), (case, sig)
        layers = [(1.0, 1.0, 0.0)] + [(et, ez, d) for et, ez, d in reversed(PW_FILMS[case["stack"]])] + [
            (1.0, 1.0, 4 / 64)]
        Rs, Rp = rs_rp(layers, W0, q)
        worst = max(worst, relerr(R, np.diag([Rs, Rp])))
        Rg = reflection_sp(pw_stack(case["stack"], general_eigen=True), q)
        worst = max(worst, relerr(Rg, np.diag([Rs, Rp])))
    assert worst < tol_tmm, f"s/p reflection vs TE/TM Airy {worst:.2e}"
    return worst


# ---------------------------------------------------------------------------
# The check list shared by the closed-form tests and the mutant test
# ---------------------------------------------------------------------------
C1_PAIRS = [(C1_SOURCE, z) for z in C1_HEIGHTS] + [(300 * NM, 40 * NM)]


def c1_kts():
    rng = np.random.default_rng(11)
    kr = np.concatenate([rng.uniform(0.05, 0.99, 8) * W0, rng.uniform(1.01, 1.7, 8) * W0,
                         rng.uniform(1.8, 40, 10) * W0])
    ph = rng.uniform(0, 2 * np.pi, kr.size)
    return np.stack([kr * np.cos(ph), kr * np.sin(ph)], axis=1)


def hyperbolic_5_stack():
    return [lay(IsotropicMaterial(1.0 + 0.01j), 1.0), lay(UniaxialMaterial(4.0 + 0.1j, -3.2 + 0.3j), 60 * NM),
            lay(IsotropicMaterial(2.25 + 0.01j), 80 * NM), lay(UniaxialMaterial(2.5 + 0.01j, 3.0 + 0.01j), 40 * NM),
            lay(IsotropicMaterial(2.25), 1.0)]


def all_checks():
    """(name, callable) of every closed-form check."""
    gen = passive_tensor([[3.0, 0.4 - 0.2j, 0.3], [0.4 + 0.2j, 2.5, 0.2 - 0.1j], [0.3, 0.2 + 0.1j, 2.2]],
                         [0.5, -0.3 + 0.2j, 0.4j], 0.04)
    mu = passive_tensor([[1.2, 0.1j, 0.05], [-0.1j, 1.1, 0.0], [0.05, 0.0, 0.9]], [0.2, 0.1j, -0.1], 0.01)
    tilt = tilted_uniaxial(2.5 + 0.01j, 3.5 + 0.02j, 0.6)
    return [
        ("null vectors, gyrotropic", lambda: check_bulk_maxwell_null_vectors(gyro_tensor(3 + 0.02j, 0.3, 2.8 + 0.02j))),
        ("null vectors, tilted uniaxial", lambda: check_bulk_maxwell_null_vectors(tilt)),
        ("null vectors, biaxial", lambda: check_bulk_maxwell_null_vectors(np.diag([2.0 + 0.01j, 3.0 + 0.01j, 4.0 + 0.01j]))),
        ("null vectors, general eps and mu", lambda: check_bulk_maxwell_null_vectors(gen, mu)),
        ("gyrotropic normal-incidence modes", check_gyro_normal_incidence),
        ("particular solution, tilted uniaxial", lambda: check_particular_solution(tilt)),
        ("particular solution, gyrotropic", lambda: check_particular_solution(gyro_tensor(4 + 0.01j, 0.3, 3 + 0.01j))),
        ("particular solution, general", lambda: check_particular_solution(gen)),
        ("layer-1 dyadic, C1", lambda: check_layer1_dyadic(c1_stack(), c1_kts(), C1_PAIRS)),
        ("layer-1 dyadic, hyperbolic 5-layer", lambda: check_layer1_dyadic(hyperbolic_5_stack(), c1_kts(), C1_PAIRS)),
        ("gyrotropic normal-incidence reflection, C2", check_gyro_normal_reflection),
        ("3-layer semi-analytical model, C1", check_semi_analytical_c1),
        ("logged s/p reflections", check_plane_wave_logs),
    ]
