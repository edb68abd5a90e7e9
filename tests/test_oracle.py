"""Pins the CPU oracle (oracle/ytf_oracle.cpp) before it is trusted as the
parity reference: reference-held vectors, closed forms, library routines and
invariants.  Every check here can fail for a plausible mistake in the oracle
(dropped term, sign, index, transposed operand)."""
import os

import numpy as np
import pytest

from helpers import (po, relerr, ostack, lay, FOUNDATION_STACKS, FOUNDATION_KTS, IsotropicMaterial,
                     UniaxialMaterial, BiaxialMaterial, GyrotropicMaterial, FlatLayer, FlatOpticalLayer,
                     kr_rule, kr_grid, gl_rule, uniform_phi, polar_nodes, lateral_sum, E_free, graded, panels, c2_stack,
                     m_vac, m_gl, m_bi)

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
W0 = 2 * np.pi


# ---------------------------------------------------------------------------
# Eigen-solver: LAPACK zgeev (numpy) as the library check
# ---------------------------------------------------------------------------
def _lapack_sorted(A):
    w, X = np.linalg.eig(A)
    o = np.lexsort((w.imag, w.real))
    return w[o], X[:, o]


def _check_eig(A, tol_val=1e-12, tol_vec=1e-10):
    Q, V = po.eigen4(A)
    w, X = _lapack_sorted(A)
    assert np.max(np.abs(Q - w)) <= tol_val * np.max(np.abs(w))
    # unit norm, largest component real positive (zgeev normalisation)
    assert np.allclose(np.linalg.norm(V, axis=0), 1.0, atol=1e-13)
    for k in range(4):
        i = np.argmax(np.abs(V[:, k]) ** 2)
        assert abs(V[i, k].imag) <= 1e-15 and V[i, k].real > 0
        # same vector up to a unit phase (the phase is ambiguous only on ties)
        assert abs(abs(np.vdot(X[:, k], V[:, k])) - 1.0) <= tol_vec
    assert np.linalg.norm(A @ V - V * Q) <= 1e-13 * np.linalg.norm(A)


def test_eigen_random_matrices():
    rng = np.random.default_rng(1)
    for _ in range(300):
        _check_eig(rng.normal(size=(4, 4)) + 1j * rng.normal(size=(4, 4)))


@pytest.mark.parametrize("name", ["gyro", "biaxial", "hyperbolic"])
def test_eigen_berreman_matrices(name):
    eps = {"gyro": np.array([[3 + 0.02j, 0.3j, 0], [-0.3j, 3 + 0.02j, 0], [0, 0, 2.8 + 0.02j]]),
           "biaxial": np.diag([2.0 + 0.01j, 3.0 + 0.01j, 4.0 + 0.01j]),
           "hyperbolic": np.diag([-3.0 + 0.2j, 2.0 + 0.1j, 5.0 + 0.3j])}[name]
    mu = np.eye(3, dtype=complex)
    for kx in np.linspace(-50, 50, 9):
        for ky in np.linspace(-37, 41, 7):
            if kx == 0 and ky == 0:
                continue
            _check_eig(po.berreman(eps, mu, W0, kx, ky))


def test_berreman_isotropic_eigenvalues_closed_form():
    # isotropic layer: Q = +-sqrt(w^2 eps - k^2) (each twice), against the general eigen path
    eps = 2.25 + 0.01j
    for kx, ky in [(0.3, 0.0), (4.0, 5.0), (20.0, -3.0)]:
        B = po.berreman(eps * np.eye(3), np.eye(3, dtype=complex), W0, kx, ky)
        kz = np.sqrt(W0 ** 2 * eps - kx ** 2 - ky ** 2)
        Q = np.sort_complex(np.linalg.eigvals(B))
        want = np.sort_complex(np.array([kz, kz, -kz, -kz]))
        assert np.max(np.abs(Q - want)) < 1e-9 * abs(kz)


def test_closed_form_modes_satisfy_berreman():
    # _eigen_QV closed form (mod_ModeSolver.jl:243-265): B Psi = Q Psi for uniaxial and isotropic
    for (et, ez) in [(2.5 + 0.01j, 3.0 + 0.01j), (1.0 + 0j, 1.0 + 0j), (4 + 0.1j, -3.2 + 0.3j)]:
        eps = np.diag([et, et, ez])
        for kx, ky in [(0.0, 0.0), (3.0, 0.0), (2.0, -7.0), (40.0, 10.0)]:
            Q, V, _ = po.modes_layer(eps, np.eye(3), 1, 0, W0, kx, ky)
            B = po.berreman(eps, np.eye(3, dtype=complex), W0, kx, ky)
            assert np.linalg.norm(B @ V - V * Q) <= 1e-12 * np.linalg.norm(B) * np.linalg.norm(V)


# ---------------------------------------------------------------------------
# S-matrix: the reference's own reflectance vectors (R_own) and TMM/Airy
# ---------------------------------------------------------------------------
CASES = {"film": [(1.25, 0.5)], "stack": [(1.25, 0.25), (3.0, 3 / 16), (1.0 + 0.1j, 1 / 8)]}


def _own_stack(films):
    # own_stack -- comparisons/gila_airy/airy_comparison.jl:66-69
    vac = lay(IsotropicMaterial(1.0), 0.0)
    return [vac] + [lay(IsotropicMaterial(1 + chi), d) for chi, d in films] + [vac]


def _own_reflectance(st, theta, pol):
    # own_reflectance -- airy_comparison.jl:71-82
    kx = W0 * np.sin(np.deg2rad(theta))
    M, _, _, _, SR = po.stack_smatrix(st, W0, kx, 0.0)
    S = SR[0]
    c = np.cos(np.deg2rad(theta))
    psi = np.array([0, 1, -c, 0.0]) if pol == "s" else np.array([c, 0, 0, 1.0])
    a = np.linalg.lstsq(M[0]["Vf"], psi, rcond=None)[0]
    psir = M[0]["Vb"] @ (S[2:4, 0:2] @ a)
    return abs(psir[1]) ** 2 + abs(psir[0]) ** 2 / c ** 2


def _tmm(films, theta, pol):
    # tmm_reflectance -- airy_comparison.jl:35-50
    kx = W0 * np.sin(np.deg2rad(theta))
    kz0 = W0 * np.cos(np.deg2rad(theta))
    eta = (lambda e, kz: kz / W0) if pol == "s" else (lambda e, kz: e * W0 / kz)
    e0 = eta(1.0, kz0)
    M = np.eye(2, dtype=complex)
    for chi, d in films:
        e = 1 + chi
        kz = np.sqrt(complex(e * W0 ** 2 - kx ** 2))
        dl, ej = kz * d, eta(e, kz)
        M = M @ np.array([[np.cos(dl), -1j * np.sin(dl) / ej], [-1j * ej * np.sin(dl), np.cos(dl)]])
    num = e0 * M[0, 0] + e0 ** 2 * M[0, 1] - M[1, 0] - e0 * M[1, 1]
    den = e0 * M[0, 0] + e0 ** 2 * M[0, 1] + M[1, 0] + e0 * M[1, 1]
    return abs(num / den) ** 2


@pytest.mark.parametrize("case", ["film", "stack"])
def test_reflectance_matches_reference_tsv(case):
    rows = [l.split("\t") for l in open(os.path.join(GOLDEN, f"{case}.tsv")).read().strip().split("\n")[1:]]
    st = ostack(_own_stack(CASES[case]))
    worst = max(abs(_own_reflectance(st, float(a), p) - float(ro)) for p, a, rt, ro, rg in rows)
    # the TSV carries 10 decimals
    assert worst < 1e-10, worst
    assert len(rows) == 70


@pytest.mark.parametrize("case", ["film", "stack"])
def test_reflectance_matches_tmm(case):
    st = ostack(_own_stack(CASES[case]))
    for pol in "sp":
        for th in np.arange(0, 86, 5.0):
            assert abs(_own_reflectance(st, th, pol) - _tmm(CASES[case], th, pol)) < 1e-13


def test_lossless_energy_conservation():
    # lossless film at normal incidence: |S11|^2 + |S21|^2 = 1 (s-pol), computed via the cascade
    st = ostack([lay(IsotropicMaterial(1.0), 0.0), lay(IsotropicMaterial(2.25), 0.37), lay(IsotropicMaterial(1.0), 0.0)])
    for kx in [0.0, 1.3, 4.9]:
        M, _, _, _, SR = po.stack_smatrix(st, W0, kx, 0.0)
        S = SR[0]
        # power-normalised mode amplitudes: Pz of each vacuum mode
        Pf = np.array([0.25 * np.real(np.vdot(M[0]["Vf"][:, m], M[0]["Wf"][:, m])) for m in range(2)])
        Pb = np.array([0.25 * np.real(np.vdot(M[0]["Vb"][:, m], M[0]["Wb"][:, m])) for m in range(2)])
        PfN = np.array([0.25 * np.real(np.vdot(M[2]["Vf"][:, m], M[2]["Wf"][:, m])) for m in range(2)])
        for m in range(2):
            refl = sum(abs(S[2 + i, m]) ** 2 * abs(Pb[i]) for i in range(2)) / Pf[m]
            trans = sum(abs(S[i, m]) ** 2 * PfN[i] for i in range(2)) / Pf[m]
            assert abs(refl + trans - 1.0) < 1e-12


def test_redheffer_associativity():
    # SRight[0] computed right-to-left equals the left-to-right cascade (the GPU order)
    st = ostack(c2_stack())
    M, Sint, Sprop, SL, SR = po.stack_smatrix(st, W0, 3.1, -2.2)
    # reconstruct the cascade from Sprop/Sint in numpy with the same star product
    def star(A, B):
        A11, A12, A21, A22 = A[:2, :2], A[:2, 2:], A[2:, :2], A[2:, 2:]
        B11, B12, B21, B22 = B[:2, :2], B[:2, 2:], B[2:, :2], B[2:, 2:]
        I = np.eye(2)
        M1, M2 = I - A12 @ B21, I - B21 @ A12
        return np.block([[B11 @ np.linalg.solve(M1, A11), B12 + B11 @ np.linalg.solve(M1, A12) @ B22],
                         [A21 + A22 @ np.linalg.solve(M2, B21) @ A11, A22 @ np.linalg.solve(M2, B22)]])
    n = len(Sprop)
    eff = [star(Sprop[k], Sint[k]) for k in range(n - 1)]
    acc = eff[0]
    for k in range(1, n - 1):
        acc = star(acc, eff[k])
    assert relerr(acc, SR[0]) < 1e-12


# ---------------------------------------------------------------------------
# T^o and seeds: Gz_ClosedForm_Foundation_Validation.jl:38-76
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("label", list(FOUNDATION_STACKS))
def test_foundation_transfer_identity(label):
    stack = FOUNDATION_STACKS[label]
    st = ostack(stack)
    N = len(stack)
    worst_cross = worst_same = 0.0
    for kx, ky in FOUNDATION_KTS:
        for layer in range(1, N):
            d = stack[layer - 1].d
            c = po.nodes_batch(st, W0, [kx], [ky], layer, list(range(1, N + 1)))
            assert c.valid[0]
            sif = c.sf[0]
            sib_raw = -c.sb[0]
            for q in range(3):
                for zf in (0.13, 0.5, 0.91):
                    z = zf * d
                    # Modal_amp convention: the seed is (sif, -sib)
                    u = po.seed_from_source(st, W0, kx, ky, layer, z, sif[:, q], -sib_raw[:, q])
                    Ef, Eb = po.modal_amp(st, W0, kx, ky, layer, z, sif[:, q], -sib_raw[:, q])
                    for o in range(1, N + 1):
                        v = c.T[o][0] @ u
                        if o == layer:
                            zpos = d - z if layer == 1 else z
                            pf, pb = po.modal_amp_profile(st, W0, kx, ky, layer, z, sif[:, q], -sib_raw[:, q],
                                                          [0.0, d])
                            Qf, Qb = c.Qf[0], c.Qb[0]
                            Ef_refl = pf[1] - sif[:, q] * np.exp(1j * Qf * (-zpos))
                            Eb_refl = pb[0] - (-sib_raw[:, q]) * np.exp(1j * Qb * (d - zpos))
                            e = np.linalg.norm(v - np.concatenate([Ef_refl, Eb_refl])) / max(
                                np.linalg.norm(np.concatenate([pf[1], pb[0]])), 1e-300)
                            worst_same = max(worst_same, e)
                        else:
                            worst_cross = max(worst_cross, relerr(v, np.concatenate([Ef[o - 1], Eb[o - 1]])))
    assert worst_cross < 1e-11, worst_cross
    assert worst_same < 1e-11, worst_same


# The seed's particular solution against a bulk Maxwell solve (which pins the
# eps_tz term of the source vector) is in tests/test_oracle_closed_forms.py.


# ---------------------------------------------------------------------------
# Synthesis: free-space dipole (Gz_ClosedForm_Point_Validation.jl:162-182)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("wf", [2 * np.pi, np.pi, 1.0, 3.7])
def test_free_space_dipole_1d_and_2d(wf):
    ef = 1.0 + 1e-4j
    stack = [lay(IsotropicMaterial(ef), 3.0) for _ in range(3)]
    st = ostack(stack)
    zs, zo, x = 1.0, 1.7, 0.4
    k = wf * np.sqrt(ef)
    kx, wt = kr_rule([wf * np.sqrt(ef).real], kmax=36 / 0.7 + wf, width=0.5, order=16)
    Gref = np.stack([E_free([x, 0, zo - zs], np.eye(3)[q], k, ef) for q in range(3)], axis=1)
    c = po.nodes_batch(st, wf, kx, np.zeros_like(kx), 2, [2])
    G1 = po.closed_form_G_1d(c, 2, wt, zs, [zo], [x], [0.0])[0, 0, 0]
    assert relerr(G1, Gref) < 1e-10
    phis, wph = uniform_phi(48)
    kx2, ky2, W = polar_nodes(kx, wt, phis, wph)
    g = po.ghat_batch(st, wf, kx2, ky2, 2, 2, [zs], [zo])[:, 0]
    G2 = lateral_sum(g, W, kx2, ky2, [x], [0.0])[0, 0]
    assert relerr(G2, Gref) < 1e-10
    # an overall w/2pi factor would be visible here
    if abs(wf - 2 * np.pi) > 1:
        assert relerr(G2 * wf / (2 * np.pi), Gref) > 1e-3


# ---------------------------------------------------------------------------
# Reflected amplitudes: dipole above an (eps, mu) half-space vs the Sommerfeld
# integral with analytic Fresnel coefficients
# (Gz_ClosedForm_Quadrature_Validation.jl:117-175)
# ---------------------------------------------------------------------------
def _kzroot(x):
    q = np.sqrt(complex(x))
    return -q if q.imag < 0 else q


def _fresnel(k, et, ez, mt, mz):
    kz1 = _kzroot(1 - k ** 2)
    kzs = _kzroot(et * mt - k ** 2 * mt / mz)
    kzp = _kzroot(et * mt - k ** 2 * et / ez)
    return (mt * kz1 - kzs) / (mt * kz1 + kzs), (et * kz1 - kzp) / (et * kz1 + kzp)


def _branchpts(eps, mu):
    pts = list(np.sqrt(np.array([eps[0] * mu[1], mu[0] * eps[1]], dtype=complex)))
    if eps[0] == eps[1] and mu[0] == mu[1]:
        e, m = eps[0], mu[0]
        pts += list(np.sqrt(np.array([e * (m - e) / (1 - e ** 2), m * (e - m) / (1 - m ** 2)], dtype=complex)))
    return sorted(set(p.real for p in pts if p.real > 0))


def _sommerfeld(eps, mu, d):
    kb = _branchpts(eps, mu)
    szb = [np.sqrt(1 - s ** 2) for s in kb if s < 1]
    vb = [np.sqrt(s ** 2 - 1) for s in kb if s > 1]
    vmax = 50 / (2 * d)
    Izz = Ixx = 0j
    br = list(np.arange(0, 1.0001, 0.025))
    for b in szb:
        br += graded(b, 0.0, 1.0)
    x, w = panels(br, 20)
    for sz, wi in zip(x, w):
        s = np.sqrt(1 - sz ** 2)
        rs, rp = _fresnel(s, eps[0], eps[1], mu[0], mu[1])
        e = np.exp(1j * 2 * d * sz)
        Izz += wi * s ** 2 * rp * e
        Ixx += wi * (rs - sz ** 2 * rp) * e
    br = [0.0, vmax] + [t for t in 1e-4 * 1.15 ** np.arange(121) if t <= vmax]
    for b in vb:
        br += graded(b, 0.0, vmax)
    x, w = panels([t for t in br if t <= vmax], 16)
    for v, wi in zip(x, w):
        s = np.sqrt(1 + v ** 2)
        rs, rp = _fresnel(s, eps[0], eps[1], mu[0], mu[1])
        e = np.exp(-2 * d * v)
        Izz += wi * (-1j) * s ** 2 * rp * e
        Ixx += wi * (-1j) * (rs + v ** 2 * rp) * e
    return 1j / (4 * np.pi) * Izz, 1j / (8 * np.pi) * Ixx


HALF_SPACES = [((2.25 + 0j, 2.25 + 0j), (1.5 + 0j, 1.5 + 0j)),
               ((-3 + 0.2j, -3 + 0.2j), (2 + 0.1j, 2 + 0.1j)),
               ((-2 + 0.3j, -2 + 0.3j), (-1.5 + 0.3j, -1.5 + 0.3j)),
               ((3 + 0.1j, -2 + 0.1j), (1.6 + 0.05j, 0.7 + 0.05j))]


@pytest.mark.parametrize("mat", range(len(HALF_SPACES)))
@pytest.mark.parametrize("d", [0.2, 1.0])
def test_half_space_sommerfeld(mat, d):
    eps, mu = HALF_SPACES[mat]
    iso = eps[0] == eps[1] and mu[0] == mu[1]
    if iso:
        half = FlatLayer(IsotropicMaterial(eps[0]), IsotropicMaterial(mu[0]), 1.0)
    else:
        half = FlatLayer(UniaxialMaterial(*eps), UniaxialMaterial(*mu), 1.0)
    stack = [lay(IsotropicMaterial(1.0 + 0j), 2 * d), half]
    st = ostack(stack)
    # real-axis branch points of the two media, integrated with gl_rule's
    # u^2 substitution (mod_KxGrid.jl:361-384)
    sing = sorted({1.0, float(np.sqrt(complex(eps[0] * mu[1])).real), float(np.sqrt(complex(mu[0] * eps[1])).real)})
    sing = [s for s in sing if s > 0]
    kmax = 30 / (2 * d)
    br = sorted(set(list(np.arange(0.0, kmax, 0.05)) + [kmax] + sing))
    kr, wr = gl_rule(list(zip(br[:-1], br[1:])), order=16, singular=sing)
    phis, wph = uniform_phi(16)
    kx, ky, W = polar_nodes(kr, wr, phis, wph)
    g = po.ghat_batch(st, 1.0, kx, ky, 1, 1, [d], [d])[:, 0]
    G = lateral_sum(g, W, kx, ky, [0.0], [0.0])[0, 0]
    Gzz, Gxx = _sommerfeld(eps, mu, d)
    e = max(abs(G[2, 2] - Gzz) / abs(Gzz), abs(G[0, 0] - Gxx) / abs(Gxx), abs(G[1, 1] - Gxx) / abs(Gxx))
    assert e < 1e-6, e


# ---------------------------------------------------------------------------
# Invariants: reciprocity, 90-degree covariance, 1D = 2D
# (Gz_ClosedForm_Point_Validation.jl:97-160)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("mat", ["iso", "biaxial"])
def test_reciprocity(mat):
    m = m_gl if mat == "iso" else m_bi
    stack = [lay(m_vac, 1.0), lay(m, 0.7), lay(m_gl, 0.6), lay(m_vac, 1.0)]
    st = ostack(stack)
    kr, wr = kr_grid(stack, W0, kmax=15.0)
    phis, wph = uniform_phi(48)
    kx, ky, W = polar_nodes(kr, wr, phis, wph)
    z2, z3 = 0.45, 0.2
    g23 = po.ghat_batch(st, W0, kx, ky, 2, 3, [z2], [z3])[:, 0]
    g32 = po.ghat_batch(st, W0, kx, ky, 3, 2, [z3], [z2])[:, 0]
    for dx, dy in [(0.0, 0.0), (0.3, 0.0), (0.2, -0.25)]:
        G23 = lateral_sum(g23, W, kx, ky, [dx], [dy])[0, 0]
        G32 = lateral_sum(g32, W, kx, ky, [-dx], [-dy])[0, 0]
        assert np.linalg.norm(G23 - G32.T) / np.linalg.norm(G23) < 1e-10


@pytest.mark.parametrize("which", ["biaxial", "gyrotropic"])
def test_rotation_covariance(which):
    if which == "biaxial":
        A, B = m_bi, BiaxialMaterial(3.0 + 0.01j, 2.0 + 0.01j, 4.0 + 0.01j)
    else:
        A = B = GyrotropicMaterial(np.array([[2.1 + 0.1j, 1.2j, 0], [-1.2j, 2.1 + 0.1j, 0], [0, 0, 1.5 + 0.05j]]))
    stA = [lay(m_vac, 1.0), lay(A, 0.7), lay(m_gl, 0.5), lay(m_vac, 1.0)]
    stB = [lay(m_vac, 1.0), lay(B, 0.7), lay(m_gl, 0.5), lay(m_vac, 1.0)]
    kr, wr = kr_grid(stA, W0, kmax=15.0)
    phis, wph = uniform_phi(48)
    kx, ky, W = polar_nodes(kr, wr, phis, wph)
    R = np.array([[0.0, -1, 0], [1, 0, 0], [0, 0, 1]])
    xo = np.arange(-0.4, 0.41, 0.2)
    for obs, zo in [(2, 0.1), (2, 0.6), (1, 0.8), (3, 0.25)]:
        gA = po.ghat_batch(ostack(stA), W0, kx, ky, 2, obs, [0.35], [zo])[:, 0]
        gB = po.ghat_batch(ostack(stB), W0, kx, ky, 2, obs, [0.35], [zo])[:, 0]
        GA = lateral_sum(gA, W, kx, ky, xo, xo)
        GB = lateral_sum(gB, W, kx, ky, xo, xo)
        No = len(xo)
        err = max(np.linalg.norm(GB[No - 1 - j, i] - R @ GA[i, j] @ R.T) for i in range(No) for j in range(No))
        scale = max(np.linalg.norm(GA[i, j]) for i in range(No) for j in range(No))
        assert err / scale < 1e-10


def test_1d_synthesis_equals_2d_sum_on_symmetric_stack():
    stack = FOUNDATION_STACKS["uniaxial 4-layer"]
    st = ostack(stack)
    kr, wr = kr_grid(stack, W0)
    phis, wph = uniform_phi(64)
    kx, ky, W = polar_nodes(kr, wr, phis, wph)
    xs = np.array([-0.5, -0.2, 0.0, 0.3])
    for src in (1, 2, 3):
        zs = 0.4 * stack[src - 1].d
        c1 = po.nodes_batch(st, W0, kr, np.zeros_like(kr), src, [1, 2, 3, 4])
        for obs in (1, 2, 3, 4):
            zo = [0.05 * stack[obs - 1].d, 0.95 * stack[obs - 1].d] if obs == src else [0.3 * stack[obs - 1].d]
            G1 = po.closed_form_G_1d(c1, obs, wr, zs, zo, xs, xs)
            g = po.ghat_batch(st, W0, kx, ky, src, obs, [zs] * len(zo), zo)
            for zi in range(len(zo)):
                G2 = lateral_sum(g[:, zi], W, kx, ky, xs, xs)
                assert relerr(G2, G1[:, :, zi]) < 1e-10


def test_op_count_is_stable_and_plausible():
    st = ostack(c2_stack())
    c1 = po.count_node_ops(st, W0, 0.3, 0.2, 1, 1, 20 / 600, 40 / 600)
    c2 = po.count_node_ops(st, W0, 40.0, -30.0, 1, 1, 20 / 600, 40 / 600)
    f = lambda c: c["solve"]["add"] + c["solve"]["mul"] + c["solve"]["div"]
    assert 1e4 < f(c1) < 1e5
    assert abs(f(c1) - f(c2)) / f(c1) < 0.1


# ---------------------------------------------------------------------------
# Term classes of the spectral dyadic (_term_amplitudes,
# mod_ClosedFormField.jl:143-160): direct + scattered = all and
# toeplitz + hankel = all on any medium; no scattering in a homogeneous stack.
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("label", ["iso 5-layer lossy", "gyrotropic 4-layer", "biaxial 4-layer"])
def test_term_classes_add_up(label):
    stack = FOUNDATION_STACKS[label]
    st = ostack(stack)
    rng = np.random.default_rng(11)
    kx, ky = rng.uniform(-9, 9, 40), rng.uniform(-9, 9, 40)
    src = 2
    d = stack[src - 1].d
    zs, zo = [0.3 * d, 0.6 * d], [0.7 * d, 0.2 * d]
    g = {t: po.ghat_batch(st, W0, kx, ky, src, src, zs, zo, terms=t)
         for t in ("all", "direct", "scattered", "toeplitz", "hankel")}
    assert relerr(g["direct"] + g["scattered"], g["all"]) < 1e-13
    assert relerr(g["toeplitz"] + g["hankel"], g["all"]) < 1e-13
    assert relerr(g["scattered"], g["all"]) > 1e-3  # both classes are present
    # in another layer there is no direct term
    g2 = po.ghat_batch(st, W0, kx, ky, src, 1, zs, [0.5, 0.6], terms="scattered")
    g3 = po.ghat_batch(st, W0, kx, ky, src, 1, zs, [0.5, 0.6], terms="all")
    assert np.array_equal(g2, g3)


def test_homogeneous_stack_has_no_scattered_term():
    ef = 2.0 + 0.05j
    st = ostack([lay(IsotropicMaterial(ef), 1.0) for _ in range(3)])
    rng = np.random.default_rng(12)
    kx, ky = rng.uniform(-9, 9, 30), rng.uniform(-9, 9, 30)
    g = po.ghat_batch(st, W0, kx, ky, 2, 2, [0.3], [0.7], terms="all")
    gs = po.ghat_batch(st, W0, kx, ky, 2, 2, [0.3], [0.7], terms="scattered")
    assert np.linalg.norm(gs) < 1e-14 * np.linalg.norm(g)


# ---------------------------------------------------------------------------
# Nonuniform DFT (the transform FINUFFT's type 1 approximates, the 2D lateral
# sum of mod_ClosedFormField.jl:743-780): pinned by the uniform-grid case,
# where it is a plain DFT (numpy.fft), and by a brute-force sum.
# ---------------------------------------------------------------------------
def test_nudft_on_uniform_points_is_the_dft():
    N1, N2 = 12, 10
    m1, m2 = np.meshgrid(np.arange(N1), np.arange(N2), indexing="ij")
    x = 2 * np.pi * m1.ravel() / N1 - np.pi
    y = 2 * np.pi * m2.ravel() / N2 - np.pi
    rng = np.random.default_rng(4)
    c = rng.normal(size=x.size) + 1j * rng.normal(size=x.size)
    f = po.nudft2d1(x, y, c, N1=N1, N2=N2)
    # sum_m c_m e^{i k (2 pi m / N - pi)} = (-1)^k N1 N2 ifft2(c)[k mod N]
    F = np.fft.ifft2(c.reshape(N1, N2)) * N1 * N2
    k1 = np.arange(N1) - N1 // 2
    k2 = np.arange(N2) - N2 // 2
    ref = F[np.ix_(k1 % N1, k2 % N2)] * ((-1.0) ** k1)[:, None] * ((-1.0) ** k2)[None, :]
    assert relerr(f, ref) < 1e-13


def test_nudft_against_brute_force_and_linearity():
    rng = np.random.default_rng(5)
    x, y = rng.uniform(-np.pi, np.pi, 40), rng.uniform(-np.pi, np.pi, 40)
    c = rng.normal(size=40) + 1j * rng.normal(size=40)
    modes = np.array([[0, 0], [3, -2], [-5, 7], [11, 1]])
    f = po.nudft2d1(x, y, c, modes=modes)
    for m, (a, b) in enumerate(modes):
        assert abs(f[m] - sum(c[j] * np.exp(1j * (a * x[j] + b * y[j])) for j in range(40))) < 1e-12
    f2 = po.nudft2d1(x, y, (2 - 1j) * c, modes=modes)
    assert relerr(f2, (2 - 1j) * f) < 1e-14


# ---------------------------------------------------------------------------
# Lateral assembly: the linear convolution of the operator
# (mod_ClosedFormAssembly.jl:10, mod_GreenOperator.jl:674-712)
# ---------------------------------------------------------------------------
def test_linear_convolution_reproduces_free_space_fields():
    """Dipoles on pixels of a small grid in a homogeneous medium: the oracle's
    2D closed_form_G on the offset grid, convolved with the moments, equals
    the sum of the free-space dipole fields at every observation pixel and
    plane.  G_xz and G_yz are odd in the lateral offset, so a reversed offset
    or a swapped x/y axis fails this."""
    wf, ef = 2 * np.pi, 1.0 + 1e-4j
    stack = [lay(IsotropicMaterial(ef), 3.0) for _ in range(3)]
    st = ostack(stack)
    k = wf * np.sqrt(ef)
    kr, wr = kr_rule([wf * np.sqrt(ef).real], kmax=36 / 0.7 + wf, width=0.5, order=16)
    phis, wph = uniform_phi(48)
    kx, ky, W = polar_nodes(kr, wr, phis, wph)
    Nx, Ny, dx, dy = 4, 3, 0.2, 0.15
    zs, zo = [1.0, 1.1], [1.8, 1.95]
    offx = (np.arange(2 * Nx - 1) - (Nx - 1)) * dx
    offy = (np.arange(2 * Ny - 1) - (Ny - 1)) * dy
    G = np.stack([po.closed_form_G_2d(st, wf, kx, ky, W, 2, 2, z, zo, offx, offy) for z in zs], axis=2)
    rng = np.random.default_rng(3)
    P = np.zeros((Nx, Ny, 2, 3), np.complex128)
    dips = [(0, 2, 0), (3, 1, 1), (2, 0, 0)]  # (i, j, source plane)
    for (i, j, kk) in dips:
        P[i, j, kk] = rng.normal(size=3) + 1j * rng.normal(size=3)
    E = po.linear_convolution(G, P)
    Eref = np.zeros_like(E)
    for i in range(Nx):
        for j in range(Ny):
            for o, z in enumerate(zo):
                for (a, b, kk) in dips:
                    r = [(i - a) * dx, (j - b) * dy, z - zs[kk]]
                    Eref[i, j, o] += E_free(r, P[a, b, kk], k, ef)
    assert relerr(E, Eref) < 1e-9, relerr(E, Eref)


def test_padded_fft_route_is_the_linear_convolution():
    """The oracle's FFT route of the padded assembly (Ĝ from G on the 2Nx x 2Ny
    offsets, zero-padded source, bfft, crop) equals the literal sum over
    source pixels for an arbitrary G: the centring sign, the padding and the
    crop are pinned independently of the physics."""
    rng = np.random.default_rng(5)
    Nx, Ny, Nz, dx, dy = 6, 4, 2, 0.3, 0.2
    Px, Py = 2 * Nx, 2 * Ny
    G = rng.normal(size=(Px, Py, Nz, 3, 3)) + 1j * rng.normal(size=(Px, Py, Nz, 3, 3))
    J = rng.normal(size=(Nx, Ny, 3)) + 1j * rng.normal(size=(Nx, Ny, 3))
    E = po.apply_ghat(po.ghat_from_G(G, dx, dy), J)
    # offsets -(N-1)..(N-1) sit at padded index offset + N (xs[N] = 0)
    Goff = G[1:, 1:][:, :, None]  # (2Nx-1, 2Ny-1, Ns=1, Nz, 3, 3)
    Eref = po.linear_convolution(Goff, (J * dx * dy)[:, :, None, :])
    assert relerr(E, Eref) < 1e-13, relerr(E, Eref)


def test_zskeleton_of_a_family_of_known_rank():
    """r distinct exponentials on N planes span an r-dimensional space: the
    skeleton has exactly r planes, interpolates every row to rounding, and
    P restricted to the skeleton planes is the identity; a random family
    (full rank) is kept whole."""
    z = np.linspace(0.0, 1.3, 20)
    Q = np.array([1.0 + 0.01j, 2.5 + 0.2j, 4.0 + 0.05j])
    sk = po.zskeleton(z, [(Q, 1.0, 0.0)], tol=1e-10)
    assert not sk["whole"] and len(sk["idx"]) == 3 and sk["err"] < 1e-12
    assert np.allclose(sk["P"][sk["idx"]], np.eye(3), atol=1e-13)
    A = po.family_rows([(Q, 1.0, 0.0)], z)
    assert np.max(np.abs(A - A[:, sk["idx"]] @ sk["P"].T)) < 1e-12
    # two families with distinct (sigma, c) share the planes: rank 3 + 2
    sk2 = po.zskeleton(z, [(Q, 1.0, 0.0), (np.array([3.0 + 0.1j, 0.5 + 0.0j]), -1.0, 0.4)], tol=1e-10)
    assert len(sk2["idx"]) == 5
    rng = np.random.default_rng(0)
    sk3 = po.zskeleton(z, [(rng.normal(size=400) * 30 + 1j * rng.uniform(0, 0.1, 400), 1.0, 0.0)], tol=1e-10)
    assert sk3["whole"]
    # a looser tolerance gives a smaller skeleton within that tolerance
    dense_Q = np.linspace(0.1, 6.0, 200) + 0.01j
    a = po.zskeleton(z, [(dense_Q, 1.0, 0.0)], tol=1e-9)
    b = po.zskeleton(z, [(dense_Q, 1.0, 0.0)], tol=1e-4)
    assert len(b["idx"]) < len(a["idx"]) and b["err"] <= 1e-4


# ---------------------------------------------------------------------------
# Box sources (mod_ClosedFormField.jl:86-98 _boxphase, 815-862 interior slab
# and local Ez)
# ---------------------------------------------------------------------------
def _slab_average(st, kx, ky, src, obs, zc, dz, zo, split=None, n=24, terms="all"):
    """(1/dz) times the integral of the point-source g~ over source depths in
    [zc - dz/2, zc + dz/2] (Gauss-Legendre, split at `split`)."""
    x, w = np.polynomial.legendre.leggauss(n)
    edges = [zc - dz / 2, zc + dz / 2] if split is None else [zc - dz / 2, split, zc + dz / 2]
    zs, ws = [], []
    for a, b in zip(edges[:-1], edges[1:]):
        zs += list((b - a) / 2 * x + (a + b) / 2)
        ws += list((b - a) / 2 * w)
    g = po.ghat_batch(st, W0, kx, ky, src, obs, zs, [zo] * len(zs), terms=terms)
    return np.tensordot(g, np.array(ws), axes=([1], [0])) / dz


@pytest.mark.parametrize("label,src,obs,zc,dz,zo", [
    ("iso 5-layer lossy", 2, 2, 0.35, 0.2, 0.6),    # same layer, exterior (forward)
    ("iso 5-layer lossy", 2, 2, 0.35, 0.2, 0.1),    # same layer, exterior (backward)
    ("uniaxial 4-layer", 2, 3, 0.4, 0.3, 0.25),     # another layer
    ("gyrotropic 4-layer", 2, 2, 0.3, 0.2, 0.5),
    ("iso 5-layer lossy", 1, 1, 0.3, 0.2, 0.2),     # source layer 1 (reversed depth)
])
def test_box_equals_the_slab_average_of_point_sources_outside(label, src, obs, zc, dz, zo):
    st = ostack(FOUNDATION_STACKS[label])
    kx, ky = np.array([0.3, 2.2, 5.5, 9.0]), np.array([0.1, -0.7, 1.5, 4.0])
    box = po.ghat_batch(st, W0, kx, ky, src, obs, [zc], [zo], dz_box=dz)[:, 0]
    avg = _slab_average(st, kx, ky, src, obs, zc, dz, zo)
    assert relerr(box, avg) < 1e-12, relerr(box, avg)


@pytest.mark.parametrize("zo", [0.3, 0.36, 0.42])
def test_box_interior_ramps_equal_the_split_slab_average(zo):
    """Inside the slab the direct term is the face ramps: the average of the
    point sources below (forward) and above (backward) the observation plane."""
    st = ostack(FOUNDATION_STACKS["iso 5-layer lossy"])
    kx, ky = np.array([0.3, 2.2, 5.5, 9.0]), np.array([0.1, -0.7, 1.5, 4.0])
    zc, dz = 0.35, 0.2
    box = po.ghat_batch(st, W0, kx, ky, 2, 2, [zc], [zo], dz_box=dz, ez_local=False)[:, 0]
    avg = _slab_average(st, kx, ky, 2, 2, zc, dz, zo, split=zo)
    assert relerr(box, avg) < 1e-12, relerr(box, avg)
    # the local Ez is -p_z / (eps_zz dz) on the zz entry only
    with_ez = po.ghat_batch(st, W0, kx, ky, 2, 2, [zc], [zo], dz_box=dz)[:, 0]
    d = with_ez - box
    ezz = FOUNDATION_STACKS["iso 5-layer lossy"][1].eps.eps[2, 2]
    assert np.allclose(d[:, 2, 2], -1 / (ezz * dz), rtol=1e-14)
    d[:, 2, 2] = 0
    assert np.max(np.abs(d)) == 0


def test_box_direct_term_is_continuous_at_the_faces_and_local_ez_cancels_the_flat_part():
    st = ostack(FOUNDATION_STACKS["iso 5-layer lossy"])
    kx, ky = np.array([0.3, 2.2, 5.5]), np.array([0.1, -0.7, 1.5])
    zc, dz = 0.35, 0.2
    for face in (zc + dz / 2, zc - dz / 2):
        out = po.ghat_batch(st, W0, kx, ky, 2, 2, [zc], [face], dz_box=dz, terms="direct", ez_local=False)[:, 0]
        eps = 1e-9 * np.sign(zc - face)
        inn = po.ghat_batch(st, W0, kx, ky, 2, 2, [zc], [face + eps], dz_box=dz, terms="direct",
                            ez_local=False)[:, 0]
        assert relerr(inn, out) < 1e-7, relerr(inn, out)
    # deep in the evanescent range the modal zz part inside the slab tends to
    # +1/(eps_zz dz): only its sum with the local Ez decays (like 1/k_rho^2)
    ezz = FOUNDATION_STACKS["iso 5-layer lossy"][1].eps.eps[2, 2]
    k = np.array([200.0, 400.0, 800.0])
    modal = po.ghat_batch(st, W0, k, 0 * k, 2, 2, [zc], [zc + 0.01], dz_box=dz, ez_local=False, terms="direct")[:, 0]
    total = po.ghat_batch(st, W0, k, 0 * k, 2, 2, [zc], [zc + 0.01], dz_box=dz, terms="direct")[:, 0]
    assert np.allclose(modal[:, 2, 2] * ezz * dz, 1.0, rtol=2e-2)
    r = np.abs(total[:, 2, 2]) / np.abs(modal[:, 2, 2])
    assert r[0] < 5e-3 and r[1] < r[0] / 3 and r[2] < r[1] / 3, r


def test_box_must_fit_its_layer():
    st = ostack(FOUNDATION_STACKS["iso 5-layer lossy"])
    with pytest.raises(ValueError, match="fit"):
        po.ghat_batch(st, W0, [1.0], [0.0], 2, 2, [0.05], [0.3], dz_box=0.2)


def test_operator_z_coupling_against_the_definition():
    """ora_op_zmix (the z coupling of apply_operator, every block form) against
    its definition summed plane pair by plane pair in numpy, on ragged sizes
    (node count not a multiple of the chunk, No != Ns)."""
    rng = np.random.default_rng(5)

    def c(*s):
        return rng.normal(size=s) + 1j * rng.normal(size=s)
    ns, no, nxy, rs, ro = 5, 4, 301, 3, 2
    J = c(ns, 3, nxy)
    T, H = c(ns + no - 1, 3, 3, nxy), c(ns + no - 1, 3, 3, nxy)
    D, Dk, Ps, Po = c(ns * no, 3, 3, nxy), c(rs * ro, 3, 3, nxy), c(ns, rs), c(no, ro)
    mv = lambda K, v: np.einsum("cqx,qx->cx", K, v)
    ref_s = np.array([sum(mv(T[j - i + ns - 1] + H[i + j], J[i]) for i in range(ns)) for j in range(no)])
    ref_d = np.array([sum(mv(D[i + ns * j], J[i]) for i in range(ns)) for j in range(no)])
    Js = np.einsum("ia,iqx->aqx", Ps, J)
    Es = np.array([sum(mv(Dk[a + rs * b], Js[a]) for a in range(rs)) for b in range(ro)])
    ref_k = np.einsum("jb,bcx->jcx", Po, Es)
    for got, ref in ((po.op_zmix(nxy, J, no, T=T, H=H), ref_s), (po.op_zmix(nxy, J, no, D=D), ref_d),
                     (po.op_zmix(nxy, J, no, D=Dk, Ps=Ps, Po=Po), ref_k)):
        assert relerr(got, ref) < 1e-14
    # accumulation into a given E, and the Toeplitz part alone
    E0 = c(no, 3, nxy)
    got = po.op_zmix(nxy, J, no, T=T, E=E0.copy())
    ref = E0 + np.array([sum(mv(T[j - i + ns - 1], J[i]) for i in range(ns)) for j in range(no)])
    assert relerr(got, ref) < 1e-14
