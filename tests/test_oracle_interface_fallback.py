"""The oracle's interface solve on an exactly singular system: the pivoted-QR
fallback of _S_Interface_Matrix (mod_ScatteringMatrix.jl:35-44).

Julia's `qr(L, ColumnNorm()) \\ R` returns the minimum-norm least-squares
solution of the rank-truncated system, i.e. pinv(L) R whenever the rank gap is
clear; numpy's pinv (an SVD) is the independent check.  The nodes that reach
this path in a real stack: a lossless layer at k_rho equal to its wavenumber
exactly (k_z = 0), where the forward and backward mode vectors of each
polarisation coincide, next to closed-form layers."""
import math

import numpy as np
import pytest

from helpers import po, lay, NM, IsotropicMaterial, UniaxialMaterial, GyrotropicMaterial, FlatLayer, ostack

W0 = 2 * math.pi


@pytest.mark.parametrize("rank", [0, 1, 2, 3, 4])
def test_qr_fallback_is_the_minimum_norm_solution(rank):
    rng = np.random.default_rng(11 + rank)
    for _ in range(4):
        U = rng.normal(size=(4, rank)) + 1j * rng.normal(size=(4, rank))
        V = rng.normal(size=(rank, 4)) + 1j * rng.normal(size=(rank, 4))
        L = U @ V if rank else np.zeros((4, 4), complex)
        R = rng.normal(size=(4, 4)) + 1j * rng.normal(size=(4, 4))
        X = po.qr_solve4(L, R)
        ref = np.linalg.pinv(L) @ R
        assert np.abs(X - ref).max() <= 1e-13 * max(np.abs(ref).max(), 1.0)


def test_lu_is_kept_when_no_pivot_is_exactly_zero():
    rng = np.random.default_rng(2)
    L = rng.normal(size=(4, 4)) + 1j * rng.normal(size=(4, 4))
    R = rng.normal(size=(4, 4)) + 1j * rng.normal(size=(4, 4))
    S, fb = po.interface_solve(L, R)
    assert not fb
    assert np.abs(S - np.linalg.solve(L, R)).max() < 1e-13


def singular_stack():
    """Glass | air | (eps 2, mu 0.5) | uniaxial film | gyrotropic film |
    substrate.  The air and the (eps 2, mu 0.5) layer share the index 1 with
    different impedances; at k_x = omega, k_y = 0 both have k_z = 0 exactly,
    each one's forward and backward vectors of a polarisation coincide, and the
    three interfaces they touch (the glass and the uniaxial film are closed-form
    too, so TE and TM decouple with exact zeros) meet exactly zero LU pivots and
    take the fallback; the gyrotropic film's interfaces do not.  (Two identical
    layers would do as well, but at evanescent nodes the continuity matching of
    identical layers compares two overlaps that vanish, so the reference's mode
    order there is decided by rounding.)"""
    gy = np.array([[3 + 0.02j, 0.3j, 0], [-0.3j, 3 + 0.02j, 0], [0, 0, 2.8 + 0.02j]])
    return [lay(IsotropicMaterial(2.25), 1.0), lay(IsotropicMaterial(1.0), 50 * NM),
            FlatLayer(IsotropicMaterial(2.0), IsotropicMaterial(0.5), 70 * NM),
            lay(UniaxialMaterial(2.5 + 0.01j, 3.0 + 0.01j), 60 * NM), lay(GyrotropicMaterial(gy), 40 * NM),
            lay(IsotropicMaterial(4.0 + 0.02j), 1.0)]


def test_index_matched_layers_at_their_light_line_take_the_fallback():
    """At k_x = omega, k_y = 0 the interfaces of the two k_z = 0 layers are
    exactly singular ([-Vf_B  Vb_A] has rank 3 or 2), the LU meets a zero pivot
    and the interface S is pinv(L) R."""
    M, Sint, _, _, _ = po.stack_smatrix(ostack(singular_stack()), W0, W0, 0.0)
    assert np.all(M[1]["Qf"] == 0) and np.all(M[2]["Qb"] == 0)
    ranks, fell = [], []
    for j in range(len(M) - 1):
        L = np.hstack([-M[j + 1]["Vf"], M[j]["Vb"]])
        R = np.hstack([-M[j]["Vf"], M[j + 1]["Vb"]])
        S, fb = po.interface_solve(L, R)
        ranks.append(int(np.linalg.matrix_rank(L)))
        fell.append(fb)
        if fb:
            ref = np.linalg.pinv(L) @ R
            assert np.abs(S - ref).max() < 1e-14 * max(np.abs(ref).max(), 1.0), j
        assert np.abs(Sint[j] - S).max() == 0.0, j
    assert fell == [True, True, True, False, False]
    assert ranks == [3, 2, 3, 4, 4]
    # every node of the stack around it stays valid and finite in the cache
    kx = np.array([W0, 0.5 * W0, 1.3 * W0])
    ky = np.array([0.0, 0.2 * W0, -0.4 * W0])
    for src, obs in ((1, 1), (1, 6), (5, 1), (5, 5), (4, 4)):
        c = po.nodes_batch(ostack(singular_stack()), W0, kx, ky, src, [obs])
        assert c.valid.all()
        g = po.ghat_from_cache(c, obs, [0.01], [0.02])
        assert np.isfinite(g).all()
