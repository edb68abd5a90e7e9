"""Host-side logic and the C-ABI library, without a GPU: data layouts, the
uniform grid, library loading and exports, argument validation, and the
transform-pair algebra of the spectral operator (against numpy.fft)."""
import ctypes as ct
import math
import os
import re

import numpy as np
import pytest

from helpers import ROOT, po, relerr, c2_stack, ostack, NM
from yellowthinfilms_amd import _lib
from yellowthinfilms_amd.green import layer_structs, uniform_kxky_grid, fft_wavenumbers
from yellowthinfilms_amd.materials import (IsotropicMaterial, UniaxialMaterial, BiaxialMaterial, GyrotropicMaterial,
                                           FlatOpticalLayer, FlatLayer, stack_arrays)


def _header_functions():
    src = open(os.path.join(ROOT, "include", "ytf.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(ytf_[A-Za-z_0-9]+)\s*\(", src)))


def test_library_loads_and_exports_every_header_symbol():
    L = _lib.lib()
    names = _header_functions()
    assert "ytf_green_nodes" in names and "ytf_spectral_apply" in names
    for n in names:
        assert hasattr(L, n), n
    # and the ctypes binding covers the header exactly
    assert sorted(_lib.EXPORTS) == names
    assert b"gfx950" in L.ytf_version()


def test_layer_struct_layout():
    assert ct.sizeof(_lib.YtfLayer) == 304
    g = np.array([[3 + 0.02j, 0.3j, 0.1], [-0.3j, 3 + 0.02j, 0.2], [0.4, 0.5, 2.8 + 0.02j]])
    st = [FlatOpticalLayer(GyrotropicMaterial(g), 0.25),
          FlatLayer(UniaxialMaterial(2.0, 3.0), BiaxialMaterial(1.1, 1.2, 1.3), 0.5)]
    arr = layer_structs(st)
    # column-major interleaved: eps[2*(i + 3j)] = Re g[i, j]
    for i in range(3):
        for j in range(3):
            assert arr[0].eps[2 * (i + 3 * j)] == g[i, j].real
            assert arr[0].eps[2 * (i + 3 * j) + 1] == g[i, j].imag
    assert arr[0].d == 0.25 and arr[0].eps_kind == 3 and arr[0].mu_kind == 0
    assert arr[1].eps_kind == 1 and arr[1].mu_kind == 1
    assert arr[1].mu[2 * 4] == 1.2  # mu[1,1]


def test_stack_arrays_kinds():
    st = [FlatOpticalLayer(IsotropicMaterial(2.0), 1.0), FlatOpticalLayer(UniaxialMaterial(2.0, 3.0), 0.1),
          FlatOpticalLayer(BiaxialMaterial(1, 2, 3), 0.1), FlatLayer(IsotropicMaterial(1.0), IsotropicMaterial(2.0), 1)]
    eps, mu, d, ek, mk = stack_arrays(st)
    assert list(ek) == [0, 1, 2, 0] and list(mk) == [0, 0, 0, 0]
    assert np.allclose(eps[1], np.diag([2, 2, 3])) and np.allclose(mu[3], 2 * np.eye(3))


def test_uniform_grid_relations():
    # the DFT relations uniform_kxky_grid encodes (mod_UniformSource.jl:52-88)
    g = uniform_kxky_grid(16.0, 1 / 16)
    assert g["Nx"] == 512 and g["Ny"] == 512
    assert abs(g["dkx"] * g["dx"] * g["Nx"] - 2 * math.pi) < 1e-12
    assert abs(g["kx_max"] - 8 * 2 * math.pi) < 1e-12
    assert g["midx"] == 257
    g2 = uniform_kxky_grid(1.0, 0.3)
    assert g2["Nx"] == 7 and abs(g2["dx"] - 2 / 7) < 1e-15
    for N in (8, 512):
        assert np.allclose(fft_wavenumbers(N, 0.5), 0.5 * N * np.fft.fftfreq(N))
    # odd N: the reference's order (mod_UniformSource.jl:127-130) wraps at
    # fld(N, 2), i.e. -ceil(N/2) dk instead of fftfreq's +floor(N/2) dk
    assert np.allclose(fft_wavenumbers(7, 0.5), [0, 0.5, 1.0, -2.0, -1.5, -1.0, -0.5])


def test_argument_validation_without_gpu():
    L = _lib.lib()
    st = layer_structs(c2_stack())
    z = (ct.c_double * 1)(0.0)
    rc = L.ytf_green_nodes(None, st, 5, 2 * math.pi, None, None, 0, 1, 1, 0, z, z, None, None, 0)
    assert rc == _lib.YTF_EINVAL
    assert b"ctx is NULL" in L.ytf_last_error(None)
    rc = L.ytf_spectral_apply(None, 512, 512, None, None, None, None, 1.0)
    assert rc == _lib.YTF_EINVAL


oracle_spectral_apply = po.spectral_apply


def test_spectral_pair_round_trip_is_identity():
    # identity kernel: E = J exactly (the round-trip identity the grid is built for)
    grid = uniform_kxky_grid(1.0, 1 / 16)
    rng = np.random.default_rng(3)
    J = rng.normal(size=(grid["Nx"], grid["Ny"], 3)) + 1j * rng.normal(size=(grid["Nx"], grid["Ny"], 3))
    g = np.zeros((grid["Nx"], grid["Ny"], 3, 3), complex)
    for c in range(3):
        g[:, :, c, c] = 1
    assert relerr(oracle_spectral_apply(g, J, grid), J) < 1e-14


def test_spectral_apply_is_circular_convolution():
    # small grid: the spectral route equals the direct circular convolution
    # with the kernel G(x) = (1/(dx dy Nx Ny)) sum_k g(k) e^{ikx} (brute force)
    grid = uniform_kxky_grid(0.25, 1 / 16)
    Nx, Ny = grid["Nx"], grid["Ny"]
    rng = np.random.default_rng(4)
    g = rng.normal(size=(Nx, Ny, 3, 3)) + 1j * rng.normal(size=(Nx, Ny, 3, 3))
    J = rng.normal(size=(Nx, Ny, 3)) + 1j * rng.normal(size=(Nx, Ny, 3))
    E = oracle_spectral_apply(g, J, grid)
    G = np.stack([[np.fft.ifft2(g[:, :, c, q]) for q in range(3)] for c in range(3)])  # (3,3,Nx,Ny)
    Eb = np.zeros_like(J)
    for x in range(Nx):
        for y in range(Ny):
            sh = np.roll(np.roll(J, x, axis=0), y, axis=1)  # J[(i - x), (j - y)]
            Eb += np.einsum("cq,ijq->ijc", G[:, :, x, y], sh)
    assert relerr(E, Eb) < 1e-12


def test_rotation_fill_is_refused_on_a_stack_without_symmetry():
    """build_cavity_transfer_2d(...; rotation=true) on a stack with a biaxial
    layer errors before any device work (mod_CavityTransfer.jl:322-324)."""
    from helpers import FOUNDATION_STACKS
    from yellowthinfilms_amd.green_operator import CavityTransfer2D
    with pytest.raises(ValueError, match="requires an azimuthally symmetric stack"):
        CavityTransfer2D(FOUNDATION_STACKS["biaxial 4-layer"], 6.28, 2, [1.0], [1.0], [0.0], [1.0], rotation=True)
    with pytest.raises(ValueError, match="φ nodes/weights length mismatch"):
        CavityTransfer2D(FOUNDATION_STACKS["iso 3-layer vac/glass/vac"], 6.28, 2, [1.0], [1.0], [0.0], [1.0, 2.0])


def test_ktable_needs_a_rotation_invariant_stack():
    """The tabulated k_rho surrogate accepts isotropic, uniaxial-about-z and
    gyrotropic-about-z layers and refuses the rest before any device work."""
    from helpers import FOUNDATION_STACKS, c2_stack
    from yellowthinfilms_amd.ktable import KrhoTable, rotation_invariant
    assert rotation_invariant(FOUNDATION_STACKS["uniaxial 4-layer"]) and rotation_invariant(c2_stack())
    assert not rotation_invariant(FOUNDATION_STACKS["biaxial 4-layer"])
    with pytest.raises(ValueError, match="invariant under rotations about z"):
        KrhoTable(FOUNDATION_STACKS["biaxial 4-layer"], 6.28, 20.0)
