"""yellowthinfilms_amd -- MI355X (gfx950) implementation of the layered-media
Green's-function hot path of HortenseBeaudoin/YellowThinFilms.

The compute runs in the hand-written HIP kernels of libytf.so (C-ABI in
include/ytf.h); this package is the host-side mirror of the reference's Julia
interface for that path (material/layer types, the k-node cache build, the
uniform-grid operator build/apply), driving the C-ABI through ctypes with
torch device buffers.
"""
from .materials import (IsotropicMaterial, UniaxialMaterial, BiaxialMaterial, GyrotropicMaterial,
                        FlatOpticalLayer, FlatLayer, stack_arrays)
from .green import (green_nodes, build_cavity_transfer, uniform_kxky_grid, fft_wavenumbers,
                    SpectralGreenOperator, NodeCache)
from .synthesis import CavityTransfer1D, closed_form_G, radial_synthesis
from .nufft import nufft2d1
from . import quadrature

__all__ = [
    "IsotropicMaterial", "UniaxialMaterial", "BiaxialMaterial", "GyrotropicMaterial", "FlatOpticalLayer",
    "FlatLayer", "stack_arrays", "green_nodes", "build_cavity_transfer", "uniform_kxky_grid", "fft_wavenumbers",
    "SpectralGreenOperator", "NodeCache", "CavityTransfer1D", "closed_form_G", "radial_synthesis", "quadrature", "nufft2d1",
]
