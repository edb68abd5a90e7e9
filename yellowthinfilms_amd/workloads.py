"""Named workloads of the benchmark configurations (units: lambda = 1, so
omega = k0 = 2 pi; 1 nm = 1/600 at lambda = 600 nm).

batched_gyro_512: a 5-layer stack with a 40 nm gyrotropic film
  (eps = [[3, ig, 0], [-ig, 3, 0], [0, 0, 2.8]] + 0.02i, g = 0.3) on the uniform
  512 x 512 (kx, ky) grid of uniform_kxky_grid(L = 16, dx = 1/16) (kmax = 8 k0):
  one Green k-node (Berreman modes of every layer, S-matrix cascade, T^o,
  seed, basis, spectral dyadic) per grid node, then the FFT convolution of a
  512 x 512 x 3 complex N(0,1) source.  Source: layer 1 (air), 20 nm above the
  film; observation: layer 1, 40 nm above the film.
"""
import math

import numpy as np

from .materials import IsotropicMaterial, UniaxialMaterial, GyrotropicMaterial, FlatOpticalLayer

NM = 1.0 / 600.0
OMEGA = 2 * math.pi


def gyro_512_stack():
    gy = np.array([[3 + 0.02j, 0.3j, 0], [-0.3j, 3 + 0.02j, 0], [0, 0, 2.8 + 0.02j]])
    return [FlatOpticalLayer(IsotropicMaterial(1 + 0.02j), 1.0),
            FlatOpticalLayer(IsotropicMaterial(2.25 + 0.02j), 60 * NM),
            FlatOpticalLayer(GyrotropicMaterial(gy), 40 * NM),
            FlatOpticalLayer(IsotropicMaterial(4.0 + 0.02j), 80 * NM),
            FlatOpticalLayer(IsotropicMaterial(2.25 + 0.02j), 1.0)]


BATCHED_GYRO_512 = dict(
    name="batched_gyro_512",
    description=("5-layer stack with a 40 nm gyrotropic film (g=0.3, Im eps=0.02), uniform 512x512 kx,ky grid "
                 "(L=16, dx=1/16, kmax=8k0), one Berreman eig + S-matrix cascade + T^o + seed + spectral dyadic "
                 "per node (262144 Green k-nodes), then the FFT convolution of a 512x512x3 N(0,1) complex source; "
                 "complex FP64"),
    L=16.0, dx=1.0 / 16.0,
    source_layer=1, zsrc=20 * NM,          # Dipole.z convention: height above the film
    obs_layer=1, zobs=1.0 - 40 * NM,       # closed_form_G z_obs convention (layer-1 local depth)
    seed=1,
)

# Algorithmic FP64 operations per Green k-node of batched_gyro_512 (node solve
# + one spectral dyadic), counted by running the oracle's reference-order
# arithmetic with an op-counting real type over 512 sampled grid nodes:
# real adds + multiplies + divides (sqrt and exp/sin/cos/atan2 reported
# separately).  Pinned by tests/test_workloads.py.
BATCHED_GYRO_512_FLOPS_PER_NODE = 32350.0
BATCHED_GYRO_512_SPECIAL_PER_NODE = 278.0
# HBM bytes per launch of the node kernel on batched_gyro_512 (rocprofv3
# FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, scales from a 256 MiB device copy):
# 6.03 MB read (kx, ky, stack) + 37.75 MB written (g~).
BATCHED_GYRO_512_NODE_HBM_BYTES = 43.78e6


def uniform_c128_source(nx, ny, seed):
    """Seeded complex N(0,1) source (re, im each N(0, 1/2)) laid out as the
    reference's column-major (Nx, Ny, 3) array = contiguous (3, Ny, Nx)."""
    rng = np.random.default_rng(seed)
    return ((rng.standard_normal((3, ny, nx)) + 1j * rng.standard_normal((3, ny, nx))) / math.sqrt(2.0))
