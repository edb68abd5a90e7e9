"""Named workloads of the benchmark configurations (units: lambda = 1, so
omega = k0 = 2 pi; 1 nm = 1/600 at lambda = 600 nm).

batched_gyro_512: a 5-layer stack with a 40 nm gyrotropic film
  (eps = [[3, ig, 0], [-ig, 3, 0], [0, 0, 2.8]] + 0.02i, g = 0.3) on the uniform
  512 x 512 (kx, ky) grid of uniform_kxky_grid(L = 16, dx = 1/16) (kmax = 8 k0):
  one Green k-node (Berreman modes of every layer, S-matrix cascade, T^o,
  seed, basis, spectral dyadic) per grid node, then the FFT convolution of a
  512 x 512 x 3 complex N(0,1) source.  Source: layer 1 (air), 20 nm above the
  film; observation: layer 1, 40 nm above the film.

point_uniaxial_1d: air / 50 nm uniaxial film eps = diag(2.5, 2.5, 3.0) + 0.01i /
  glass 2.25; a dipole 10 nm above the film; 4096 Gauss-Legendre k-rho nodes on
  [0, 6 k0] in panels broken at the light lines (u^2 substitution at the
  lossless ones) plus a 512-node cosh tail; the J0/J1/J2 synthesis at 256
  radii in (0, 2] and 8 observation heights in the air, complex FP64.

bessel_gemm: the same stack; 16384 k-rho nodes (14336 in the band + a
  2048-node tail), 8192 radii in (0, 2], 64 depth pairs (8 source heights x 8
  observation heights in the air): the J0/J1/J2 contraction as a real x complex
  GEMM on the matrix cores with split fp16 products (mode solves FP64).

uneven_256: 256 independent stacks, each with U{2..24} layers of thickness
  U[5, 200] nm and random hyperbolic diagonal permittivities (each diagonal
  entry +-U[1, 6] + i U[0.01, 0.5], at least one negative entry per stack),
  and U{1024..16384} nodes kt = k_rho (cos phi, sin phi), k_rho ~ U[0, 8 k0],
  phi ~ U[0, 2 pi): modes + total S-matrix per node, FP64, one persistent launch.
"""
import math

import numpy as np

from .materials import (IsotropicMaterial, UniaxialMaterial, BiaxialMaterial, GyrotropicMaterial,
                        FlatOpticalLayer)

NM = 1.0 / 600.0
OMEGA = 2 * math.pi


def gyro_512_stack():
    gy = np.array([[3 + 0.02j, 0.3j, 0], [-0.3j, 3 + 0.02j, 0], [0, 0, 2.8 + 0.02j]])
    return [FlatOpticalLayer(IsotropicMaterial(1 + 0.02j), 1.0),
            FlatOpticalLayer(IsotropicMaterial(2.25 + 0.02j), 60 * NM),
            FlatOpticalLayer(GyrotropicMaterial(gy), 40 * NM),
            FlatOpticalLayer(IsotropicMaterial(4.0 + 0.02j), 80 * NM),
            FlatOpticalLayer(IsotropicMaterial(2.25 + 0.02j), 1.0)]


def point_uniaxial_stack():
    return [FlatOpticalLayer(IsotropicMaterial(1.0), 1.0),
            FlatOpticalLayer(UniaxialMaterial(2.5 + 0.01j, 3.0 + 0.01j), 50 * NM),
            FlatOpticalLayer(IsotropicMaterial(2.25), 1.0)]


POINT_UNIAXIAL_1D = dict(
    name="point_uniaxial_1d",
    description=("air / 50 nm uniaxial film (eps diag(2.5,2.5,3.0)+0.01i) / glass 2.25, lambda=600 nm, dipole 10 nm "
                 "above the film; 4096 GL k-rho nodes on [0, 6k0] broken at the light lines + 512-node cosh tail; "
                 "J0/J1/J2 synthesis at 256 radii x 8 heights; complex FP64"),
    source_layer=1, zsrc=10 * NM,
    obs_layer=1, heights_nm=(20, 30, 40, 60, 80, 120, 160, 240),
    n_band=4096, order=16, kmax_k0=6.0, n_tail=512, n_rho=256, rho_max=2.0,
)


def point_uniaxial_rule(wl=POINT_UNIAXIAL_1D):
    """(kx, wt) of point_uniaxial_1d: the band [0, 6 k0] and the cosh tail
    (mod_KxGrid.jl:126-145, decay length = the smallest source-observation
    separation)."""
    from . import quadrature as Q
    k0 = OMEGA
    kmax = wl["kmax_k0"] * k0
    lossless = [k0, 1.5 * k0]                       # air, glass
    breaks = [0.0, kmax, k0, math.sqrt(2.5) * k0, 1.5 * k0]
    kb, wb = Q.panel_rule(breaks, wl["n_band"], order=wl["order"], singular=lossless)
    Leff = (min(wl["heights_nm"]) - wl["zsrc"] / NM) * NM
    umax = Q.evanescent_umax(kmax, Leff)
    kt, wt = Q.evanescent_band_nodes(kmax, umax, wl["n_tail"])
    return np.concatenate([kb, kt]), np.concatenate([wb, wt])


def point_uniaxial_geometry(wl=POINT_UNIAXIAL_1D):
    """(zobs local depths in layer 1, radii) of point_uniaxial_1d."""
    zobs = np.array([1.0 - h * NM for h in wl["heights_nm"]])
    rho = wl["rho_max"] * np.arange(1, wl["n_rho"] + 1) / wl["n_rho"]
    return zobs, rho


BESSEL_GEMM = dict(
    name="bessel_gemm",
    description=("J0/J1/J2 contraction of the air / 50 nm uniaxial film / glass stack: 8192 radii x 16384 k-rho "
                 "nodes x 64 depth pairs (8 source x 8 observation heights), real x complex GEMM on the matrix "
                 "cores, split fp16 products with FP32 accumulation; mode solves and g~ in FP64"),
    source_layer=1, obs_layer=1,
    src_heights_nm=(5, 10, 15, 20, 30, 40, 60, 80),
    obs_heights_nm=(20, 30, 40, 60, 80, 120, 160, 240),
    n_band=14336, order=16, kmax_k0=6.0, n_tail=2048, n_rho=8192, rho_max=2.0,
)


def bessel_gemm_rule(wl=BESSEL_GEMM):
    """(kx, wt): band [0, 6 k0] broken at the light lines, cosh tail with the
    decay length of the closest source/observation heights."""
    rule = dict(POINT_UNIAXIAL_1D, n_band=wl["n_band"], order=wl["order"], kmax_k0=wl["kmax_k0"],
                n_tail=wl["n_tail"], heights_nm=(min(wl["obs_heights_nm"]),),
                zsrc=max(h for h in wl["src_heights_nm"] if h < min(wl["obs_heights_nm"])) * NM)
    return point_uniaxial_rule(rule)


def bessel_gemm_geometry(wl=BESSEL_GEMM):
    """(zsrc, zobs) of every depth pair (source height major) and the radii."""
    zs = np.repeat([h * NM for h in wl["src_heights_nm"]], len(wl["obs_heights_nm"]))
    zo = np.tile([1.0 - h * NM for h in wl["obs_heights_nm"]], len(wl["src_heights_nm"]))
    rho = wl["rho_max"] * np.arange(1, wl["n_rho"] + 1) / wl["n_rho"]
    return zs, zo, rho


UNEVEN_256 = dict(
    name="uneven_256",
    description=("256 independent stacks, U{2..24} layers of U[5,200] nm with random hyperbolic diagonal eps "
                 "(+-U[1,6] + i U[0.01,0.5]), U{1024..16384} nodes each (k_rho ~ U[0,8k0], phi ~ U[0,2pi)): "
                 "Berreman modes + S-matrix cascade to the total S per node, FP64, one persistent launch"),
    n_stacks=256, layers=(2, 24), d_nm=(5.0, 200.0), eps_abs=(1.0, 6.0), eps_im=(0.01, 0.5),
    nodes=(1024, 16384), kmax_k0=8.0, seed=1,
)


def uneven_batch(wl=UNEVEN_256):
    """(stacks, node_counts, kx, ky) of the uneven batch, seeded."""
    rng = np.random.default_rng(wl["seed"])
    stacks, counts = [], []
    for _ in range(wl["n_stacks"]):
        nl = int(rng.integers(wl["layers"][0], wl["layers"][1] + 1))
        sg = rng.choice([-1.0, 1.0], size=(nl, 3))
        if not (sg < 0).any():
            sg[rng.integers(nl), rng.integers(3)] = -1.0
        re = sg * rng.uniform(*wl["eps_abs"], size=(nl, 3))
        im = rng.uniform(*wl["eps_im"], size=(nl, 3))
        d = rng.uniform(*wl["d_nm"], size=nl) * NM
        stacks.append([FlatOpticalLayer(BiaxialMaterial(*(re[l] + 1j * im[l])), d[l]) for l in range(nl)])
        counts.append(int(rng.integers(wl["nodes"][0], wl["nodes"][1] + 1)))
    n = sum(counts)
    kr = rng.uniform(0.0, wl["kmax_k0"] * OMEGA, n)
    ph = rng.uniform(0.0, 2 * math.pi, n)
    return stacks, counts, kr * np.cos(ph), kr * np.sin(ph)


SURROGATE_1024 = dict(
    name="surrogate_1024",
    description=("learned surrogate of the scattered spectral dyadic of a seeded 4-layer gyrotropic family "
                 "(g ~ U[0,0.5] drawn once): MLP 6->256x4->18 SiLU in BF16 with FP32 accumulation, trained on 2^22 "
                 "exact FP64 solver samples (kx, ky, z, z', omega, loss); evaluated on a 1024x1024 uniform grid "
                 "(|k| <= 3 omega) at z = 40 nm, z' = 50 nm, omega = 2pi, loss = 0.1, with the exact solve on "
                 "flagged nodes (light line, origin, cavity-response pole windows, large response), against the plain GPU "
                 "solve"),
    family_seed=1, seed=7, data_seed=2, width=256, depth=4, n_train=1 << 22, epochs=8, batch=8192, lr=2e-3,
    f=1.0, L=1024 / 12.0, dx=1.0 / 6.0, h_src_nm=40.0, h_obs_nm=50.0, loss=0.1,
    pole_window=8.0,  # nodes within this many half-widths of a cavity-response pole take the exact solve
    ktable_tol=1e-10, ktable_deg=16,  # the tabulated k_rho surrogate of the same stack
)


NUFFT_1E6 = dict(
    name="nufft_1e6",
    description=("2D type-1 nonuniform FFT (iflag +1) of 10^6 seeded points (x, y) ~ U[-pi, pi)^2 with complex "
                 "N(0,1) strengths onto 1024x1024 modes, eps 1e-9, FP64"),
    M=1_000_000, N=1024, eps=1e-9, seed=1,
)


def nufft_points(wl=NUFFT_1E6):
    rng = np.random.default_rng(wl["seed"])
    M = wl["M"]
    x = rng.uniform(-math.pi, math.pi, M)
    y = rng.uniform(-math.pi, math.pi, M)
    c = (rng.standard_normal(M) + 1j * rng.standard_normal(M)) / math.sqrt(2.0)
    return x, y, c


# Algorithmic FP64 operations of one uneven_256 launch: per stack, the mean
# op count (real adds + multiplies + divides) of Modes_Stack + the total S in
# the reference's formulation over 32 seeded sample nodes (rng seed 0),
# times the stack's node count.  Pinned by tests/test_workloads.py.
UNEVEN_256_FLOPS_PER_LAUNCH = 2.8317e11


OPERATOR_512 = dict(
    name="operator_512",
    description=("padded linear-convolution Green operator (apply_operator) of the batched_gyro_512 stack on its "
                 "512x512 grid (L=16, dx=1/16; G on the 1024x1024 padded offsets): one source plane (layer 1, "
                 "20 nm above the film) onto one observation plane (layer 1, 40 nm above the film), built from "
                 "262144 polar nodes (1024 Gauss-Legendre k_rho on [0, 8 k0] broken at the light lines x 256 "
                 "uniform phi) with the NUFFT lateral sum (eps 1e-9); one apply = the FFT convolution of a "
                 "512x512x3 N(0,1) complex moment sheet; complex FP64"),
    L=16.0, dx=1.0 / 16.0,
    source_layer=1, zsrc=20 * NM, obs_layer=1, zobs=1.0 - 40 * NM,
    n_krho=1024, n_phi=256, kmax_k0=8.0, eps=1e-9, seed=1,
)


def operator_nodes(wl=OPERATOR_512):
    """(k_rho, w_rho, phi, w_phi) of operator_512."""
    from . import quadrature as Q
    stack = gyro_512_stack()
    kmax = wl["kmax_k0"] * OMEGA
    breaks = [0.0, kmax] + [k for k in Q.light_lines(stack, OMEGA) if 0 < k < kmax]
    kr, wr = Q.panel_rule(breaks, wl["n_krho"], order=16)
    phi = 2 * math.pi * np.arange(wl["n_phi"]) / wl["n_phi"]
    return kr, wr, phi, np.full(wl["n_phi"], 2 * math.pi / wl["n_phi"])


OPERATOR_VOLUME_64 = dict(
    name="operator_volume_64",
    description=("z-structured Green operator of the reference's operator benchmark at size L "
                 "(Bench_GreenOperator.jl): iso vac/glass/vac, every layer 0.5 thick, Im eps = 1e-3 on every layer, "
                 "box source voxels of side 1/32 (box_lateral = (1/32, 1/32), dz_box = 1/32) on the 16 planes of "
                 "layer 2, observed on the 16 planes of each of layers 1, 2 and 3; 64x64 pixels (L=1, dx=1/32, "
                 "padded 128x128); the benchmark's own k_rho x phi grid (box_krho_grid_grouped: N_phi 256, rtol "
                 "1e-3, k_rho cap 3000, rho_far = L, maxevals 200000, 4 mirror classes per group), built by "
                 "build_layer_operator_grouped (node budget 900000, phi-rotation fill on this symmetric stack); "
                 "NUFFT lateral sum (eps 1e-12), z skeletons at tol 1e-6; one apply = apply_operator of a full "
                 "64x64x16x3 complex N(0,1) moment volume; complex FP64"),
    L=1.0, dx=1.0 / 32.0, delta=1e-3, thick=0.5, source_layer=2, obs_layers=(1, 2, 3),
    n_phi=256, rtol=1e-3, kcap=3000.0, maxevals=200_000, classes_per_group=4, node_budget=900_000,
    eps=1e-12, skeleton_tol=1e-6, seed=1,
    # the reference's own apply time for this case: 90 ms on an Apple M5, 4 threads
    # (STATUS_closed_form_rewrite.md:75), i.e. 11.1 applies/s
    reference_applies_per_s=1.0 / 0.090,
)


def operator_volume_stack(wl=OPERATOR_VOLUME_64):
    """vac / glass / vac with Im eps = delta on every layer (m_vac, m_gl, Bench_GreenOperator.jl:42-55)."""
    from .materials import IsotropicMaterial, FlatOpticalLayer
    d = wl["delta"]
    return [FlatOpticalLayer(IsotropicMaterial(complex(1.0, d)), wl["thick"]),
            FlatOpticalLayer(IsotropicMaterial(complex(2.25, d)), wl["thick"]),
            FlatOpticalLayer(IsotropicMaterial(complex(1.0, d)), wl["thick"])]


def operator_volume_planes(wl=OPERATOR_VOLUME_64):
    """planes(d, dx) = ((1:round(d/dx)) - 0.5) dx (Bench_GreenOperator.jl:91): source planes and
    {obs layer: planes}."""
    n = int(round(wl["thick"] / wl["dx"]))
    z = (np.arange(1, n + 1) - 0.5) * wl["dx"]
    return z, {o: z.copy() for o in wl["obs_layers"]}


def operator_volume_groups(wl=OPERATOR_VOLUME_64, device=0):
    """The benchmark's k_rho x phi grid: box_krho_grid_grouped(st, omega, 2; box=(dx, dx, dx),
    N_phi, rtol, k_rho_cap, rho_far = L, maxevals, classes_per_group) (Bench_GreenOperator.jl:
    163-168).  Returns (groups, info)."""
    from .adaptive import box_krho_grid_grouped
    stack = operator_volume_stack(wl)
    d = wl["dx"]
    return box_krho_grid_grouped(stack, OMEGA, wl["source_layer"], (d, d, d), n_phi=wl["n_phi"],
                                 classes_per_group=wl["classes_per_group"], rtol=wl["rtol"], krho_cap=wl["kcap"],
                                 rho_far=wl["L"], maxevals=wl["maxevals"], device=device)


POLE_SCAN_HYPERBOLIC = dict(
    name="pole_scan_hyperbolic",
    description=("real-axis pole search (real_axis_poles, mod_KSpaceFeatures.jl:151-214) of the field of a source "
                 "in the hyperbolic biaxial layer of the reference's operator benchmark (vac / biax "
                 "eps=(2.25, 4.6, -3.2) + 1e-3i / vac, every layer 0.5 thick, source layer 2): the cavity response "
                 "on a log-spaced k_rho grid (relative step 1e-3) from 0.5 k0 to the benchmark's k_rho cap 3000 "
                 "along 9 directions in [0, pi/2] (a diagonal stack: kspace_features' N_phi_scan = 32), then "
                 "golden-section location and half-width of every prominent maximum; complex FP64"),
    delta=1e-3, thick=0.5, source_layer=2, n_phi=9, k_lo_k0=0.5, k_hi=3000.0, scan_step=1e-3,
)


def pole_scan_stack(wl=POLE_SCAN_HYPERBOLIC):
    """m_vac / m_hyp / m_vac of Bench_GreenOperator.jl:42-55 at loss delta."""
    from .materials import IsotropicMaterial, BiaxialMaterial, FlatOpticalLayer
    d = wl["delta"]
    vac = IsotropicMaterial(complex(1.0, d))
    return [FlatOpticalLayer(vac, wl["thick"]),
            FlatOpticalLayer(BiaxialMaterial(complex(2.25, d), complex(4.6, d), complex(-3.2, d)), wl["thick"]),
            FlatOpticalLayer(vac, wl["thick"])]


def pole_scan_args(wl=POLE_SCAN_HYPERBOLIC):
    """(phis, k_lo, k_hi, scan_step, exclude) of pole_scan_hyperbolic."""
    phis = list(np.linspace(0.0, math.pi / 2, wl["n_phi"]))
    return phis, wl["k_lo_k0"] * OMEGA, wl["k_hi"], wl["scan_step"], [OMEGA]


PROBE_BATCH_BOX = dict(
    name="probe_batch_box",
    description=("probe integrands of the adaptive box k_rho grid (box_kρ_grid, mod_ClosedFormGrid.jl:155-255) on "
                 "the reference's grid validation setup (vac 1.0 / glass 2.25+0.05i 0.6 / vac 1.0, omega = 2 pi, box "
                 "0.1^3 centred at 0.3 in layer 2, rho_far = hypot(0.5, 0.5): 8 probes in 3 observation layers, "
                 "96-node phi grid): one batch of 4096 log-spaced k_rho in [1e-3, 6000] -> the 3x3 box Green's "
                 "tensor of every probe at every k_rho (one-k_rho-node caches: node solves of every (k_rho, phi) "
                 "with the box factors, interior slab and local Ez, then the phi sums); complex FP64"),
    n_krho=4096, k_lo=1e-3, k_hi=6000.0, n_phi=96, box=(0.1, 0.1, 0.1), zbar=0.3, source_layer=2,
)
# Algorithmic FP64 operations per (k_rho, phi) node of probe_batch_box: the
# oracle's reference-order op count (real adds + multiplies + divides) of the
# node solve for each observation layer plus the spectral dyadic of each of
# that layer's probes (point-source count: the box factors come on top), mean
# over 64 seeded nodes.  Pinned by tests/test_workloads.py.
PROBE_BATCH_BOX_FLOPS_PER_NODE = 52578.0


def probe_batch_stack(wl=PROBE_BATCH_BOX):
    """vac / glass / vac of Gz_ClosedForm_Grid_Validation.jl:25-28."""
    from .materials import IsotropicMaterial, FlatOpticalLayer
    return [FlatOpticalLayer(IsotropicMaterial(1.0 + 0.001j), 1.0), FlatOpticalLayer(IsotropicMaterial(2.25 + 0.05j), 0.6),
            FlatOpticalLayer(IsotropicMaterial(1.0 + 0.001j), 1.0)]


def probe_batch_args(wl=PROBE_BATCH_BOX):
    """(k_rho nodes, phi nodes, phi weights, probes) of probe_batch_box."""
    from .probes import box_probes
    from .quadrature import build_kphi_grid
    kr = np.geomspace(wl["k_lo"], wl["k_hi"], wl["n_krho"])
    phis, wph = build_kphi_grid(wl["n_phi"])
    probes = box_probes(probe_batch_stack(wl), wl["source_layer"], wl["box"], zbar=[wl["zbar"]],
                        rho_far=math.hypot(0.5, 0.5))
    return kr, phis, wph, probes


BOX_KRHO_GRID = dict(
    name="box_krho_grid",
    description=("the adaptive box k_rho grid (box_kρ_grid, mod_ClosedFormGrid.jl:155-255) of a 1/8-wavelength voxel in "
                 "the glass layer of vac 1.0 / glass 2.25+0.05i 0.6 / vac 1.0 (omega = 2 pi): the default box-centre "
                 "depths (22 probes in 3 observation layers), 64-node phi grid, rtol 1e-6, k_rho cap 2000, rho_far "
                 "0.25: light lines, the real-axis pole scan, the seeds, then the lock-step G7-K15 refinement of every "
                 "probe, each step's new k_rho solved as one batch; complex FP64"),
    box=(0.125, 0.125, 0.125), source_layer=2, n_phi=64, rtol=1e-6, krho_cap=2000.0, rho_far=0.25,
)
# Algorithmic FP64 operations per (k_rho, phi) node of box_krho_grid, counted
# as for probe_batch_box over its 22 probes (mean over 64 seeded nodes,
# k_rho log-spaced in [1e-3, 2000]).  Pinned by tests/test_workloads.py.
BOX_KRHO_GRID_FLOPS_PER_NODE = 64763.125


def box_grid_probes(wl=BOX_KRHO_GRID):
    """The probes box_krho_grid builds for box_krho_grid's setup."""
    from .probes import box_probes
    return box_probes(probe_batch_stack(), wl["source_layer"], wl["box"], rho_far=wl["rho_far"])


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
# FP64 flops the node kernel executes per node on batched_gyro_512 (rocprofv3
# counters of one launch: 64 x (SQ_INSTS_VALU_ADD_F64 + SQ_INSTS_VALU_MUL_F64 +
# 2 SQ_INSTS_VALU_FMA_F64) / 262,144 nodes = 64 x (3.555e6 + 8.151e6 + 2 x 11.362e6)
# / 262,144).  The kernel solves each node with less arithmetic than the reference
# formulation above (block-form eigen, 2x2 recursions, closed-form interfaces).
BATCHED_GYRO_512_EXECUTED_FLOPS_PER_NODE = 8406.0


def uniform_c128_source(nx, ny, seed):
    """Seeded complex N(0,1) source (re, im each N(0, 1/2)) laid out as the
    reference's column-major (Nx, Ny, 3) array = contiguous (3, Ny, Nx)."""
    rng = np.random.default_rng(seed)
    return ((rng.standard_normal((3, ny, nx)) + 1j * rng.standard_normal((3, ny, nx))) / math.sqrt(2.0))
