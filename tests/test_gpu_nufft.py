"""GPU parity of the 2D type-1 nonuniform FFT (ytf_nufft2d1) against the
oracle's direct sum, through the C-ABI.  Tolerance: the relative l2 error over
all modes is at most 10 x eps (the spreading kernel's design accuracy), for
eps = 1e-6, 1e-9 and 1e-12."""
import numpy as np
import pytest

from helpers import po, relerr

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


def _pts(M, seed, span=np.pi):
    rng = np.random.default_rng(seed)
    x, y = rng.uniform(-span, span, M), rng.uniform(-span, span, M)
    c = rng.normal(size=M) + 1j * rng.normal(size=M)
    return x, y, c


@pytest.mark.parametrize("eps", [1e-6, 1e-9, 1e-12])
@pytest.mark.parametrize("N1,N2", [(64, 48), (33, 20)])
def test_against_direct_sum(eps, N1, N2):
    print(f"case: nufft N=({N1},{N2}) eps={eps}", flush=True)
    from yellowthinfilms_amd import nufft2d1
    x, y, c = _pts(3000, 1)
    f = nufft2d1(x, y, c, N1, N2, eps=eps)[0].cpu().numpy().T  # (N1, N2)
    ref = po.nudft2d1(x, y, c, N1=N1, N2=N2)
    assert relerr(f, ref) < 10 * eps, relerr(f, ref)


def test_several_transforms_wrapping_and_empty():
    print("case: nufft ntrans=3, points outside [-pi, pi), M=0", flush=True)
    from yellowthinfilms_amd import nufft2d1
    x, y, _ = _pts(2000, 2, span=7.0)  # folded mod 2 pi
    rng = np.random.default_rng(3)
    c = rng.normal(size=(3, 2000)) + 1j * rng.normal(size=(3, 2000))
    f = nufft2d1(x, y, c, 40, 56, eps=1e-10).cpu().numpy()
    for t in range(3):
        ref = po.nudft2d1(x, y, c[t], N1=40, N2=56)
        assert relerr(f[t].T, ref) < 1e-9
    # points exactly on -pi and just below +pi
    xb = np.array([-np.pi, np.pi - 1e-15, 0.0, -np.pi])
    yb = np.array([np.pi - 1e-15, -np.pi, -np.pi, 0.0])
    cb = np.array([1.0, 2.0j, -1.0, 0.5 + 0.5j])
    fb = nufft2d1(xb, yb, cb, 16, 16, eps=1e-11)[0].cpu().numpy().T
    assert relerr(fb, po.nudft2d1(xb, yb, cb, N1=16, N2=16)) < 1e-10
    f0 = nufft2d1(np.zeros(0), np.zeros(0), np.zeros(0, complex), 8, 8)
    assert float(f0.abs().max()) == 0.0


def test_million_points_on_a_1024_grid_sampled_modes():
    """The benchmark shape (10^6 points, 1024^2 modes, eps 1e-9) checked on
    400 sampled modes, and linearity over the whole output."""
    print("case: nufft 1e6 points, 1024^2", flush=True)
    from yellowthinfilms_amd import nufft2d1
    x, y, c = _pts(1_000_000, 4)
    f = nufft2d1(x, y, c, 1024, 1024, eps=1e-9)[0]
    rng = np.random.default_rng(5)
    modes = np.stack([rng.integers(-512, 512, 400), rng.integers(-512, 512, 400)], axis=1)
    ref = po.nudft2d1(x, y, c, modes=modes)
    got = f.cpu().numpy()[modes[:, 1] + 512, modes[:, 0] + 512]
    assert relerr(got, ref) < 1e-8, relerr(got, ref)
    f2 = nufft2d1(x, y, (1 - 2j) * c, 1024, 1024, eps=1e-9)[0]
    assert float(torch.linalg.norm(f2 - (1 - 2j) * f) / torch.linalg.norm(f)) < 1e-12


def test_clustered_points_many_binning_blocks():
    """300,000 points, 90% of them in a disc of radius 0.05 around the origin
    (a few tiles hold most points: many subproblems per tile, and every
    binning block of the sort contributes to the same tiles), checked on 200
    sampled modes against the direct sum."""
    print("case: nufft clustered 3e5 points, 256^2", flush=True)
    from yellowthinfilms_amd import nufft2d1
    rng = np.random.default_rng(6)
    M, Mc = 300_000, 270_000
    r, th = 0.05 * np.sqrt(rng.uniform(0, 1, Mc)), rng.uniform(0, 2 * np.pi, Mc)
    x = np.concatenate([r * np.cos(th), rng.uniform(-np.pi, np.pi, M - Mc)])
    y = np.concatenate([r * np.sin(th), rng.uniform(-np.pi, np.pi, M - Mc)])
    perm = rng.permutation(M)
    x, y = x[perm], y[perm]
    c = rng.normal(size=M) + 1j * rng.normal(size=M)
    f = nufft2d1(x, y, c, 256, 256, eps=1e-9)[0]
    modes = np.stack([rng.integers(-128, 128, 200), rng.integers(-128, 128, 200)], axis=1)
    ref = po.nudft2d1(x, y, c, modes=modes)
    got = f.cpu().numpy()[modes[:, 1] + 128, modes[:, 0] + 128]
    assert relerr(got, ref) < 1e-8, relerr(got, ref)
