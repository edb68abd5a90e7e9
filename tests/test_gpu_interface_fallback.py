"""GPU parity at an exactly singular interface solve: the pivoted-QR fallback of
_S_Interface_Matrix (mod_ScatteringMatrix.jl:35-44), in every kernel that
solves interfaces (the node kernel's four cases, the uneven batch and the
cavity response), against the oracle through the C-ABI.

The stack holds two lossless layers of index 1 (air, and eps 2 with mu 0.5);
at k_x = omega, k_y = 0 their k_z is exactly 0, the LUs of the three
interfaces they touch meet zero pivots, and the reference takes the
minimum-norm QR solve (the oracle's fallback is
pinned against numpy's pinv in test_oracle_interface_fallback.py).  The node
must come out valid and within 1e-10 of the oracle, as every other node.
Tolerance: 1e-10 relative (the plain-path bar)."""
import math

import numpy as np
import pytest

from helpers import po, relerr, ostack
from test_gpu_nodes import _compare
from test_oracle_interface_fallback import singular_stack

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

W0 = 2 * math.pi


def _nodes():
    """The singular node first, then 255 ordinary nodes (the wave shares the
    fallback branch with lanes that do not take it)."""
    rng = np.random.default_rng(7)
    kr = rng.uniform(0.05, 3.0, 255) * W0
    ph = rng.uniform(0, 2 * np.pi, 255)
    kx = np.concatenate([[W0], kr * np.cos(ph)])
    ky = np.concatenate([[0.0], kr * np.sin(ph)])
    return kx, ky


@pytest.mark.parametrize("src,obs,zs,zo", [(1, 1, 0.3, 0.2), (1, 6, 0.3, 0.1), (5, 1, 0.01, 0.3), (5, 5, 0.01, 0.02),
                                           (4, 4, 0.05, 0.08)],
                         ids=["top", "right", "left", "same-gyro", "same-uniaxial"])
def test_node_kernel_at_a_singular_interface(src, obs, zs, zo):
    """g~ and every cache entry (modes, seeds, T^o, basis) of every node, the
    singular one included, against the oracle."""
    print(f"case: singular interface, node kernel src={src} obs={obs}", flush=True)
    kx, ky = _nodes()
    stack = singular_stack()
    c, g = _compare(stack, kx, ky, src, obs, [zs], [zo])
    assert c["valid"][0]
    gg = g.reshape(1, 3, 3, -1).transpose(3, 0, 2, 1)
    assert np.isfinite(gg[0]).all()
    og = po.ghat_from_cache(po.nodes_batch(ostack(stack), W0, kx[:1], ky[:1], src, [obs]), obs, [zs], [zo])
    assert np.abs(gg[0] - og[0]).max() <= 1e-10 * max(np.abs(og[0]).max(), 1e-300)


def test_batch_at_a_singular_interface():
    print("case: singular interface, uneven batch", flush=True)
    from yellowthinfilms_amd.batch import StackBatch
    kx, ky = _nodes()
    b = StackBatch([singular_stack()], [len(kx)], W0)
    out = {k: v.cpu().numpy() for k, v in b.solve(kx, ky, modes=True).items()}
    o = po.stack_S_batch(ostack(singular_stack()), W0, kx, ky)
    valid = (out["flags"] & 1) != 0
    assert valid.all() and o["ok"].all()
    S = out["S"].T.reshape(-1, 4, 4).transpose(0, 2, 1)
    assert relerr(S, o["S"]) < 1e-10, relerr(S, o["S"])
    assert np.abs(S[0] - o["S"][0]).max() <= 1e-10 * np.abs(o["S"][0]).max()


@pytest.mark.parametrize("s", [2, 4])
def test_cavity_response_at_a_singular_interface(s):
    print(f"case: singular interface, cavity response s={s}", flush=True)
    from yellowthinfilms_amd.kspace import cavity_response
    kx, ky = _nodes()
    g = cavity_response(singular_stack(), W0, s, kx, ky).cpu().numpy()
    o = po.cavity_response(ostack(singular_stack()), W0, s, kx, ky)
    assert o[0] > 0 and np.isfinite(o[0])
    ok = o > 0
    assert np.abs(g[ok] - o[ok]).max() <= 1e-9 * np.abs(o[ok]).max()
    assert abs(g[0] - o[0]) <= 1e-10 * o[0]
