"""The Green k-node solve of the GPU kernel (yellowthinfilms_amd/csrc/
node_core.hpp) built as host C++ with AddressSanitizer and
UndefinedBehaviorSanitizer (tests/host_node), one lane per node, run over the
foundation stacks, C1 and C2 for every kernel case (source above the stack,
same layer, observation below or above the source), point and box sources,
and compared with the oracle.

The kernel's LDS block is a heap array that ends at the last slot the case
declares, so a slot index past it is a reported overflow, and every slot is
NaN before each node, so a slot read before the node wrote it shows up as a
NaN g~.  Tolerance: 1e-10 relative, the plain path's bar."""
import math
import os
import struct
import subprocess

import numpy as np
import pytest

from helpers import po, relerr, ostack, c1_stack, c2_stack, FOUNDATION_STACKS, FOUNDATION_KTS, NM

W0 = 2 * math.pi
HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "host_node")
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "yellowthinfilms_amd", "csrc")
CLANG = "/opt/rocm/llvm/bin/clang++"
TERMS = {"all": 0, "direct": 1, "scattered": 2, "toeplitz": 3, "hankel": 4}


@pytest.fixture(scope="module")
def host_node():
    if not os.path.exists(CLANG):
        pytest.skip("no host clang++ with sanitizer runtimes")
    exe = os.path.join(HERE, "_build", "host_node")
    deps = [os.path.join(HERE, "host_node.cpp"), os.path.join(HERE, "include", "hip", "hip_runtime.h")] + \
        [os.path.join(CSRC, f) for f in ("node_core.hpp", "ytf_device.hpp", "ytf_kernels.h")] + \
        [os.path.join(ROOT, "include", "ytf.h")]
    if not os.path.exists(exe) or os.path.getmtime(exe) < max(os.path.getmtime(d) for d in deps):
        os.makedirs(os.path.dirname(exe), exist_ok=True)
        cmd = [CLANG, "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
               "-fno-omit-frame-pointer", "-I", os.path.join(HERE, "include"), "-I", os.path.join(ROOT, "include"),
               "-I", CSRC, "-o", exe, os.path.join(HERE, "host_node.cpp")]
        p = subprocess.run(cmd, capture_output=True, text=True)
        assert p.returncode == 0, p.stderr[-4000:]
    return exe


def _run(exe, tmp_path, stack, kx, ky, src, obs, zs, zo, terms="all", dz=0.0, ez_local=True):
    from yellowthinfilms_amd.green import layer_structs
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    L = layer_structs(stack)
    n, npairs = len(kx), len(zs)
    with open(inp, "wb") as f:
        f.write(struct.pack("=i", len(stack)))
        f.write(bytes(L))
        f.write(struct.pack("=dq", W0, n))
        f.write(np.asarray(kx, np.float64).tobytes())
        f.write(np.asarray(ky, np.float64).tobytes())
        f.write(struct.pack("=iii", src, obs, npairs))
        f.write(np.asarray(zs, np.float64).tobytes())
        f.write(np.asarray(zo, np.float64).tobytes())
        f.write(struct.pack("=idii", TERMS[terms], dz, 1 if ez_local else 0, 1))
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:halt_on_error=1:abort_on_error=0",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    env.pop("LD_PRELOAD", None)
    p = subprocess.run([exe, str(inp), str(out)], capture_output=True, text=True, env=env, timeout=120)
    assert p.returncode == 0 and "runtime error" not in p.stderr and "AddressSanitizer" not in p.stderr, \
        p.stderr[-4000:]
    g = np.fromfile(out, np.complex128).reshape(npairs, 3, 3, n).transpose(3, 0, 2, 1)  # (n, npairs, c, q)
    og = po.ghat_batch(ostack(stack), W0, kx, ky, src, obs, zs, zo, terms=terms, dz_box=dz, ez_local=ez_local)
    _run.redo = int(p.stdout.split()[-1])  # nodes solved again with the interface QR fallback
    return g, og


def _nodes(seed, kmax=12.0, n=12):
    rng = np.random.default_rng(seed)
    kts = list(FOUNDATION_KTS) + [tuple(rng.uniform(-kmax, kmax, 2)) for _ in range(n)]
    return np.array([k[0] for k in kts]), np.array([k[1] for k in kts])


STACKS = dict(FOUNDATION_STACKS, **{"C1 uniaxial film": None, "C2 gyrotropic film": None})


@pytest.mark.parametrize("label", list(STACKS))
def test_node_solve_under_sanitizers_every_case(host_node, tmp_path, label):
    stack = {"C1 uniaxial film": c1_stack, "C2 gyrotropic film": c2_stack}.get(label, lambda: STACKS[label])()
    N = len(stack)
    kx, ky = _nodes(11)
    worst = 0.0
    for src in range(1, N):
        ds = stack[src - 1].d
        for obs in range(1, N + 1):
            do = stack[obs - 1].d
            zs, zo = [0.3 * ds, 0.7 * ds], [0.2 * do, 0.9 * do]
            g, og = _run(host_node, tmp_path, stack, kx, ky, src, obs, zs, zo)
            assert np.all(np.isfinite(g) | ~np.isfinite(og)), (src, obs)
            e = relerr(g, og)
            worst = max(worst, e)
            assert e < 1e-10, (label, src, obs, e)
    print(f"{label}: {N - 1} source layers x {N} observation layers, worst rel err {worst:.1e}")


def test_node_solve_under_sanitizers_box_sources_and_term_classes(host_node, tmp_path):
    stack = FOUNDATION_STACKS["iso 5-layer lossy"]
    kx, ky = _nodes(12)
    d2 = stack[1].d
    for terms in ("all", "direct", "scattered", "toeplitz", "hankel"):
        g, og = _run(host_node, tmp_path, stack, kx, ky, 2, 2, [0.3 * d2, 0.35 * d2], [0.32 * d2, 0.6 * d2],
                     terms=terms, dz=0.1 * d2)
        assert relerr(g, og) < 1e-10, (terms, relerr(g, og))
    for obs in (1, 3, 5):
        do = stack[obs - 1].d
        g, og = _run(host_node, tmp_path, stack, kx, ky, 2, obs, [0.4 * d2], [0.5 * do], dz=0.2 * d2)
        assert relerr(g, og) < 1e-10, (obs, relerr(g, og))


def test_node_solve_under_sanitizers_singular_interfaces(host_node, tmp_path):
    """Exactly singular interface solves (two lossless air layers at k_z = 0):
    the pivoted-QR fallback in every kernel case, the singular node first in a
    batch of ordinary ones."""
    from test_oracle_interface_fallback import singular_stack
    stack = singular_stack()
    rng = np.random.default_rng(7)
    kr = rng.uniform(0.05, 3.0, 31) * W0
    ph = rng.uniform(0, 2 * np.pi, 31)
    kx = np.concatenate([[W0], kr * np.cos(ph)])
    ky = np.concatenate([[0.0], kr * np.sin(ph)])
    for src, obs, zs, zo in ((1, 1, 0.3, 0.2), (1, 6, 0.3, 0.1), (5, 1, 0.01, 0.3), (5, 5, 0.01, 0.02), (4, 4, 0.05, 0.08)):
        g, og = _run(host_node, tmp_path, stack, kx, ky, src, obs, [zs], [zo])
        assert _run.redo == 1, (src, obs, _run.redo)
        assert np.isfinite(g[0]).all(), (src, obs)
        assert relerr(g, og) < 1e-10, (src, obs, relerr(g, og))
        assert np.abs(g[0] - og[0]).max() <= 1e-10 * max(np.abs(og[0]).max(), 1e-300), (src, obs)
