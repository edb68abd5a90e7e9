"""GPU parity of the uneven-batch solve (ytf_batch_*): the total S-matrix and
end-layer modes of many stacks in one persistent launch, against the
reference's golden reflectances and the oracle, through the C-ABI.

Tolerances: FP64; reflectances at the golden files' 10 decimals (<= 1e-10);
S and modes <= 1e-10 relative per stack against the oracle."""
import os

import numpy as np
import pytest

from helpers import po, relerr, ostack
from test_oracle import CASES, GOLDEN, _own_stack

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

W0 = 2 * np.pi


def _m(t, k):
    """(e, n) complex tensor column k -> 4x4 (column-major entries)."""
    return t[:, k].reshape(4, 4).T


def test_reflectance_from_the_batch_matches_reference_tsv():
    print("case: batch reflectance vs golden TSV", flush=True)
    from yellowthinfilms_amd.batch import StackBatch
    rows = {c: [l.split("\t") for l in open(os.path.join(GOLDEN, f"{c}.tsv")).read().strip().split("\n")[1:]]
            for c in ("film", "stack")}
    stacks = [_own_stack(CASES[c]) for c in ("film", "stack")]
    th = {c: np.array([float(r[1]) for r in rows[c]]) for c in rows}
    kx = np.concatenate([W0 * np.sin(np.deg2rad(th[c])) for c in ("film", "stack")])
    b = StackBatch(stacks, [len(th["film"]), len(th["stack"])], W0)
    out = b.solve(kx, np.zeros_like(kx), modes=True)
    S = out["S"].cpu().numpy()
    V1 = out["V1"].cpu().numpy()
    worst, k = 0.0, 0
    for c in ("film", "stack"):
        for r in rows[c]:
            pol, theta = r[0], float(r[1])
            cth = np.cos(np.deg2rad(theta))
            psi = np.array([0, 1, -cth, 0.0]) if pol == "s" else np.array([cth, 0, 0, 1.0])
            V = _m(V1, k)
            a = np.linalg.lstsq(V[:, :2], psi, rcond=None)[0]
            psir = V[:, 2:] @ (_m(S, k)[2:4, 0:2] @ a)
            R = abs(psir[1]) ** 2 + abs(psir[0]) ** 2 / cth ** 2
            worst = max(worst, abs(R - float(r[3])))
            k += 1
    assert worst < 1e-10, worst


def _small_uneven(seed=3, n_stacks=12):
    from yellowthinfilms_amd import workloads as W
    wl = dict(W.UNEVEN_256, n_stacks=n_stacks, nodes=(37, 300), seed=seed)
    return W.uneven_batch(wl)


def test_uneven_batch_against_oracle():
    print("case: uneven batch (12 stacks) vs oracle", flush=True)
    from yellowthinfilms_amd.batch import StackBatch
    stacks, counts, kx, ky = _small_uneven()
    b = StackBatch(stacks, counts, W0)
    out = b.solve(kx, ky, modes=True)
    g = {k: v.cpu().numpy() for k, v in out.items()}
    off = np.concatenate([[0], np.cumsum(counts)])
    for s, st in enumerate(stacks):
        sl = slice(off[s], off[s + 1])
        o = po.stack_S_batch(ostack(st), W0, kx[sl], ky[sl])
        valid = (g["flags"][sl] & 1) != 0
        assert np.array_equal(valid, o["ok"]), s
        S = g["S"][:, sl].T.reshape(-1, 4, 4).transpose(0, 2, 1)
        assert relerr(S[valid], o["S"][valid]) < 1e-10, (s, len(st), relerr(S[valid], o["S"][valid]))
        for key in ("V1", "VN"):
            V = g[key][:, sl].T.reshape(-1, 4, 4).transpose(0, 2, 1)
            assert relerr(V[valid], o[key][valid]) < 1e-10, (s, key)
        for key in ("Q1", "QN"):
            assert relerr(g[key][:, sl].T[valid], o[key][valid]) < 1e-10, (s, key)


def test_every_launch_shape_and_graph_replay_give_identical_bits(tmp_path):
    print("case: batch launch shapes, graph replays, tune db", flush=True)
    from yellowthinfilms_amd.batch import StackBatch
    stacks, counts, kx, ky = _small_uneven(seed=5, n_stacks=9)
    b = StackBatch(stacks, counts, W0)
    ref = b.solve(kx, ky)
    S0 = ref["S"].clone()
    F0 = ref["flags"].clone()
    for npi, wpb, blocks in [(64, 1, 7), (128, 2, 64), (512, 4, 1), (256, 1, 1024)]:
        b.set_config(npi, wpb, blocks)
        out = b.solve(kx, ky)
        assert torch.equal(out["S"], S0) and torch.equal(out["flags"], F0), (npi, wpb, blocks)
    # replays of the captured graph on the same buffers, with eager work between
    out = b.outputs()
    for r in range(4):
        out["S"].fill_(float("nan"))
        b.solve(kx, ky, out=out)
        torch.cuda.synchronize()
        _ = (out["S"].abs() + 1).sum().item()
        assert torch.equal(out["S"], S0), r
    db = str(tmp_path / "tune.txt")
    rep = b.tune(kx, ky, tune_db=db)
    assert rep.startswith("tuned"), rep
    cfg = b.config
    assert cfg["nodes_per_item"] in (64, 128, 256, 512) and cfg["waves_per_block"] in (1, 2, 4)
    assert torch.equal(b.solve(kx, ky)["S"], S0)
    b2 = StackBatch(stacks, counts, W0)
    rep2 = b2.tune(kx, ky, tune_db=db)
    assert rep2.startswith("db "), rep2
    assert b2.config == cfg


def test_bad_batches_raise():
    print("case: batch argument errors", flush=True)
    from yellowthinfilms_amd.batch import StackBatch
    from yellowthinfilms_amd._lib import YtfError
    stacks, counts, kx, ky = _small_uneven(seed=7, n_stacks=2)
    with pytest.raises(YtfError):
        StackBatch([stacks[0][:1]], [10], W0)  # one layer
    b = StackBatch(stacks, counts, W0)
    with pytest.raises(YtfError):
        b.set_config(100, 1, 10)
    with pytest.raises(ValueError):
        b.solve(kx[:-1], ky[:-1])
