"""GPU parity of ragged batches (one plan for many 2D type-1 problems).

A ragged batch has no reference counterpart; its baseline is a loop of
``nufft1`` calls (the reference's multi-transform test pattern,
tests/ops_test.py:380-397, extended to problems of different sizes).  Every
problem of a batch must agree with its own single call to fp rounding, and
with the FP64 direct sum within eps.  Graph replays are checked against the
eager result on every replay, with eager work and synchronisations between
them.
"""

import numpy as np
import pytest
import torch

import oracle
from gpu_helpers import CDT, NPR, clustered_points, cnormal, dev, host, rel_l2, uniform_points
from jax_finufft_amd import RaggedPlan, nufft1, nufft1_ragged

pytestmark = pytest.mark.gpu

TOL_SAME = {"f32": 2e-6, "f64": 1e-13}


def _problems(seed, prec, specs):
    """specs: (shape, M, ntransf, distribution) per problem."""
    rng = np.random.default_rng(seed)
    out = []
    for shape, M, T, dist in specs:
        gen = uniform_points if dist == "uniform" else clustered_points
        x, y = gen(rng, 2, M, prec)
        c = cnormal(rng, (T, M) if T > 1 else (M,), prec)
        out.append((shape, x, y, c))
    return out


SPECS = [
    ((32, 32), 1000, 1, "uniform"),
    ((64, 48), 20000, 3, "uniform"),
    ((100, 128), 5000, 1, "clustered"),
    ((64, 64), 0, 2, "uniform"),          # a problem without points
    ((128, 96), 40000, 2, "clustered"),
    ((33, 71), 3000, 1, "uniform"),       # odd mode counts
    ((64, 48), 7000, 4, "uniform"),       # same grid as problem 1
]


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("iflag", [1, -1])
def test_ragged_matches_single_calls_and_direct_sum(prec, graph, iflag):
    eps = 1e-6 if prec == "f32" else 1e-10
    probs = _problems(31, prec, SPECS)
    shapes = [p[0] for p in probs]
    xs = [dev(p[1]) for p in probs]
    ys = [dev(p[2]) for p in probs]
    cs = [dev(p[3]) for p in probs]
    got = nufft1_ragged(shapes, cs, xs, ys, iflag=iflag, eps=eps, graph=graph)
    torch.cuda.synchronize()
    for i, (shape, x, y, c) in enumerate(probs):
        assert tuple(got[i].shape) == tuple(c.shape[:-1]) + shape
        single = host(nufft1(shape, cs[i], xs[i], ys[i], iflag=iflag, eps=eps))
        g = host(got[i])
        if x.size == 0:
            assert np.count_nonzero(g) == 0
            continue
        assert rel_l2(g, single) < TOL_SAME[prec], (i, rel_l2(g, single))
        cc = c.reshape(-1, x.size)
        for t in range(cc.shape[0]):
            ref = oracle.type1([x, y], cc[t], shape, iflag)
            gt = g.reshape((-1,) + shape)[t]
            assert rel_l2(gt, ref) <= eps, (i, t, rel_l2(gt, ref))


def test_graph_replays_match_eager_every_time():
    probs = _problems(32, "f32", SPECS)
    shapes = [p[0] for p in probs]
    ntr = [p[3].size // max(p[1].size, 1) if p[1].size else 2 for p in probs]
    npts = [p[1].size for p in probs]
    gplan = RaggedPlan(shapes, npts, ntr, eps=1e-6, graph=True)
    eplan = RaggedPlan(shapes, npts, ntr, eps=1e-6, graph=False)
    rng = np.random.default_rng(33)
    for rep in range(5):
        # new strengths and points every replay, same shapes
        xs, ys, cs = [], [], []
        for (shape, x, y, c), m, t in zip(probs, npts, ntr):
            xs.append(dev(rng.uniform(-np.pi, np.pi, m).astype(np.float32)))
            ys.append(dev(rng.uniform(-np.pi, np.pi, m).astype(np.float32)))
            cs.append(dev(cnormal(rng, (t, m) if t > 1 else (m,), "f32")))
        a = gplan.run(xs, ys, cs)
        torch.cuda.synchronize()
        b = eplan.run(xs, ys, cs)
        torch.cuda.synchronize()
        for i in range(len(probs)):
            assert torch.equal(a[i], b[i]), (rep, i)
        # unrelated eager work between replays
        _ = (torch.randn(1 << 20, device="cuda") * 2).sum().item()
    assert gplan.uses_graph


def test_ragged_heavy_units_and_output_reuse():
    # one problem with every point in one cell splits its unit into pieces that
    # flush with atomics; outputs written into caller-provided tensors
    rng = np.random.default_rng(34)
    M = 30000
    x = np.full(M, 0.25, np.float32)
    y = np.full(M, -1.0, np.float32)
    c = cnormal(rng, M, "f32")
    x2, y2 = uniform_points(rng, 2, 5000, "f32")
    c2 = cnormal(rng, 5000, "f32")
    shapes = [(64, 64), (96, 80)]
    plan = RaggedPlan(shapes, [M, 5000], [1, 1], eps=1e-6, graph=True)
    out = [torch.empty(s, dtype=torch.complex64, device="cuda") for s in shapes]
    for _ in range(2):
        plan.run([dev(x), dev(x2)], [dev(y), dev(y2)], [dev(c), dev(c2)], out=out)
    torch.cuda.synchronize()
    assert rel_l2(host(out[0]), oracle.type1([x, y], c, shapes[0], 1)) <= 1e-6
    assert rel_l2(host(out[1]), oracle.type1([x2, y2], c2, shapes[1], 1)) <= 1e-6


def test_ragged_rejects_mismatched_inputs():
    plan = RaggedPlan([(64, 64)], [100], [1], eps=1e-6, graph=False)
    x = torch.zeros(100, device="cuda")
    c = torch.zeros(100, dtype=torch.complex64, device="cuda")
    with pytest.raises(ValueError):
        plan.run([x[:50]], [x[:50]], [c])
    with pytest.raises(TypeError):
        plan.run([x.double()], [x.double()], [c])
