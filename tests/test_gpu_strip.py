"""GPU parity of the 2D strip spreaders (the default 2D type-1 spreading).

The strip method is chosen automatically for 2D spreading; here it is forced
(gpu_method=3) through the plan interface and checked against the FP64 direct
sum at every kernel width class: ns <= 8 runs the modular-ownership kernel,
wider kernels the column-per-lane kernel.  A small gpu_maxsubprobsize splits
crowded units into many pieces, each with its own scratch tile that the merge
sums, so the multi-piece path is covered; the output is bitwise reproducible.
"""

from itertools import product

import numpy as np
import pytest
import torch

import oracle
from gpu_helpers import CDT, RDT, clustered_points, cnormal, dev, host, rel_l2, uniform_points
from jax_finufft_amd import nufft1
from jax_finufft_amd._native import Plan
from jax_finufft_amd.options import Opts

pytestmark = pytest.mark.gpu

EPS = {"f32": [1e-3, 1e-5, 1e-6], "f64": [1e-6, 1e-9, 1e-12]}


def _type1_plan(n_modes, x, c, eps, prec, iflag=1, ntransf=1, **opts):
    plan = Plan(1, n_modes, iflag=iflag, ntransf=ntransf, eps=eps, dtype=CDT[prec], **opts)
    plan.setpts(*[dev(a) for a in x])
    out = torch.empty((ntransf,) + tuple(n_modes), dtype=CDT[prec], device="cuda")
    plan.execute(dev(c), out)
    torch.cuda.synchronize()
    info = plan.info()
    plan.destroy()
    return host(out), info


@pytest.mark.parametrize("prec,dist,pmax", product(["f32", "f64"], ["uniform", "clustered"],
                                                   [0, 64]))
@pytest.mark.parametrize("n_modes", [(40, 52), (150, 97)])
def test_strip_spreader_vs_direct_sum(prec, dist, pmax, n_modes):
    rng = np.random.default_rng(21)
    M = 4000
    gen = uniform_points if dist == "uniform" else clustered_points
    x = gen(rng, 2, M, prec)
    c = cnormal(rng, M, prec)
    ref = oracle.type1(x, c, n_modes, 1)
    for eps in EPS[prec]:
        got, info = _type1_plan(n_modes, x, c, eps, prec, gpu_method=3,
                                gpu_maxsubprobsize=pmax)
        assert rel_l2(got[0], ref) <= eps, (eps, info, rel_l2(got[0], ref))


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_strip_matches_lds_method(prec):
    # same plan parameters, two spreaders: the results agree to rounding
    rng = np.random.default_rng(22)
    M = 20000
    x = clustered_points(rng, 2, M, prec)
    c = cnormal(rng, (3, M), prec)
    eps = 1e-6
    a, _ = _type1_plan((128, 96), x, c, eps, prec, ntransf=3, gpu_method=3)
    b, _ = _type1_plan((128, 96), x, c, eps, prec, ntransf=3, gpu_method=2)
    assert rel_l2(a, b) < (2e-6 if prec == "f32" else 1e-13)


@pytest.mark.parametrize("pmax", [0, 100])
def test_strip_dense_rows_and_one_crowded_cell(pmax):
    # every point in a handful of rows (long runs of equal (row, block) keys),
    # plus one cell holding a third of all points
    rng = np.random.default_rng(23)
    M = 30000
    y = rng.choice(np.linspace(-3.0, 3.0, 5), M)
    x = rng.uniform(-np.pi, np.pi, M)
    x[: M // 3] = 0.7
    y[: M // 3] = -2.2
    c = cnormal(rng, M, "f32")
    pts = [x.astype(np.float32), y.astype(np.float32)]
    got, _ = _type1_plan((256, 200), pts, c, 1e-6, "f32", gpu_method=3, gpu_maxsubprobsize=pmax)
    ref = oracle.type1(pts, c, (256, 200), 1)
    assert rel_l2(got[0], ref) <= 1e-6


def test_auto_method_is_strip_for_2d_type1():
    # the default (auto) spreading of a 2D type-1 plan agrees with the forced strip plan
    rng = np.random.default_rng(24)
    M = 5000
    x = uniform_points(rng, 2, M, "f32")
    c = cnormal(rng, M, "f32")
    a, _ = _type1_plan((100, 120), x, c, 1e-6, "f32", gpu_method=3)
    b = host(nufft1((100, 120), dev(c), *[dev(v) for v in x], eps=1e-6, opts=Opts()))
    assert rel_l2(a[0], b) < 1e-7


@pytest.mark.parametrize("pmax", [0, 50])
def test_strip_is_bitwise_reproducible(pmax):
    # no float atomics: the same inputs give the same bits, crowded units included
    rng = np.random.default_rng(25)
    M = 30000
    x = clustered_points(rng, 2, M, "f32")
    c = cnormal(rng, (2, M), "f32")
    a, _ = _type1_plan((128, 96), x, c, 1e-6, "f32", ntransf=2, gpu_method=3,
                       gpu_maxsubprobsize=pmax)
    b, _ = _type1_plan((128, 96), x, c, 1e-6, "f32", ntransf=2, gpu_method=3,
                       gpu_maxsubprobsize=pmax)
    assert np.array_equal(a.view(np.float32), b.view(np.float32))


@pytest.mark.parametrize("prec,dist,pmax", product(["f32", "f64"], ["uniform", "clustered"],
                                                   [0, 64]))
def test_strip_3d_slices_vs_direct_sum(prec, dist, pmax):
    # 3D spreading as ns 2D slice points per point: every z-slice of the fine
    # grid a stack of strips (the fastest axis, the last mode count, needs a
    # fine grid of >= 64 cells)
    rng = np.random.default_rng(26)
    M = 3000
    n_modes = (24, 20, 40)
    gen = uniform_points if dist == "uniform" else clustered_points
    x = gen(rng, 3, M, prec)
    c = cnormal(rng, M, prec)
    ref = oracle.type1(x, c, n_modes, 1)
    for eps in EPS[prec][:2]:
        got, info = _type1_plan(n_modes, x, c, eps, prec, gpu_method=3, gpu_maxsubprobsize=pmax)
        assert rel_l2(got[0], ref) <= eps, (eps, info, rel_l2(got[0], ref))


def test_strip_3d_matches_lds_method_and_is_reproducible():
    rng = np.random.default_rng(27)
    M = 20000
    x = clustered_points(rng, 3, M, "f32")
    c = cnormal(rng, (2, M), "f32")
    a, _ = _type1_plan((30, 32, 64), x, c, 1e-6, "f32", ntransf=2, gpu_method=3)
    b, _ = _type1_plan((30, 32, 64), x, c, 1e-6, "f32", ntransf=2, gpu_method=2)
    a2, _ = _type1_plan((30, 32, 64), x, c, 1e-6, "f32", ntransf=2, gpu_method=3)
    assert rel_l2(a, b) < 2e-6
    assert np.array_equal(a.view(np.float32), a2.view(np.float32))
