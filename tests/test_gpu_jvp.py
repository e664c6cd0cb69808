"""GPU tests of forward-mode differentiation (the reference's jvp rule,
src/jax_finufft/ops.py:158-277): torch.autograd.gradcheck with forward-mode
AD checked against finite differences in fp64 (the counterpart of
jtu.check_grads with modes=["fwd"] in tests/ops_test.py:128-219), and
torch.func.jvp against FP64 central differences of the transform itself.
"""

from itertools import product

import numpy as np
import pytest
import torch

from gpu_helpers import cnormal, dev, rel_l2
from jax_finufft_amd import nufft1, nufft2, nufft3
from jax_finufft_amd.options import Opts

pytestmark = pytest.mark.gpu


def _modes(ndim, n=12):
    return tuple(int(v) for v in (n // ndim + 5 * np.arange(ndim)))


def _fwd_check(fn, inputs):
    return torch.autograd.gradcheck(fn, inputs, eps=1e-6, atol=1e-5, rtol=1e-4,
                                    check_forward_ad=True, check_backward_ad=False,
                                    check_undefined_grad=False, check_batched_grad=False)


@pytest.mark.parametrize("ndim,iflag,modeord", product([1, 2, 3], [-1, 1], [0, 1]))
def test_nufft1_forward_mode(ndim, iflag, modeord):
    rng = np.random.default_rng(657)
    M = 12
    n_modes = _modes(ndim)
    x = [dev(rng.uniform(-np.pi, np.pi, M)).requires_grad_() for _ in range(ndim)]
    c = dev(cnormal(rng, M, "f64")).requires_grad_()
    opts = Opts(modeord=bool(modeord))
    fn = lambda c, *x: nufft1(n_modes, c, *x, eps=1e-12, iflag=iflag, opts=opts)  # noqa: E731
    assert _fwd_check(fn, (c, *x))


@pytest.mark.parametrize("ndim,iflag,modeord", product([1, 2, 3], [-1, 1], [0, 1]))
def test_nufft2_forward_mode(ndim, iflag, modeord):
    rng = np.random.default_rng(657)
    M = 12
    n_modes = _modes(ndim)
    x = [dev(rng.uniform(-np.pi, np.pi, M)).requires_grad_() for _ in range(ndim)]
    f = dev(cnormal(rng, n_modes, "f64")).requires_grad_()
    opts = Opts(modeord=bool(modeord))
    fn = lambda f, *x: nufft2(f, *x, eps=1e-12, iflag=iflag, opts=opts)  # noqa: E731
    assert _fwd_check(fn, (f, *x))


@pytest.mark.parametrize("ndim,iflag", product([1, 2, 3], [-1, 1]))
def test_nufft3_forward_mode(ndim, iflag):
    rng = np.random.default_rng(657)
    M, N = 10, 9
    x = [dev(rng.uniform(-2.0, 2.0, M)).requires_grad_() for _ in range(ndim)]
    s = [dev(rng.uniform(-3.0, 3.0, N)).requires_grad_() for _ in range(ndim)]
    c = dev(cnormal(rng, M, "f64")).requires_grad_()
    fn = lambda c, *p: nufft3(c, *p, eps=1e-12, iflag=iflag)  # noqa: E731
    assert _fwd_check(fn, (c, *x, *s))


@pytest.mark.parametrize("nufft_type", [1, 2])
def test_func_jvp_matches_central_differences(nufft_type):
    # one direction through source and both coordinate arrays, batched layout
    rng = np.random.default_rng(5)
    M, T = 300, 2
    n_modes = (24, 30)
    x = [dev(rng.uniform(-np.pi, np.pi, M)) for _ in range(2)]
    dx = [dev(rng.normal(size=M)) for _ in range(2)]
    if nufft_type == 1:
        src = dev(cnormal(rng, (T, M), "f64"))
        dsrc = dev(cnormal(rng, (T, M), "f64"))
        fn = lambda s, a, b: nufft1(n_modes, s, a, b, eps=1e-13)  # noqa: E731
    else:
        src = dev(cnormal(rng, (T,) + n_modes, "f64"))
        dsrc = dev(cnormal(rng, (T,) + n_modes, "f64"))
        fn = lambda s, a, b: nufft2(s, a, b, eps=1e-13)  # noqa: E731
    _, tangent = torch.func.jvp(fn, (src, *x), (dsrc, *dx))
    h = 1e-6
    plus = fn(src + h * dsrc, *[a + h * d for a, d in zip(x, dx)])
    minus = fn(src - h * dsrc, *[a - h * d for a, d in zip(x, dx)])
    fd = (plus - minus) / (2 * h)
    assert rel_l2(tangent.cpu().numpy(), fd.cpu().numpy()) < 1e-6
