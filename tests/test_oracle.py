"""Pins for the CPU oracle (no GPU needed).

The oracle is checked against what it must reproduce independently of its own
formula: the reference tests' seeded closed-form fixtures, special cases that
reduce to numpy's FFT, the type-3/type-1 identity at integer targets, and the
bilinear adjointness of type 1 and type 2.  A few checks plant a mistake in
the oracle's inputs (flipped sign, swapped axes) to show the pins catch it.
"""

import os
from itertools import product

import numpy as np
import pytest
import scipy.integrate

import oracle
from oracle import finufft_cpu as fc

GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def fwd():
    return np.load(os.path.join(GOLD, "forward.npz"))


@pytest.fixture(scope="module")
def mo():
    return np.load(os.path.join(GOLD, "modeord.npz"))


# ------------------------------------------------------------------ golden


@pytest.mark.parametrize("ndim,prec,iflag", product([1, 2, 3], ["f32", "f64"], [-1, 1]))
def test_direct_type1_golden(fwd, ndim, prec, iflag):
    key = f"t1_d{ndim}_{prec}_i{iflag:+d}"
    x, c, f = fwd[key + "_x"], fwd[key + "_c"], fwd[key + "_f"]
    got = oracle.type1(list(x), c, f.shape, iflag)
    assert oracle.rel_l2(got, f) < 1e-13


@pytest.mark.parametrize("ndim,prec,iflag", product([1, 2, 3], ["f32", "f64"], [-1, 1]))
def test_direct_type2_golden(fwd, ndim, prec, iflag):
    key = f"t2_d{ndim}_{prec}_i{iflag:+d}"
    x, f, c = fwd[key + "_x"], fwd[key + "_f"], fwd[key + "_c"]
    got = oracle.type2(f, list(x), iflag)
    assert oracle.rel_l2(got, c) < 1e-13


@pytest.mark.parametrize("ndim,prec,iflag", product([1, 2, 3], ["f32", "f64"], [-1, 1]))
def test_direct_type3_golden(fwd, ndim, prec, iflag):
    key = f"t3_d{ndim}_{prec}_i{iflag:+d}"
    x, s, c, f = fwd[key + "_x"], fwd[key + "_s"], fwd[key + "_c"], fwd[key + "_f"]
    got = oracle.type3(c, list(x), list(s), iflag)
    assert oracle.rel_l2(got, f) < 1e-13


@pytest.mark.parametrize(
    "modeord,n,ndim,t", product([0, 1], [16, 17], [1, 2, 3], [1, 2])
)
def test_direct_modeord_golden(mo, modeord, n, ndim, t):
    key = f"mo{modeord}_n{n}_d{ndim}_t{t}"
    x = list(mo[key + "_x"])
    if t == 1:
        got = oracle.type1(x, mo[key + "_c"], mo[key + "_f"].shape, 1, modeord)
        assert oracle.rel_l2(got, mo[key + "_f"]) < 1e-13
    else:
        got = oracle.type2(mo[key + "_f"], x, 1, modeord)
        assert oracle.rel_l2(got, mo[key + "_c"]) < 1e-13


def test_multi_transform_golden():
    d = np.load(os.path.join(GOLD, "multi_transform.npz"))
    for a in range(d["mt_x"].shape[0]):
        x = [d["mt_x"][a]]
        assert oracle.rel_l2(oracle.type1(x, d["mt_c"][a, 0], (12,), 1), d["mt_t1"][a, 0]) < 1e-13
        assert oracle.rel_l2(oracle.type2(d["mt_f"][a, 0], x, -1), d["mt_t2"][a, 0]) < 1e-13
        assert (
            oracle.rel_l2(oracle.type3(d["mt_c"][a, 0], x, [d["mt_xt"][a]], -1), d["mt_t3"][a, 0])
            < 1e-13
        )


def test_planted_sign_mistake_is_caught(fwd):
    key = "t1_d2_f64_i+1"
    x, c, f = fwd[key + "_x"], fwd[key + "_c"], fwd[key + "_f"]
    assert oracle.rel_l2(oracle.type1(list(x), c, f.shape, -1), f) > 0.1


def test_planted_axis_swap_is_caught(fwd):
    key = "t1_d3_f64_i-1"
    x, c, f = fwd[key + "_x"], fwd[key + "_c"], fwd[key + "_f"]
    # same mode shape, but pair the points with the wrong axes
    got = oracle.type1([x[1], x[0], x[2]], c, f.shape, -1)
    assert oracle.rel_l2(got, f) > 0.1


# ----------------------------------------------- closed-form reductions


@pytest.mark.parametrize("n", [8, 15, 64])
def test_type1_on_grid_is_dft(n):
    rng = np.random.default_rng(1)
    c = rng.normal(size=n) + 1j * rng.normal(size=n)
    x = 2 * np.pi * np.arange(n) / n
    # modeord 1 (FFT order), iflag -1  ==  numpy forward DFT
    np.testing.assert_allclose(oracle.type1([x], c, (n,), -1, modeord=1), np.fft.fft(c),
                               rtol=0, atol=1e-12 * np.abs(c).sum())
    # iflag +1  ==  unnormalised inverse DFT
    np.testing.assert_allclose(oracle.type1([x], c, (n,), +1, modeord=1), n * np.fft.ifft(c),
                               rtol=0, atol=1e-12 * np.abs(c).sum())
    # modeord 0 is the fftshift of modeord 1
    np.testing.assert_allclose(oracle.type1([x], c, (n,), -1, modeord=0),
                               np.fft.fftshift(np.fft.fft(c)), atol=1e-12 * np.abs(c).sum())


def test_type1_2d_on_grid_is_dft2():
    rng = np.random.default_rng(2)
    n0, n1 = 6, 10
    c = rng.normal(size=(n0, n1)) + 1j * rng.normal(size=(n0, n1))
    g0, g1 = np.meshgrid(2 * np.pi * np.arange(n0) / n0, 2 * np.pi * np.arange(n1) / n1,
                         indexing="ij")
    got = oracle.type1([g0.ravel(), g1.ravel()], c.ravel(), (n0, n1), -1, modeord=1)
    np.testing.assert_allclose(got, np.fft.fft2(c), atol=1e-11 * np.abs(c).sum())


def test_type2_on_grid_is_dft():
    rng = np.random.default_rng(3)
    n = 20
    f = rng.normal(size=n) + 1j * rng.normal(size=n)
    x = 2 * np.pi * np.arange(n) / n
    np.testing.assert_allclose(oracle.type2(f, [x], +1, modeord=1), n * np.fft.ifft(f),
                               atol=1e-12 * np.abs(f).sum())


@pytest.mark.parametrize("ndim", [1, 2, 3])
def test_type3_integer_targets_equal_type1(ndim):
    rng = np.random.default_rng(4)
    M = 40
    n_modes = (7, 6, 5)[:ndim]
    x = [rng.uniform(-np.pi, np.pi, M) for _ in range(ndim)]
    c = rng.normal(size=M) + 1j * rng.normal(size=M)
    f1 = oracle.type1(x, c, n_modes, 1)
    grids = np.meshgrid(*[np.arange(-(n // 2), (n + 1) // 2) for n in n_modes], indexing="ij")
    s = [g.ravel().astype(float) for g in grids]
    f3 = oracle.type3(c, x, s, 1)
    assert oracle.rel_l2(f3, f1.ravel()) < 1e-13


@pytest.mark.parametrize("ndim,iflag", product([1, 2, 3], [-1, 1]))
def test_type1_type2_bilinear_adjoint(ndim, iflag):
    rng = np.random.default_rng(5)
    M = 30
    n_modes = (9, 8, 7)[:ndim]
    x = [rng.uniform(-np.pi, np.pi, M) for _ in range(ndim)]
    c = rng.normal(size=M) + 1j * rng.normal(size=M)
    f = rng.normal(size=n_modes) + 1j * rng.normal(size=n_modes)
    lhs = np.sum(oracle.type1(x, c, n_modes, iflag) * f)
    rhs = np.sum(c * oracle.type2(f, x, iflag))
    assert abs(lhs - rhs) < 1e-11 * abs(lhs)


def test_type1_at_matches_full():
    rng = np.random.default_rng(6)
    x = [rng.uniform(-np.pi, np.pi, 25) for _ in range(2)]
    c = rng.normal(size=25) + 1j * rng.normal(size=25)
    full = oracle.type1(x, c, (11, 13), -1, modeord=1)
    idx = np.array([0, 5, 17, 142])
    np.testing.assert_allclose(oracle.type1_at(x, c, (11, 13), -1, idx, modeord=1),
                               full.ravel()[idx], rtol=1e-14)


# ------------------------------------------- FINUFFT-method restatement


def test_published_width_rule():
    # sigma = 2: w = ceil(log10(10/eps)), beta = 2.30 w (2.20/2.26/2.38 for w=2/3/4)
    assert fc.params(1e-6)[:2] == (7, pytest.approx(16.1))
    assert fc.params(1e-9)[:2] == (10, pytest.approx(23.0))
    assert fc.params(1e-1)[:2] == (2, pytest.approx(4.4))
    assert fc.params(1e-2)[:2] == (3, pytest.approx(6.78))
    assert fc.params(1e-3)[:2] == (4, pytest.approx(9.52))
    assert fc.params(1e-20)[0] == 16 and fc.params(1e-20)[2] == 1  # capped, warning


def test_next235even_bruteforce():
    def smooth(m):
        for p in (2, 3, 5):
            while m % p == 0:
                m //= p
        return m == 1

    for n in range(1, 400):
        m = fc.next235even(n)
        assert m % 2 == 0 and smooth(m) and m >= n
        assert not any(k % 2 == 0 and smooth(k) for k in range(max(n, 2), m))


@pytest.mark.parametrize("n", [1, 5, 20, 61])
def test_gauss_legendre_matches_numpy(n):
    x, w = fc.gauss_legendre(n)
    xe, we = np.polynomial.legendre.leggauss(n)
    order = np.argsort(x)
    np.testing.assert_allclose(x[order], xe, atol=1e-14)
    np.testing.assert_allclose(w[order], we, atol=1e-14)


@pytest.mark.parametrize("eps", [1e-3, 1e-6, 1e-10])
def test_phihat_matches_adaptive_quadrature(eps):
    ns, beta, _ = fc.params(eps)
    nf, n = 64, 31
    ph = fc.phihat(ns, beta, nf, n)

    def kernel_ft(k):
        h = 2 * np.pi / nf
        # psi_hat(kh) = int psi(u) cos(k h u) du, psi(u) = phi(2u/ns)
        g = lambda u: np.exp(beta * (np.sqrt(max(0.0, 1 - (2 * u / ns) ** 2)) - 1)) * np.cos(
            k * h * u)
        return scipy.integrate.quad(g, -ns / 2, ns / 2, limit=200, epsabs=1e-14, epsrel=1e-13)[0]

    for q, k in enumerate(range(-(n // 2), (n + 1) // 2)):
        assert ph[q] == pytest.approx(kernel_ft(k), rel=1e-11)


def _es(z, beta):
    return np.where(np.abs(z) <= 1, np.exp(beta * (np.sqrt(np.clip(1 - z * z, 0, None)) - 1)), 0.0)


@pytest.mark.parametrize("ndim", [1, 2, 3])
def test_spread_matches_bruteforce(ndim):
    rng = np.random.default_rng(7)
    eps = 1e-5
    n_modes = (8, 6, 5)[:ndim]
    plan = fc.Plan(n_modes, eps, 1)
    M = 13
    x = [rng.uniform(-3 * np.pi, 3 * np.pi, M) for _ in range(ndim)]  # exercise folding
    c = rng.normal(size=M) + 1j * rng.normal(size=M)
    fw = plan.spread(x, c)
    # brute force: every grid cell, periodic images, the ceil(u - ns/2) tap window
    ns, beta = plan.ns, plan.beta
    shape = plan.fine_shape
    expect = np.zeros(shape, dtype=complex)
    for j in range(M):
        w_axes = []
        for d in range(ndim):
            nf = shape[d]
            u = np.mod(x[d][j] * nf / (2 * np.pi), nf)
            l0 = int(np.ceil(u - ns / 2))
            wd = np.zeros(nf)
            for i in range(ns):
                wd[(l0 + i) % nf] += _es(2 * (u - (l0 + i)) / ns, beta)
            w_axes.append(wd)
        tens = w_axes[0]
        for wd in w_axes[1:]:
            tens = np.multiply.outer(tens, wd)
        expect += c[j] * tens
    np.testing.assert_allclose(fw, expect, atol=1e-13)


@pytest.mark.parametrize("ndim", [1, 2, 3])
def test_interp_is_adjoint_of_spread(ndim):
    rng = np.random.default_rng(8)
    n_modes = (10, 8, 6)[:ndim]
    plan = fc.Plan(n_modes, 1e-8, -1)
    M = 50
    x = [rng.uniform(-np.pi, np.pi, M) for _ in range(ndim)]
    c = rng.normal(size=M) + 1j * rng.normal(size=M)
    g = rng.normal(size=plan.fine_shape) + 1j * rng.normal(size=plan.fine_shape)
    lhs = np.sum(plan.spread(x, c) * g)
    rhs = np.sum(c * plan.interp(x, g))
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)


@pytest.mark.parametrize("ndim,iflag,eps", product([1, 2, 3], [-1, 1], [1e-4, 1e-7, 1e-11]))
def test_restatement_accuracy_vs_direct(ndim, iflag, eps):
    rng = np.random.default_rng(9)
    M = 200
    n_modes = (40, 16, 10)[:ndim]
    x = [rng.uniform(-np.pi, np.pi, M) for _ in range(ndim)]
    c = rng.normal(size=M) + 1j * rng.normal(size=M)
    f = rng.normal(size=n_modes) + 1j * rng.normal(size=n_modes)
    # FINUFFT's published width rule reaches about eps (measured <= 3 eps)
    tol = 4 * eps
    assert oracle.rel_l2(fc.nufft1(x, c, n_modes, iflag, eps), oracle.type1(x, c, n_modes, iflag)) < tol
    assert oracle.rel_l2(fc.nufft2(f, x, iflag, eps), oracle.type2(f, x, iflag)) < tol
