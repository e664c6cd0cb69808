"""Regenerate the golden fixtures in this directory.

The reference stores no golden files; its tests rebuild their inputs from
seeded generators and compare against closed-form direct sums
(tests/ops_test.py:25-125 forward tests, :494-567 modeord, :380-397
multi-transform).  This script re-creates those seeded inputs with the same
generator calls and evaluates the same closed forms in plain NumPy (an
implementation independent of the C oracle in ``oracle/direct.c``), then saves
inputs and FP64 expected outputs as small ``.npz`` files.

Run from the repository root:  python tests/golden/make_golden.py
"""

import os
from itertools import product

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def cmcl_freqs(n):
    # modeord 0 frequencies, written as in tests/ops_test.py:37 / :509-510
    return np.arange(-np.floor(n / 2), np.floor((n - 1) / 2 + 1))


def fft_freqs(n):
    # modeord 1 frequencies, tests/ops_test.py:512-525
    lo = np.arange(0, np.floor(n / 2)) if n % 2 == 0 else np.arange(0, np.floor(n / 2) + 1)
    return np.concatenate((lo, np.arange(-np.floor(n / 2), 0)))


def t1_closed_form(x, c, n_modes, ks, iflag):
    """f[coords] = sum c exp(i iflag k.x), looping over every output mode."""
    x = np.asarray(x, dtype=np.float64)
    c = np.asarray(c, dtype=np.complex128)
    f = np.zeros(n_modes, dtype=np.complex128)
    for coords in product(*map(range, n_modes)):
        kv = np.array([k[n] for (n, k) in zip(coords, ks)])
        f[coords] = np.sum(c * np.exp(iflag * 1j * np.dot(kv, x)))
    return f


def t2_closed_form(f, x, ks, iflag):
    """c_n = sum_k f_k prod_d exp(i iflag k_d x_d[n]), built axis by axis."""
    f = np.asarray(f, dtype=np.complex128)
    ndim = f.ndim
    M = len(x[0])
    c = np.zeros(M, dtype=np.complex128)
    for n in range(M):
        arg = np.copy(f)
        for i, k in enumerate(ks):
            sl = [None] * ndim
            sl[i] = slice(None)
            arg *= np.exp(iflag * 1j * k * float(x[i][n]))[tuple(sl)]
        c[n] = np.sum(arg)
    return c


def t3_closed_form(c, x, s, iflag):
    c = np.asarray(c, dtype=np.complex128)
    xm = np.array([np.asarray(xd, dtype=np.float64) for xd in x])
    out = np.zeros(len(s[0]), dtype=np.complex128)
    for k in range(len(s[0])):
        sv = np.array([float(sd[k]) for sd in s])
        out[k] = np.sum(c * np.exp(iflag * 1j * np.dot(sv, xm)))
    return out


def forward_cases():
    data = {}
    # tests/ops_test.py:25-54 (type 1) and :57-93 (type 2): M=50, modes 75
    for ndim, x64, iflag in product([1, 2, 3], [False, True], [-1, 1]):
        dtype = np.double if x64 else np.single
        cdtype = np.cdouble if x64 else np.csingle
        n_modes = tuple(int(v) for v in (75 // ndim + 5 * np.arange(ndim)))
        ks = [cmcl_freqs(n) for n in n_modes]

        rng = np.random.default_rng(657)
        x = rng.uniform(-np.pi, np.pi, size=(ndim, 50)).astype(dtype)
        c = (rng.normal(size=50) + 1j * rng.normal(size=50)).astype(cdtype)
        key = f"t1_d{ndim}_{'f64' if x64 else 'f32'}_i{iflag:+d}"
        data[key + "_x"] = x
        data[key + "_c"] = c
        data[key + "_f"] = t1_closed_form(x, c, n_modes, ks, iflag)

        rng = np.random.default_rng(657)
        xs = [rng.uniform(-np.pi, np.pi, size=50).astype(dtype) for _ in range(ndim)]
        f = (rng.normal(size=n_modes) + 1j * rng.normal(size=n_modes)).astype(cdtype)
        key = f"t2_d{ndim}_{'f64' if x64 else 'f32'}_i{iflag:+d}"
        data[key + "_x"] = np.array(xs)
        data[key + "_f"] = f
        data[key + "_c"] = t2_closed_form(f, xs, ks, iflag)

    # tests/ops_test.py:96-125 (type 3): 25 sources, 20 targets in [-1, 1]
    for ndim, x64, iflag in product([1, 2, 3], [False, True], [-1, 1]):
        dtype = np.double if x64 else np.single
        cdtype = np.cdouble if x64 else np.csingle
        rng = np.random.default_rng(657)
        x = [rng.uniform(-1.0, 1.0, size=25).astype(dtype) for _ in range(ndim)]
        s = [rng.uniform(-1.0, 1.0, size=20).astype(dtype) for _ in range(ndim)]
        c = (rng.normal(size=25) + 1j * rng.normal(size=25)).astype(cdtype)
        key = f"t3_d{ndim}_{'f64' if x64 else 'f32'}_i{iflag:+d}"
        data[key + "_x"] = np.array(x)
        data[key + "_s"] = np.array(s)
        data[key + "_c"] = c
        data[key + "_f"] = t3_closed_form(c, x, s, iflag)
    return data


def modeord_cases():
    # tests/ops_test.py:494-567: Nf in {16, 17}, modeord {0, 1}, iflag = 1, fp64
    data = {}
    for modeord, Nf, ndim, nufft_type in product([0, 1], [16, 17], [1, 2, 3], [1, 2]):
        rng = np.random.default_rng(657)
        n_modes = tuple(int(v) for v in (Nf // ndim + 5 * np.arange(ndim)))
        ks = [(cmcl_freqs if modeord == 0 else fft_freqs)(n) for n in n_modes]
        key = f"mo{modeord}_n{Nf}_d{ndim}_t{nufft_type}"
        x = rng.uniform(-np.pi, np.pi, size=(ndim, 50))
        if nufft_type == 1:
            c = rng.normal(size=50) + 1j * rng.normal(size=50)
            data[key + "_x"] = x
            data[key + "_c"] = c
            data[key + "_f"] = t1_closed_form(x, c, n_modes, ks, 1)
        else:
            f = rng.normal(size=n_modes) + 1j * rng.normal(size=n_modes)
            data[key + "_x"] = x
            data[key + "_f"] = f
            data[key + "_c"] = t2_closed_form(f, list(x), ks, 1)
    return data


def multi_transform_case():
    # tests/ops_test.py:380-397: seed 314, n_tot=5, n_tr=1, n_j=50, n_k=12, 35 targets
    rng = np.random.default_rng(314)
    n_tot, n_tr, n_j, n_k, n_target = 5, 1, 50, 12, 35
    f = rng.standard_normal(size=(n_tot, n_tr, n_k)) + 1j * rng.standard_normal(
        size=(n_tot, n_tr, n_k)
    )
    c = rng.standard_normal(size=(n_tot, n_tr, n_j)) + 1j * rng.standard_normal(
        size=(n_tot, n_tr, n_j)
    )
    x = rng.uniform(-np.pi, np.pi, (n_tot, n_j))
    xt = rng.uniform(-1.0, 1.0, (n_tot, n_target))
    ks = [cmcl_freqs(n_k)]
    e1 = np.zeros((n_tot, n_tr, n_k), dtype=np.complex128)
    e2 = np.zeros((n_tot, n_tr, n_j), dtype=np.complex128)
    e3 = np.zeros((n_tot, n_tr, n_target), dtype=np.complex128)
    for a in range(n_tot):
        for b in range(n_tr):
            e1[a, b] = t1_closed_form(x[a][None], c[a, b], (n_k,), ks, 1)
            e2[a, b] = t2_closed_form(f[a, b], [x[a]], ks, -1)
            e3[a, b] = t3_closed_form(c[a, b], [x[a]], [xt[a]], -1)
    return {"mt_f": f, "mt_c": c, "mt_x": x, "mt_xt": xt, "mt_t1": e1, "mt_t2": e2,
            "mt_t3": e3}


def main():
    np.savez_compressed(os.path.join(HERE, "forward.npz"), **forward_cases())
    np.savez_compressed(os.path.join(HERE, "modeord.npz"), **modeord_cases())
    np.savez_compressed(os.path.join(HERE, "multi_transform.npz"), **multi_transform_case())


if __name__ == "__main__":
    main()
