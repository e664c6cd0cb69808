"""FINUFFT-method CPU restatement: spread/interp/deconv in C, FFT by scipy.fft.

TEST INFRASTRUCTURE ONLY (stage-level checker and bench.py's cpu_baseline).

Follows the FINUFFT method that the reference delegates to
(lib/jax_finufft_cpu.cc:27-76 drives plan -> setpts -> execute per n_tot
index; the arithmetic is in the absent vendor/finufft submodule).  The
published method is restated in ``finufft_cpu.c``; this module strings the
stages together in the same order FINUFFT's execute does:

* type 1: spread -> FFT -> deconvolve;
* type 2: amplify -> FFT -> interpolate.

The FFT sign: B_k = sum_l b_l exp(i iflag k l h).  With numpy conventions that
is ``nf * ifft(b)`` for iflag = +1 and ``fft(b)`` for iflag = -1.
"""

import ctypes
import os

import numpy as np
import scipy.fft

from . import _c128, _f64, _native_modes, _native_points, _ptr, lib


def params(eps, upsampfac=2.0):
    """(ns, beta, warn) from the published width/shape rules."""
    ns = ctypes.c_int(0)
    beta = ctypes.c_double(0.0)
    ret = lib().fcpu_setup(float(eps), float(upsampfac), ctypes.byref(ns), ctypes.byref(beta))
    return ns.value, beta.value, ret


def fine_size(n, ns, upsampfac=2.0):
    return int(lib().fcpu_fine_size(int(n), float(upsampfac), int(ns)))


def next235even(n):
    return int(lib().fcpu_next235even(int(n)))


def gauss_legendre(n):
    x = np.zeros(n)
    w = np.zeros(n)
    lib().fcpu_gauss_legendre(int(n), _ptr(x), _ptr(w))
    return x, w


def phihat(ns, beta, nf, n):
    """psi_hat(k h) for k = -(n//2) .. (n-1)//2."""
    out = np.zeros(n)
    lib().fcpu_phihat(int(ns), float(beta), int(nf), -(n // 2), int(n), _ptr(out))
    return out


class Plan:
    """One FINUFFT-method problem (type 1 or 2) on the CPU, user axis order."""

    def __init__(self, n_modes, eps, iflag, modeord=0, upsampfac=2.0, workers=None, ns=None):
        self.n_modes = tuple(int(n) for n in n_modes)
        self.dim = len(self.n_modes)
        self.iflag = int(iflag)
        self.modeord = int(modeord)
        self.ns, self.beta, self.warn = params(eps, upsampfac)
        if ns is not None and upsampfac == 2.0:
            # a forced width, with the published beta/width ratio for it
            self.ns = int(ns)
            self.beta = {2: 2.20, 3: 2.26, 4: 2.38}.get(self.ns, 2.30) * self.ns
        self.nf_native = [fine_size(n, self.ns, upsampfac) for n in self.n_modes[::-1]]
        while len(self.nf_native) < 3:
            self.nf_native.append(1)
        nat = _native_modes(self.n_modes)
        self.ph = [
            phihat(self.ns, self.beta, self.nf_native[d], nat[d]) if d < self.dim else np.ones(1)
            for d in range(3)
        ]
        self.workers = workers or os.cpu_count()

    @property
    def fine_shape(self):
        """C-ordered fine-grid shape (slowest first)."""
        return tuple(self.nf_native[: self.dim][::-1])

    def _fft(self, b):
        axes = tuple(range(-self.dim, 0))
        if self.iflag > 0:
            return scipy.fft.ifftn(b, axes=axes, norm="forward", workers=self.workers)
        return scipy.fft.fftn(b, axes=axes, workers=self.workers)

    def spread(self, points, c):
        dim, (x, y, z) = _native_points(points)
        c = _c128(c).ravel()
        fw = np.zeros(self.fine_shape, dtype=np.complex128)
        n1, n2, n3 = self.nf_native
        lib().fcpu_spread(
            dim, c.size, _ptr(x), _ptr(y), _ptr(z), _ptr(c.view(np.float64)),
            self.ns, self.beta, n1, n2, n3, _ptr(fw.view(np.float64)),
        )
        return fw

    def interp(self, points, fw):
        dim, (x, y, z) = _native_points(points)
        fw = _c128(fw)
        M = x.size
        c = np.zeros(M, dtype=np.complex128)
        n1, n2, n3 = self.nf_native
        lib().fcpu_interp(
            dim, M, _ptr(x), _ptr(y), _ptr(z), _ptr(fw.view(np.float64)),
            self.ns, self.beta, n1, n2, n3, _ptr(c.view(np.float64)),
        )
        return c

    def deconv(self, B):
        B = _c128(B)
        out = np.zeros(self.n_modes, dtype=np.complex128)
        n1, n2, n3 = _native_modes(self.n_modes)
        f1, f2, f3 = self.nf_native
        lib().fcpu_deconv(
            self.dim, n1, n2, n3, f1, f2, f3, _ptr(self.ph[0]), _ptr(self.ph[1]),
            _ptr(self.ph[2]), self.modeord, _ptr(B.view(np.float64)),
            _ptr(out.view(np.float64)),
        )
        return out

    def amplify(self, f):
        f = _c128(f)
        bhat = np.zeros(self.fine_shape, dtype=np.complex128)
        n1, n2, n3 = _native_modes(self.n_modes)
        f1, f2, f3 = self.nf_native
        lib().fcpu_amplify(
            self.dim, n1, n2, n3, f1, f2, f3, _ptr(self.ph[0]), _ptr(self.ph[1]),
            _ptr(self.ph[2]), self.modeord, _ptr(f.view(np.float64)),
            _ptr(bhat.view(np.float64)),
        )
        return bhat

    def type1(self, points, c):
        return self.deconv(self._fft(self.spread(points, c)))

    def type2(self, points, f):
        return self.interp(points, self._fft(self.amplify(f)))


def nufft1(points, c, n_modes, iflag, eps, modeord=0, upsampfac=2.0):
    return Plan(n_modes, eps, iflag, modeord, upsampfac).type1(points, c)


def nufft2(f, points, iflag, eps, modeord=0, upsampfac=2.0):
    return Plan(np.shape(f), eps, iflag, modeord, upsampfac).type2(points, f)


__all__ = ["Plan", "nufft1", "nufft2", "params", "fine_size", "phihat", "gauss_legendre",
           "next235even", "_f64"]
