"""Shared helpers for the GPU parity tests."""

import numpy as np
import torch

import oracle

CDT = {"f32": torch.complex64, "f64": torch.complex128}
RDT = {"f32": torch.float32, "f64": torch.float64}
NPC = {"f32": np.complex64, "f64": np.complex128}
NPR = {"f32": np.float32, "f64": np.float64}


def dev(a, dtype=None):
    t = torch.as_tensor(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.cuda()


def host(t):
    return t.detach().cpu().numpy()


def rel_l2(a, b):
    return oracle.rel_l2(a, b)


def uniform_points(rng, ndim, M, prec):
    return [rng.uniform(-np.pi, np.pi, M).astype(NPR[prec]) for _ in range(ndim)]


def clustered_points(rng, ndim, M, prec, sigma=0.1):
    """Gaussian cluster about the origin, wrapped into [-pi, pi)."""
    pts = []
    for _ in range(ndim):
        x = rng.normal(0.0, sigma, M)
        x = np.mod(x + np.pi, 2 * np.pi) - np.pi
        pts.append(x.astype(NPR[prec]))
    return pts


def cnormal(rng, shape, prec):
    return (rng.normal(size=shape) + 1j * rng.normal(size=shape)).astype(NPC[prec])
