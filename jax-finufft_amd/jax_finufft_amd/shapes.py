"""Broadcasting and flattening of user arrays to the canonical layout.

Same semantics as src/jax_finufft/shapes.py (BroadcastIndex :11-24,
broadcast_and_flatten_inputs :27-126, abstract_eval :129-170), on torch
tensors: point axes that are broadcast against the strengths are grouped
into the stacked-transform axis ``n_transf`` so they share one setpts, and the
remaining leading axes become ``n_tot`` independent problems.

Canonical layout handed to the backend:
  type 1/3 source  (n_tot, n_transf, M)
  type 2 source    (n_tot, n_transf, *N)
  points           (n_tot, M);  type-3 targets (n_tot, N)
"""

from dataclasses import dataclass
from math import prod
from typing import Sequence

import torch


@dataclass
class BroadcastIndex:
    broadcast_from: Sequence[int]
    broadcast_to: Sequence[int]
    expected_output_shape: Sequence[int]

    def unflatten(self, result):
        result = torch.reshape(result, tuple(self.expected_output_shape))
        if len(self.broadcast_to):
            return torch.movedim(result, tuple(self.broadcast_to), tuple(self.broadcast_from))
        return result


def _bshape(*shapes):
    return tuple(torch.broadcast_shapes(*[tuple(s) for s in shapes]))


def broadcast_and_flatten_inputs(nufft_type, output_shape, source, *points):
    if nufft_type == 3:
        num_dim = len(points) // 2
        points3 = points[num_dim:]
        points = points[:num_dim]
    else:
        num_dim = len(points)
    assert num_dim

    points = torch.broadcast_tensors(*points)
    *input_shape, num_points = points[0].shape
    input_shape = tuple(input_shape)

    if nufft_type == 3:
        points3 = torch.broadcast_tensors(*points3)
        *input_shape3, num_points3 = points3[0].shape
        input_shape = _bshape(input_shape, tuple(input_shape3))
        points = [p.broadcast_to(input_shape + (num_points,)) for p in points]
        points3 = [p.broadcast_to(input_shape + (num_points3,)) for p in points3]
    else:
        points3 = []
        num_points3 = 0

    # unpadded points: an implicit transform axis in the strengths
    if (nufft_type == 2 and source.ndim == len(input_shape) + num_dim + 1) or (
        nufft_type in (1, 3) and source.ndim == len(input_shape) + 2
    ):
        input_shape = tuple(input_shape) + (1,)
        points = tuple(p[..., None, :] for p in points)

    target_shape = _bshape(tuple(source.shape[: len(input_shape)]), input_shape)

    broadcast_from = tuple(
        n
        for n, (input_dim, target_dim) in enumerate(zip(input_shape, target_shape))
        if input_dim != target_dim
    )
    broadcast_to = tuple(
        len(target_shape) - len(broadcast_from) + n for n in range(len(broadcast_from))
    )
    assert all(input_shape[n] == 1 for n in broadcast_from)

    source = source.broadcast_to(target_shape + tuple(source.shape[len(target_shape):]))

    if len(broadcast_to):
        source = torch.movedim(source, broadcast_from, broadcast_to)
        points = tuple(torch.movedim(p, broadcast_from, broadcast_to) for p in points)
        points3 = tuple(torch.movedim(p, broadcast_from, broadcast_to) for p in points3)

    num_in = len(target_shape)
    num_axes = num_in - len(broadcast_from)
    size_in = prod(source.shape[:num_axes])
    size_bcast = prod(source.shape[num_axes:num_in])

    if nufft_type == 3:
        assert source.ndim == len(target_shape) + 1
        assert source.shape[-1] == num_points
        expected_output_shape = tuple(source.shape[:num_in]) + (num_points3,)
        source_extra_shape = (num_points,)
    elif nufft_type == 2:
        assert source.ndim == num_in + num_dim
        expected_output_shape = tuple(source.shape[:num_in]) + (num_points,)
        source_extra_shape = tuple(source.shape[num_in:])
    elif nufft_type == 1:
        assert source.ndim == len(target_shape) + 1
        assert source.shape[-1] == num_points
        assert output_shape is not None
        expected_output_shape = tuple(source.shape[:num_in]) + tuple(output_shape)
        source_extra_shape = (num_points,)

    source = torch.reshape(source, (size_in, size_bcast) + tuple(source_extra_shape))
    points = tuple(torch.reshape(p, (size_in, num_points)) for p in points)
    points3 = tuple(torch.reshape(p, (size_in, num_points3)) for p in points3)

    return (
        BroadcastIndex(
            broadcast_from=broadcast_from,
            broadcast_to=broadcast_to,
            expected_output_shape=expected_output_shape,
        ),
        source,
        *points,
        *points3,
    )


def abstract_eval(source, *points, output_shape, nufft_type):
    """Output shape and dtype of a canonical-layout call (shapes.py:129-170).

    Raises TypeError for an unsupported or inconsistent dtype pairing (the
    reference asserts; complex64 goes with float32, complex128 with float64).
    """
    ndim = len(points) // 2 if nufft_type == 3 else len(points)
    if not 1 <= ndim <= 3:
        raise ValueError("Only 1-, 2-, and 3-dimensions are supported")
    single = source.dtype == torch.complex64 and all(p.dtype == torch.float32 for p in points)
    double = source.dtype == torch.complex128 and all(p.dtype == torch.float64 for p in points)
    if not (single or double):
        raise TypeError(
            "source must be complex64 with float32 points or complex128 with float64 points; "
            f"got {source.dtype} and {[p.dtype for p in points]}"
        )
    assert all(p.ndim == 2 for p in points)
    assert all(p.shape == points[0].shape for p in points[1:ndim])
    assert source.shape[0] == points[0].shape[0]
    if nufft_type == 3:
        assert source.ndim == 3
        assert all(p.shape == points[ndim].shape for p in points[ndim + 1:])
        return tuple(source.shape[:2]) + (points[ndim].shape[-1],), source.dtype
    if nufft_type == 2:
        assert source.ndim == 2 + ndim
        return tuple(source.shape[:2]) + (points[0].shape[-1],), source.dtype
    assert source.ndim == 3
    assert source.shape[2] == points[0].shape[1]
    return tuple(source.shape[:2]) + tuple(output_shape), source.dtype
