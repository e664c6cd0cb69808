"""Option resolution, the cases of the reference's tests/options_test.py:5-52."""

import pytest

from jax_finufft_amd import options


@pytest.mark.parametrize("opts", [None, options.Opts()])
def test_default_options(opts):
    assert options.unpack_opts(opts, 1, True) == opts
    assert options.unpack_opts(opts, 2, True) == opts
    assert options.unpack_opts(opts, 1, False) == opts
    assert options.unpack_opts(opts, 2, False) == opts


def test_nested_by_type():
    opts = options.NestedOpts(
        type1=options.Opts(spread_debug=True),
        type2=options.Opts(debug=True),
    )
    assert options.unpack_opts(opts, 1, True) == options.Opts(spread_debug=True)
    assert options.unpack_opts(opts, 2, True) == options.Opts(debug=True)


def test_nested_by_direction():
    opts = options.NestedOpts(
        forward=options.Opts(spread_debug=True),
        backward=options.Opts(debug=True),
    )
    assert options.unpack_opts(opts, 1, True) == options.Opts(spread_debug=True)
    assert options.unpack_opts(opts, 1, False) == options.Opts(debug=True)


def test_nested_multi():
    opts = options.NestedOpts(
        forward=options.Opts(spread_debug=True),
        backward=options.NestedOpts(
            type1=options.Opts(spread_debug=True, debug=True),
            type2=options.Opts(debug=True),
        ),
    )
    assert options.unpack_opts(opts, 1, True) == options.Opts(spread_debug=True)
    assert options.unpack_opts(opts, 1, False) == options.NestedOpts(
        type1=options.Opts(spread_debug=True, debug=True),
        type2=options.Opts(debug=True),
    )
    assert options.unpack_opts(options.unpack_opts(opts, 1, False), 1, True) == options.Opts(
        spread_debug=True, debug=True
    )
    assert options.unpack_opts(options.unpack_opts(opts, 1, False), 2, True) == options.Opts(
        debug=True
    )


def test_gpu_attribute_set_matches_reference():
    # the seven GPU attributes the reference forwards (options.py:105-119)
    o = options.Opts(modeord=True, gpu_upsampfac=1.25, gpu_method=2, gpu_sort=False,
                     gpu_kerevalmeth=False, gpu_maxbatchsize=3, gpu_debug=1).to_cufinufft_opts()
    assert (o.modeord, o.upsampfac, o.gpu_method, o.gpu_sort, o.gpu_kerevalmeth,
            o.gpu_maxbatchsize, o.debug) == (1, 1.25, 2, 0, 0, 3, 1)


def test_opts_are_frozen_and_hashable():
    o = options.Opts()
    with pytest.raises(Exception):
        o.modeord = True
    assert hash(o) == hash(options.Opts())
