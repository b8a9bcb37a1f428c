"""plbamd_reduce_moments_noalloc (the allocation-free variant of
plbamd_reduce_moments) against plbamd_reduce_moments through the C ABI: equal
sums bit for bit on two calls in a row (the second on a smaller box), in one-
and two-level reductions, and its refusals (out-of-date device set-up, more
partial records than the fixed scratch holds)."""
import ctypes

import numpy as np
import pytest

import palabos_amd as plb
from palabos_amd import native
from cases import TAG_MEBB, channel_case, gpu_lattice, random_state, seeded_uneven_blocks

pytestmark = pytest.mark.gpu


def _sums(lat, box, noalloc):
    out = np.full(5, np.nan)
    b = native.Box.of(box)
    fn = native.lib().plbamd_reduce_moments_noalloc if noalloc else native.lib().plbamd_reduce_moments
    rc = fn(lat.handle, ctypes.byref(b), out.ctypes.data)
    return rc, out


def _check_equal(lat, box):
    rc, want = _sums(lat, box, False)
    assert rc == native.OK, native.lib().plbamd_last_error()
    rc, got = _sums(lat, box, True)
    assert rc == native.OK, native.lib().plbamd_last_error()
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), (got, want)
    return got


def test_noalloc_equals_reduce_moments_single_block(gpu):
    """FP64 channel with Zou-He planes and a momentum-exchange sphere (plane
    records included): whole lattice, then a smaller box."""
    case = channel_case(24, 14, 15, 1.7, 0.04, sphere=((9, 7, 7), 3.0), sphere_tag=TAG_MEBB)
    lat = gpu_lattice(case)
    try:
        lat.setPopulations(random_state(case, 3))
        lat.collideAndStream(5)
        full = _check_equal(lat, (0, 23, 0, 13, 0, 14))
        small = _check_equal(lat, (2, 9, 1, 6, 3, 12))
        assert not np.array_equal(full, small)
    finally:
        lat.close()


def test_noalloc_equals_reduce_moments_two_level_multiblock(gpu):
    """FP16 storage, 12 uneven blocks, more than 4096 partial records (two-level
    reduction), at an odd time (swapped layout); then a smaller box and a
    second, smaller lattice."""
    shape = (128, 96, 96)
    blocks = seeded_uneven_blocks(shape, (3, 2, 2), 32, 64, seed=5)
    lat = plb.Block3D(*shape, plb.D3Q19, plb.BGK(1.6), storage="f16", periodic=(True, True, True),
                           blocks=blocks, keep_blocks=True)
    try:
        rho = 1.0 + 0.01 * np.sin(np.arange(np.prod(shape)).reshape(shape) * 0.01)
        u = np.zeros(shape + (3,))
        u[..., 0] = 0.03
        lat.initializeAtEquilibrium(rho, u)
        lat.collideAndStream(3)
        _check_equal(lat, (0, 127, 0, 95, 0, 95))
        _check_equal(lat, (10, 70, 5, 50, 20, 90))
    finally:
        lat.close()
    small = channel_case(16, 10, 11, 1.8, 0.05)
    lat2 = gpu_lattice(small)
    try:
        lat2.setPopulations(random_state(small, 4))
        lat2.collideAndStream(2)
        _check_equal(lat2, (0, 15, 0, 9, 0, 10))
    finally:
        lat2.close()


def test_noalloc_refusals(gpu):
    """Out-of-date device set-up: PLBAMD_ESTATE; a box needing more partial
    records than the fixed scratch holds: PLBAMD_EINVAL; neither writes `out`."""
    case = channel_case(16, 10, 11, 1.8, 0.05)
    lat = gpu_lattice(case)
    try:
        rc, out = _sums(lat, (0, 15, 0, 9, 0, 10), True)
        assert rc == native.ESTATE and np.isnan(out).all()
        assert b"changed since" in native.lib().plbamd_last_error()
        lat.collideAndStream(1)
        rc, _ = _sums(lat, (0, 15, 0, 9, 0, 10), True)
        assert rc == native.OK
    finally:
        lat.close()
    # three 256^3 blocks: 3 x 65536 records > 131072
    shape = (768, 256, 256)
    blocks = [(0, 255, 0, 255, 0, 255), (256, 511, 0, 255, 0, 255), (512, 767, 0, 255, 0, 255)]
    big = plb.Block3D(*shape, plb.D3Q19, plb.BGK(1.6), storage="f16", periodic=(True, True, True),
                           blocks=blocks, keep_blocks=True)
    try:
        big.collideAndStream(2)
        rc, out = _sums(big, (0, 767, 0, 255, 0, 255), True)
        assert rc == native.EINVAL and np.isnan(out).all()
        assert b"partial records" in native.lib().plbamd_last_error()
        _check_equal(big, (0, 400, 0, 255, 0, 255))   # two blocks' worth fits
    finally:
        big.close()
