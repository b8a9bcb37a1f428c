"""Multi-block lattices: uneven blocks swept by one launch, two-cell envelopes
exchanged every second step. The bulk must stay bit-identical to one block
(Palabos's decomposition invariance), so every case compares against the
oracle (one block) and, for the Palabos multi-block fixture, against Palabos."""
import os

import numpy as np
import pytest

import palabos_amd
from cases import (Case, TAG_WALL, channel_case, random_state, seeded_uneven_blocks, sphere_tags,
                   uneven_blocks)
from golden_cases import MULTIBLOCK_BLOCKS, load
from oracle import oracle

pytestmark = pytest.mark.gpu
GOLDEN_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _lattice_with_blocks(case, blocks, keep=True):
    dyn = palabos_amd.BGK(case.omega) if case.dynamics == "bgk" else palabos_amd.RegularizedBGK(case.omega)
    force = case.force if any(case.force) else None
    lat = palabos_amd.Block3D(case.nx, case.ny, case.nz, palabos_amd.D3Q19, dyn, storage=case.storage,
                              periodic=case.periodic, force=force, blocks=blocks, keep_blocks=keep)
    if case.tags is not None:
        lat.setTags(case.tags)
    for p in case.planes:
        if p.kind == "velocity":
            pid = lat.addVelocityBoundary(p.dir, p.ori, p.box, family=p.family)
            lat.setBoundaryVelocity(pid, p.values)
        else:
            pid = lat.addPressureBoundary(p.dir, p.ori, p.box, family=p.family)
            lat.setBoundaryDensity(pid, p.values)
    if case.alpha is not None:
        lat.setPorosity(case.alpha)
    return lat


def _dtype(storage):
    return {"f64": np.float64, "f32": np.float32}.get(storage, np.uint16)


def _check(case, blocks, steps_list, seed=3, keep=True, expect_blocks=None):
    f0 = oracle.to_storage(random_state(case, seed), case.storage)
    lat = _lattice_with_blocks(case, blocks, keep)
    try:
        assert lat.blockCount() == (len(blocks) if expect_blocks is None else expect_blocks)
        lat.setPopulations(f0)
        done = 0
        want = f0
        for s in steps_list:
            lat.collideAndStream(s - done)
            want, wst = oracle.run(case, want, s - done)
            done = s
            got = lat.getPopulations(dtype=_dtype(case.storage))
            if case.tags is not None:
                wall = case.tags == TAG_WALL
                got[wall] = 0
                want = want.copy()
                want[wall] = 0
            assert np.array_equal(got, want), f"step {s}: {np.count_nonzero(got != want)} differ"
            gst = lat.getStatistics()
            assert gst["count"] == wst["count"]
            np.testing.assert_allclose(gst["avg_rhobar"], wst["avg_rhobar"], rtol=1e-12, atol=1e-15)
            np.testing.assert_allclose(gst["force"], wst["force"], rtol=1e-10, atol=1e-14)
            np.testing.assert_allclose(gst["porous"], wst["porous"], rtol=1e-10, atol=1e-14)
    finally:
        lat.close()


@pytest.mark.parametrize("storage", ["f64", "f32", "f16", "bf16"])
def test_periodic_uneven_blocks(gpu, storage):
    case = Case(20, 18, 16, storage=storage, omega=1.7, periodic=(True, True, True))
    blocks = uneven_blocks(case.shape, [[7, 13], [5], [9]])
    _check(case, blocks, [1, 2, 5, 8])


@pytest.mark.parametrize("storage", ["f16", "bf16"])
def test_even_z_blocks_16bit(gpu, storage):
    """Every block (with its envelope) even along z: the two-cells-per-thread
    kernel runs every block, with link walls and a body force across blocks."""
    case = Case(20, 18, 16, storage=storage, dynamics="rbgk", omega=1.6, periodic=(True, True, False),
                force=(1e-5, 2e-6, 0))
    case.tags = sphere_tags(case.shape, (9.5, 8.5, 7.5), 3.4, TAG_WALL)
    blocks = uneven_blocks(case.shape, [[7, 13], [5], [8]])
    _check(case, blocks, [1, 4, 7])


def test_mixed_periodicity_rbgk_force_porosity(gpu):
    rng = np.random.RandomState(2)
    case = Case(18, 14, 12, storage="f64", dynamics="rbgk", omega=1.6, periodic=(True, False, True),
                force=(2e-5, 0, 1e-5), alpha=rng.rand(18, 14, 12) * 0.3)
    blocks = uneven_blocks(case.shape, [[4, 11], [6], [3, 9]])
    _check(case, blocks, [3, 6])


def test_channel_planes_and_sphere_across_blocks(gpu):
    """Zou-He inlet/outlet planes and a momentum-exchange sphere cut by block faces."""
    case = channel_case(24, 12, 13, 1.8, 0.05, sphere=((9, 6, 6), 3.1))
    blocks = uneven_blocks(case.shape, [[8, 10, 17], [5], [7]])
    _check(case, blocks, [4, 9])


def test_link_walls_across_blocks(gpu):
    case = Case(16, 14, 12, storage="f32", omega=1.5, periodic=(True, True, False))
    case.tags = sphere_tags(case.shape, (7.5, 6.5, 5.5), 3.3, TAG_WALL)
    blocks = uneven_blocks(case.shape, [[6, 9], [7], [5]])
    _check(case, blocks, [2, 7])


def test_single_cell_thick_and_full_span_blocks(gpu):
    """Blocks one cell thick, and a block spanning a whole periodic axis (its
    envelope is its own periodic image)."""
    case = Case(9, 10, 11, storage="f64", omega=1.3, periodic=(True, True, True))
    blocks = uneven_blocks(case.shape, [[1, 2, 8], [], [5]])
    _check(case, blocks, [1, 4])


def test_palabos_multiblock_fixture(gpu):
    """The 6-block Palabos run (SparseBlockStructure3D, momentum-exchange sphere,
    periodic in x) on the same 6 blocks: bit-identical populations and drag."""
    case, init, data = load("multiblock_mebb_f64", GOLDEN_DIR)
    lat = _lattice_with_blocks(case, MULTIBLOCK_BLOCKS)
    try:
        lat.setPopulations(init)
        lat.collideAndStream(10)
        assert np.array_equal(lat.getPopulations(), data["s10"])
        np.testing.assert_allclose(lat.getStatistics()["force"], data["stats_s10"][4:7], rtol=1e-10, atol=1e-14)
    finally:
        lat.close()


def test_sixty_seeded_blocks(gpu):
    """60 uneven blocks (5 x 4 x 3 pieces, edges drawn in [6, 16]) on a 48 x 40 x 30
    porous lattice with a body force, FP16 storage, against the oracle."""
    shape = (48, 40, 30)
    blocks = seeded_uneven_blocks(shape, (5, 4, 3), 6, 16, seed=11)
    assert len(blocks) == 60
    case = Case(*shape, storage="f16", dynamics="rbgk", omega=1.7, periodic=(True, True, True),
                force=(1e-5, 0, 0))
    case.tags = sphere_tags(shape, (20.5, 18.5, 14.5), 6.2, TAG_WALL)
    _check(case, blocks, [3, 10])


@pytest.mark.parametrize("storage", ["f64", "f16"])
def test_coalesced_tiling_runs_as_one_block(gpu, storage):
    """Without keep_blocks a tiling below 2^32 cells runs as one block: the
    60-block porous lattice (link walls, RBGK + Guo) reports one block and
    equals the oracle bit for bit."""
    shape = (48, 40, 30)
    blocks = seeded_uneven_blocks(shape, (5, 4, 3), 6, 16, seed=11)
    case = Case(*shape, storage=storage, dynamics="rbgk", omega=1.7, periodic=(True, True, True),
                force=(1e-5, 0, 0))
    case.tags = sphere_tags(shape, (20.5, 18.5, 14.5), 6.2, TAG_WALL)
    _check(case, blocks, [3, 10], keep=False, expect_blocks=1)


def test_partial_merge_under_a_cell_limit(gpu, monkeypatch):
    """Above the per-block cell limit (lowered here to 1500 cells) only
    z-neighbours with equal x and y extents merge: twelve 8 x 10 x 9 blocks run
    as six of 8 x 10 x 18, bit-identical to the oracle."""
    monkeypatch.setenv("PLBAMD_MAX_BLOCK_CELLS", "1500")
    case = Case(24, 20, 18, storage="f64", omega=1.4, periodic=(True, False, True))
    case.tags = sphere_tags(case.shape, (11.5, 9.5, 8.5), 4.2, TAG_WALL)
    blocks = uneven_blocks(case.shape, [[8, 16], [10], [9]])
    assert len(blocks) == 12
    _check(case, blocks, [2, 7], keep=False, expect_blocks=6)


def test_upload_after_odd_step_is_refused(gpu):
    case = Case(10, 10, 10, storage="f64", omega=1.2, periodic=(True, True, True))
    lat = _lattice_with_blocks(case, uneven_blocks(case.shape, [[5], [], []]))
    try:
        lat.collideAndStream(1)
        with pytest.raises(palabos_amd.PlbError):
            lat.setPopulations(np.zeros(case.shape + (19,)))
        lat.collideAndStream(1)
        lat.setPopulations(np.zeros(case.shape + (19,)))
    finally:
        lat.close()


def test_bad_tilings_are_refused(gpu):
    case = Case(8, 8, 8, storage="f64", omega=1.2)
    with pytest.raises(palabos_amd.PlbError):
        _lattice_with_blocks(case, [(0, 3, 0, 7, 0, 7), (3, 7, 0, 7, 0, 7)])   # overlap
    with pytest.raises(palabos_amd.PlbError):
        _lattice_with_blocks(case, [(0, 3, 0, 7, 0, 7)])                       # gap
