"""Checkpoint files (saveBinaryBlock / loadBinaryBlock) through the native
library, against the files the compiled reference wrote
(tests/golden/checkpoint_*.npz) and the restatement oracle/checkpoint.py."""
import glob
import os

import numpy as np
import pytest

import palabos_amd as plb
from palabos_amd.native import PlbError
from cases import uneven_blocks
from oracle import checkpoint

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIXTURES = sorted(glob.glob(os.path.join(ROOT, "tests", "golden", "checkpoint_*.npz")))
FORCE = (1.25e-5, -2.5e-6, 3.0e-7)   # the force externals of the forced fixtures

pytestmark = pytest.mark.gpu


def load(path):
    z = np.load(path, allow_pickle=False)
    T = {"double": "f64", "float": "f32"}[str(z["T"])]
    return (T, bool(z["forced"]), bool(z["enforce_uint"]), tuple(int(v) for v in z["shape"]), z["records"],
            z["text"].tobytes())


def lattice(shape, storage, forced, blocks=None, periodic=False):
    return plb.Block3D(*shape, plb.D3Q19, plb.BGK(1.7), storage=storage, periodic=periodic,
                       force=FORCE if forced else None, blocks=blocks, keep_blocks=blocks is not None)


def tiling(shape):
    """Two blocks along every axis of at least 4 cells."""
    return uneven_blocks(shape, [[n // 2] if n >= 4 else [] for n in shape])


@pytest.mark.parametrize("tiled", [False, True], ids=["one_block", "blocks"])
@pytest.mark.parametrize("path", FIXTURES, ids=lambda p: os.path.basename(p)[:-4])
def test_save_matches_reference_file(gpu, tmp_path, path, tiled):
    T, forced, uint, shape, rec, text = load(path)
    print(f"[case] save {os.path.basename(path)} tiled={tiled}", flush=True)
    lat = lattice(shape, T, forced, tiling(shape) if tiled else None)
    try:
        lat.setPopulations(np.ascontiguousarray(rec[..., :19]))
        out = tmp_path / "lattice.dat"
        lat.saveBinaryBlock(str(out), T, forceExternals=forced, enforceUint=uint)
        assert out.read_bytes() == text
    finally:
        lat.close()


@pytest.mark.parametrize("tiled", [False, True], ids=["one_block", "blocks"])
@pytest.mark.parametrize("path", FIXTURES, ids=lambda p: os.path.basename(p)[:-4])
def test_load_reference_file(gpu, tmp_path, path, tiled):
    T, forced, uint, shape, rec, text = load(path)
    print(f"[case] load {os.path.basename(path)} tiled={tiled}", flush=True)
    src = tmp_path / "ref.dat"
    src.write_bytes(text)
    lat = lattice(shape, T, forced, tiling(shape) if tiled else None)
    try:
        lat.loadBinaryBlock(str(src), T, forceExternals=forced, enforceUint=uint)
        got = lat.getPopulations(dtype=rec.dtype)
        np.testing.assert_array_equal(got, rec[..., :19])
    finally:
        lat.close()


@pytest.mark.parametrize("storage", ["f16", "bf16", "f32"])
def test_save_of_narrower_storage(gpu, tmp_path, storage):
    """16-bit and FP32 storage saved as an FP64 file: the stored values, widened
    exactly, coded as the restatement codes them; loading restores them."""
    shape = (6, 5, 8)
    rng = np.random.RandomState(3)
    lat = lattice(shape, storage, False)
    lat2 = lattice(shape, storage, False)
    try:
        lat.setPopulations(rng.standard_normal(shape + (19,)) * 0.02)
        stored = lat.getPopulations()
        out = tmp_path / "narrow.dat"
        lat.saveBinaryBlock(str(out), "f64")
        assert out.read_bytes() == checkpoint.encode(stored)
        lat2.loadBinaryBlock(str(out), "f64")
        np.testing.assert_array_equal(lat2.getPopulations(), stored)
    finally:
        lat.close()
        lat2.close()


def test_restart_continues_the_run(gpu, tmp_path):
    """A checkpoint of a multi-block run after 6 steps, loaded into a one-block
    lattice, continues exactly as the original run."""
    shape = (12, 10, 14)
    rng = np.random.RandomState(8)
    f0 = rng.standard_normal(shape + (19,)) * 0.01
    a = plb.Block3D(*shape, plb.D3Q19, plb.RegularizedBGK(1.6), storage="f64", periodic=(True, True, True),
                    blocks=uneven_blocks(shape, [[5], [4], [6]]), keep_blocks=True)
    b = plb.Block3D(*shape, plb.D3Q19, plb.RegularizedBGK(1.6), storage="f64", periodic=(True, True, True))
    try:
        a.setPopulations(f0)
        a.collideAndStream(6)
        out = tmp_path / "restart.dat"
        a.saveBinaryBlock(str(out))
        b.loadBinaryBlock(str(out))
        a.collideAndStream(5)
        b.collideAndStream(5)
        np.testing.assert_array_equal(a.getPopulations(), b.getPopulations())
    finally:
        a.close()
        b.close()


def test_round_trip_large(gpu, tmp_path):
    """Size-independent property at 3.9 M cells (FP32, 297 MB of cells): save,
    load into a fresh lattice, identical populations; the file length is the
    base64 length of the data."""
    shape = (192, 160, 128)
    rng = np.random.RandomState(1)
    f0 = (rng.standard_normal(shape + (19,)) * 0.02).astype(np.float32)
    a = lattice(shape, "f32", False)
    b = lattice(shape, "f32", False)
    try:
        a.setPopulations(f0)
        out = tmp_path / "large.dat"
        a.saveBinaryBlock(str(out), "f32")
        assert out.stat().st_size == 12 + (f0.nbytes + 2) // 3 * 4
        b.loadBinaryBlock(str(out), "f32")
        np.testing.assert_array_equal(b.getPopulations(dtype=np.float32), f0)
    finally:
        a.close()
        b.close()


def test_load_errors(gpu, tmp_path):
    T, forced, uint, shape, rec, text = load([p for p in FIXTURES if "f64_plain" in p][0])
    lat = lattice(shape, "f64", False)
    try:
        good = tmp_path / "good.dat"
        good.write_bytes(text)
        with pytest.raises(PlbError, match="bytes of cells"):
            lat.loadBinaryBlock(str(good), "f32")                 # size header disagrees
        with pytest.raises(PlbError, match="bytes of cells"):
            lat.loadBinaryBlock(str(good), "f64", forceExternals=True)
        cut = tmp_path / "cut.dat"
        cut.write_bytes(text[:-100])
        with pytest.raises(PlbError, match="truncated"):
            lat.loadBinaryBlock(str(cut), "f64")
        bad = tmp_path / "bad.dat"
        bad.write_bytes(text[:500] + b"*" + text[501:])
        with pytest.raises(PlbError, match="not base64"):
            lat.loadBinaryBlock(str(bad), "f64")
        with pytest.raises(PlbError, match="cannot open"):
            lat.loadBinaryBlock(str(tmp_path / "missing.dat"), "f64")
        with pytest.raises(ValueError):
            lat.saveBinaryBlock(str(tmp_path / "x.dat"), "f16")
        # white space in the text is skipped
        spaced = tmp_path / "spaced.dat"
        spaced.write_bytes(b"\n".join(text[k:k + 64] for k in range(0, len(text), 64)) + b"\n")
        lat.loadBinaryBlock(str(spaced), "f64")
        np.testing.assert_array_equal(lat.getPopulations(), rec)
    finally:
        lat.close()
    # force externals that differ from the lattice's force
    T, forced, uint, shape, rec, text = load([p for p in FIXTURES if "f64_forced" in p][0])
    other = plb.Block3D(*shape, plb.D3Q19, plb.BGK(1.7), storage="f64", force=(2e-5, 0, 0))
    try:
        src = tmp_path / "forced.dat"
        src.write_bytes(text)
        with pytest.raises(PlbError, match="force externals"):
            other.loadBinaryBlock(str(src), "f64", forceExternals=True, enforceUint=uint)
    finally:
        other.close()
