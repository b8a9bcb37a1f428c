"""Writes the checkpoint-file golden vectors of tests/golden/ by running the
compiled Palabos reference (oracle/_ref/ref_driver --checkpoint, built by
oracle/Makefile.ref from the reference's own sources): seeded cell records are
set into a MultiBlockLattice3D and written by saveBinaryBlock
(io/serializerIO_3D.cpp:37-49); the file is read back by loadBinaryBlock to
check that the reference restores the records.

Each fixture checkpoint_<name>.npz holds the element type, descriptor, header
kind and shape, the cell records (T values, [nx][ny][nz][19 or 22]) and the
file's bytes.

Run from the repository root:  python tests/golden/make_checkpoint_golden.py
"""
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import reference  # noqa: E402

# name: (T, forced, enforceUint, shape, seed); shapes whose byte counts leave
# 0, 1 and 2 bytes over a base64 group
CASES = {
    "f64_plain": ("double", False, False, (4, 5, 7), 11),
    "f32_forced_uint": ("float", True, True, (4, 5, 7), 12),
    "f64_forced": ("double", True, False, (3, 4, 5), 13),
    "f32_plain_uint": ("float", False, True, (6, 3, 5), 14),
}
FORCE = (1.25e-5, -2.5e-6, 3.0e-7)


def records(T, forced, shape, seed):
    rng = np.random.RandomState(seed)
    dt = np.float64 if T == "double" else np.float32
    f = (rng.standard_normal(shape + (19,)) * 0.02).astype(dt)
    if not forced:
        return f
    ext = np.broadcast_to(np.array(FORCE, dtype=dt), shape + (3,))
    return np.concatenate([f, ext], axis=-1)


def main():
    if not reference.available():
        raise SystemExit("oracle/_ref/ref_driver is missing: build it with make -f oracle/Makefile.ref")
    for name, (T, forced, uint, shape, seed) in CASES.items():
        rec = records(T, forced, shape, seed)
        with tempfile.TemporaryDirectory() as tmp:
            src, dat, back = (os.path.join(tmp, n) for n in ("rec.bin", "out.dat", "back.bin"))
            rec.tofile(src)
            reference.checkpoint("save", T, forced, shape, src, dat, uint)
            reference.checkpoint("load", T, forced, shape, dat, back, uint)
            text = np.frombuffer(open(dat, "rb").read(), dtype=np.uint8)
            restored = np.fromfile(back, dtype=rec.dtype).reshape(rec.shape)
        if not np.array_equal(restored, rec):
            raise SystemExit(f"{name}: loadBinaryBlock did not restore the records")
        np.savez_compressed(os.path.join(HERE, f"checkpoint_{name}.npz"), T=T, forced=forced, enforce_uint=uint,
                            shape=np.array(shape), records=rec, text=text)
        print(name, rec.shape, len(text), "characters")


if __name__ == "__main__":
    main()
