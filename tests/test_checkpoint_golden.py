"""The checkpoint-file restatement (oracle/checkpoint.py) against the files the
compiled reference's saveBinaryBlock wrote (tests/golden/checkpoint_*.npz,
made by tests/golden/make_checkpoint_golden.py), and against what base64 fixes:
length, padding and the size header."""
import glob
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from oracle import checkpoint  # noqa: E402

FIXTURES = sorted(glob.glob(os.path.join(ROOT, "tests", "golden", "checkpoint_*.npz")))


def load(path):
    z = np.load(path, allow_pickle=False)
    return (str(z["T"]), bool(z["forced"]), bool(z["enforce_uint"]), tuple(int(v) for v in z["shape"]),
            z["records"], z["text"].tobytes())


def test_fixtures_present():
    assert len(FIXTURES) == 4


@pytest.mark.parametrize("path", FIXTURES, ids=lambda p: os.path.basename(p)[:-4])
def test_encode_matches_reference_file(path):
    T, forced, uint, shape, rec, text = load(path)
    assert rec.shape == shape + (22 if forced else 19,)
    assert checkpoint.encode(rec, uint) == text


@pytest.mark.parametrize("path", FIXTURES, ids=lambda p: os.path.basename(p)[:-4])
def test_decode_reference_file(path):
    T, forced, uint, shape, rec, text = load(path)
    got = checkpoint.decode(text, rec.dtype, shape, rec.shape[-1], uint)
    np.testing.assert_array_equal(got, rec)
    # white space between characters is skipped, as the reference's reader does
    spaced = b"\n".join(text[k:k + 76] for k in range(0, len(text), 76))
    np.testing.assert_array_equal(checkpoint.decode(spaced, rec.dtype, shape, rec.shape[-1], uint), rec)


@pytest.mark.parametrize("path", FIXTURES, ids=lambda p: os.path.basename(p)[:-4])
def test_reference_file_framing(path):
    """Header: 12 characters for an 8-byte size (one '='), 8 for a 4-byte one
    ('=='); body: 4 * ceil(bytes / 3) characters, '=' only at its end."""
    T, forced, uint, shape, rec, text = load(path)
    nbytes = rec.nbytes
    head = 8 if uint else 12
    assert len(text) == head + (nbytes + 2) // 3 * 4
    assert text[:head].endswith(b"==" if uint else b"=") and b"=" not in text[:head].rstrip(b"=")
    body = text[head:]
    pad = (3 - nbytes % 3) % 3
    assert body.count(b"=") == pad and (pad == 0 or body.endswith(b"=" * pad))


def test_size_mismatch_refused():
    T, forced, uint, shape, rec, text = load(FIXTURES[0])
    with pytest.raises(ValueError):
        checkpoint.decode(text, rec.dtype, (shape[0] + 1,) + shape[1:], rec.shape[-1], uint)
