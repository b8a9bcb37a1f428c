"""CPU restatement of the Palabos checkpoint file (saveBinaryBlock /
loadBinaryBlock), the checker of the device conversion.

Test infrastructure only: tests/ use it as the checker; the product never
imports it.

Format, following the reference:
  * the file is the text of serializerToBase64Stream
    (io/serializerIO.cpp:188-194): Base64Writer::writeHeader writes the data size
    as an 8-byte unsigned (pluint) or, with enforceUint, a 4-byte unsigned
    (io/serializerIO.cpp:83-101), base64-coded on its own; the data follow as a
    second base64 stream of exactly that many bytes;
  * base64 is RFC 4648's alphabet with '=' padding and no line breaks
    (io/base64.hh:36-38 Base64Encoder::enc64, :96-114 encodeBlock /
    encodeUnfinishedBlock); the reader skips white space (operator>>,
    io/base64.hh:150-155);
  * the data are the cells of the bounding box in C order x, y, z
    (IndexOrdering::forward: multiBlock/multiBlockSerializer3D.cpp:71-105), each
    Cell::serialize: the raw T values of the 19 f-bar, then the external
    scalars (core/cell.hh:103-111; ForcedD3Q19Descriptor: the 3 force
    components); cells in no block are zeros (fillBufferWithZeros, :251-261).

The coding itself is Python's base64 module (the RFC 4648 routine): what the
golden files written by the compiled reference pin is the framing above.
"""
import base64
import struct

import numpy as np


def encode(records, enforce_uint=False):
    """File bytes of the cell records (T array [nx, ny, nz, 19 or 22])."""
    data = np.ascontiguousarray(records).astype(records.dtype.newbyteorder("<"), copy=False).tobytes()
    head = struct.pack("<I", len(data)) if enforce_uint else struct.pack("<Q", len(data))
    return base64.b64encode(head) + base64.b64encode(data)


def decode(text, dtype, shape, rec_len, enforce_uint=False):
    """Cell records of file bytes; raises ValueError on a size mismatch."""
    text = bytes(b for b in bytes(text) if b not in b" \t\r\n\v\f")
    head_chars = 8 if enforce_uint else 12
    head = base64.b64decode(text[:head_chars])
    size = struct.unpack("<I" if enforce_uint else "<Q", head[:4 if enforce_uint else 8])[0]
    dtype = np.dtype(dtype)
    want = int(np.prod(shape)) * rec_len * dtype.itemsize
    if size != want:
        raise ValueError(f"the file holds {size} bytes of cells, expected {want}")
    n_chars = (size + 2) // 3 * 4
    data = base64.b64decode(text[head_chars:head_chars + n_chars], validate=True)
    return np.frombuffer(data[:size], dtype=dtype.newbyteorder("<")).reshape(tuple(shape) + (rec_len,))
