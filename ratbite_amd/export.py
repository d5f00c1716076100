"""PNG export of a resolved frame.

The reference saves the resolved RGBA8 image (the renderResolve output read
back from its GL texture) with stb_image_write, flipped vertically because
row 0 of the film is its bottom row (src/core/ui.cpp:894-912). write_png
produces the same file content layout: 8-bit RGBA, rows from the top of the
picture (the film's last row) down, no filtering, zlib-compressed.
"""
import struct
import zlib

import numpy as np


def _chunk(kind, data):
    body = kind + data
    return struct.pack(">I", len(data)) + body + struct.pack(">I", zlib.crc32(body) & 0xFFFFFFFF)


def png_bytes(rgba, flip=True):
    """Encodes an (H, W, 4) uint8 array; flip=True writes the film's last row
    first, as the reference's export does."""
    a = np.ascontiguousarray(rgba, dtype=np.uint8)
    if a.ndim != 3 or a.shape[2] != 4:
        raise ValueError("expected an (H, W, 4) RGBA8 array, got shape %s" % (a.shape,))
    h, w = a.shape[:2]
    rows = a[::-1] if flip else a
    raw = np.zeros((h, 1 + 4 * w), np.uint8)        # filter byte 0 (None) per row
    raw[:, 1:] = rows.reshape(h, 4 * w)
    ihdr = struct.pack(">IIBBBBB", w, h, 8, 6, 0, 0, 0)   # 8-bit, colour type 6 (RGBA)
    return b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", ihdr) + _chunk(b"IDAT", zlib.compress(raw.tobytes(), 6)) + \
        _chunk(b"IEND", b"")


def write_png(path, rgba, flip=True):
    with open(path, "wb") as f:
        f.write(png_bytes(rgba, flip))
