"""PNG export (the reference's Save button, src/core/ui.cpp:894-912: the
resolved RGBA8 film written with stb_image_write, flipped vertically)."""
import struct
import zlib

import numpy as np
import pytest

from ratbite_amd.export import png_bytes


def _decode(data):
    """Minimal PNG reader for what png_bytes writes (checks every CRC)."""
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, chunks = 8, []
    while pos < len(data):
        n, = struct.unpack(">I", data[pos:pos + 4])
        kind, body = data[pos + 4:pos + 8], data[pos + 8:pos + 8 + n]
        crc, = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])
        assert crc == zlib.crc32(kind + body) & 0xFFFFFFFF
        chunks.append((kind, body))
        pos += 12 + n
    assert [k for k, _ in chunks] == [b"IHDR", b"IDAT", b"IEND"]
    w, h, depth, ctype, comp, filt, inter = struct.unpack(">IIBBBBB", chunks[0][1])
    assert (depth, ctype, comp, filt, inter) == (8, 6, 0, 0, 0)
    raw = np.frombuffer(zlib.decompress(chunks[1][1]), np.uint8).reshape(h, 1 + 4 * w)
    assert np.all(raw[:, 0] == 0)
    return raw[:, 1:].reshape(h, w, 4)


@pytest.mark.parametrize("shape", [(1, 1), (3, 5), (64, 48)])
def test_png_round_trip_flips_rows(shape):
    rng = np.random.default_rng(sum(shape))
    a = rng.integers(0, 256, size=shape + (4,), dtype=np.uint8)
    img = _decode(png_bytes(a))
    assert np.array_equal(img, a[::-1])
    assert np.array_equal(_decode(png_bytes(a, flip=False)), a)


def test_png_rejects_wrong_layout():
    with pytest.raises(ValueError):
        png_bytes(np.zeros((4, 4, 3), np.uint8))


@pytest.mark.gpu
def test_save_png_of_a_render(tmp_path):
    from ratbite_amd import scenes
    from ratbite_amd.renderer import Renderer
    r = Renderer(scenes.cornell(width=40, height=24, spp=4))
    try:
        r.render(spp=4, offset=0)
        p = tmp_path / "frame.png"
        r.save_png(str(p), exposure=1.0)
        img = _decode(p.read_bytes())
        assert np.array_equal(img, r.image(exposure=1.0)[::-1])
        assert img[..., :3].max() > 0
    finally:
        r.close()
