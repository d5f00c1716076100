"""Radiance .hdr reading (ratbite_amd/hdr.py) against the format's
definition and stb_image's decoding rule (what the reference's
stbi_loadf returns): RGBE bytes decode to (r, g, b) 2^(e - 136) with alpha 1
and e = 0 giving black; run-length and flat scanlines decode alike; a
scanline without the run-length marker switches the rest of the file to flat
texels; the encoder's round trip stays within the 8-bit mantissa; headers
with other formats or orientations are refused."""
import numpy as np
import pytest

from ratbite_amd import hdr


def _file(tmp_path, body, w, h, header=b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n", res=None):
    p = tmp_path / "x.hdr"
    p.write_bytes(header + (res or ("-Y %d +X %d\n" % (h, w)).encode()) + body)
    return str(p)


def test_decoding_rule(tmp_path):
    px = np.array([[128, 64, 32, 136], [255, 0, 1, 129], [10, 20, 30, 0], [1, 1, 1, 200]], np.uint8)
    img = hdr.read_hdr(_file(tmp_path, px.tobytes(), 2, 2))
    want = np.array([[128, 64, 32], [255 * 2.0 ** -7, 0, 2.0 ** -7], [0, 0, 0], [2.0 ** 64] * 3], np.float64)
    assert img.shape == (2, 2, 4) and img.dtype == np.float32
    assert np.array_equal(img[..., :3].reshape(-1, 3), want.astype(np.float32))
    assert np.all(img[..., 3] == 1.0)


@pytest.mark.parametrize("shape", [(4, 8), (3, 37), (5, 300)])
def test_round_trip_and_rle_equals_flat(tmp_path, shape):
    rng = np.random.default_rng(shape[1])
    img = rng.lognormal(0.0, 2.0, shape + (3,)).astype(np.float32)
    img[0, :4] = 0.0
    img[1, :] = img[1, 0]                       # long runs
    a, b = str(tmp_path / "rle.hdr"), str(tmp_path / "flat.hdr")
    hdr.write_hdr(a, img, rle=True)
    hdr.write_hdr(b, img, rle=False)
    ra, rb = hdr.read_hdr(a), hdr.read_hdr(b)
    assert np.array_equal(ra, rb)
    mx = img.max(axis=2, keepdims=True)
    assert np.all(np.abs(ra[..., :3] - img) <= mx * 2.0 ** -7 + 1e-30)
    if shape[1] >= 8:
        with open(a, "rb") as f:
            assert len(f.read()) < len(open(b, "rb").read())


def test_unmarked_scanline_reads_the_rest_flat(tmp_path):
    w, h = 8, 3
    rng = np.random.default_rng(3)
    rgbe = rng.integers(1, 255, (h, w, 4)).astype(np.uint8)
    rgbe[..., 3] = 130
    rgbe[1, 0, :2] = 1                         # the second scanline does not start with (2, 2)
    row0 = bytes([2, 2, 0, w])
    for k in range(4):
        row0 += bytes([w]) + rgbe[0, :, k].tobytes()
    img = hdr.read_hdr(_file(tmp_path, row0 + rgbe[1:].tobytes(), w, h))
    want = rgbe[..., :3].astype(np.float32) * np.float32(2.0 ** -6)
    assert np.array_equal(img[..., :3], want)


def test_refused_headers(tmp_path):
    with pytest.raises(ValueError):
        hdr.read_hdr(_file(tmp_path, b"\0" * 16, 2, 2, header=b"#?OTHER\n\n"))
    with pytest.raises(ValueError):
        hdr.read_hdr(_file(tmp_path, b"\0" * 16, 2, 2, header=b"#?RADIANCE\nFORMAT=32-bit_rle_xyze\n\n"))
    with pytest.raises(ValueError):
        hdr.read_hdr(_file(tmp_path, b"\0" * 16, 2, 2, res=b"+Y 2 +X 2\n"))
    with pytest.raises(ValueError):
        hdr.read_hdr(_file(tmp_path, b"\0" * 8, 2, 2))          # truncated
