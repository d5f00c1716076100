"""Radiance RGBE (.hdr) environment maps.

The reference loads environment maps with stb_image's stbi_loadf (v2.30,
vendored under its dependencies/; src/core/rendering_interface.cpp:1349),
four channels. This module restates that decoder's behaviour for .hdr files:

  * header: "#?RADIANCE" or "#?RGBE", lines up to a blank line, FORMAT must be
    32-bit_rle_rgbe; the resolution line must be "-Y <height> +X <width>";
  * scanlines: flat RGBE quadruples, or (for widths 8..32767) the "new" RLE
    form (2, 2, width >> 8, width & 255, then each of the four channels as
    runs (count > 128: repeat the next byte count - 128 times) and literals);
    a scanline that does not open with that marker is read flat, together
    with the rest of the image;
  * a texel (r, g, b, e) decodes to (r, g, b) * 2^(e - 136) (no half-unit
    offset), 0 when e == 0; alpha is 1.

write_hdr encodes float images the usual way (frexp of the largest channel)
for tests and callers.
"""
import numpy as np


def _decode(rgbe):
    rgbe = np.asarray(rgbe, np.uint8)
    e = rgbe[..., 3].astype(np.int32)
    f = np.ldexp(np.float32(1.0), e - 136).astype(np.float32)
    out = np.empty(rgbe.shape[:-1] + (4,), np.float32)
    out[..., :3] = rgbe[..., :3].astype(np.float32) * f[..., None]
    out[..., :3][e == 0] = 0.0
    out[..., 3] = 1.0
    return out


def parse_hdr(data):
    """Decodes the bytes of a Radiance .hdr file into (height, width, 4)
    float32 texels (top row first)."""
    mv = memoryview(data)
    pos = 0

    def line():
        nonlocal pos
        end = data.index(b"\n", pos)
        s = bytes(mv[pos:end]).decode("latin-1")
        pos = end + 1
        return s

    magic = line()
    if magic not in ("#?RADIANCE", "#?RGBE"):
        raise ValueError("not a Radiance HDR file")
    fmt_ok = False
    while True:
        s = line()
        if s == "":
            break
        if s == "FORMAT=32-bit_rle_rgbe":
            fmt_ok = True
    if not fmt_ok:
        raise ValueError("unsupported HDR format (want FORMAT=32-bit_rle_rgbe)")
    parts = line().split()
    if len(parts) != 4 or parts[0] != "-Y" or parts[2] != "+X":
        raise ValueError("unsupported HDR orientation (want -Y h +X w)")
    h, w = int(parts[1]), int(parts[3])
    if w <= 0 or h <= 0:
        raise ValueError("bad HDR size")
    buf = np.frombuffer(data, np.uint8, offset=pos)
    if w < 8 or w >= 32768:
        flat = buf[:w * h * 4]
        if len(flat) < w * h * 4:
            raise ValueError("truncated HDR data")
        return _decode(flat.reshape(h, w, 4))
    out = np.zeros((h, w, 4), np.uint8)
    p = 0
    for j in range(h):
        if p + 4 > len(buf):
            raise ValueError("truncated HDR data")
        c1, c2, c3 = int(buf[p]), int(buf[p + 1]), int(buf[p + 2])
        if c1 != 2 or c2 != 2 or (c3 & 0x80):
            # not run-length encoded: this and every following texel is flat
            rest = buf[p:p + (h - j) * w * 4]
            if len(rest) < (h - j) * w * 4:
                raise ValueError("truncated HDR data")
            out[j:] = rest.reshape(h - j, w, 4)
            return _decode(out)
        if ((c3 << 8) | int(buf[p + 3])) != w:
            raise ValueError("invalid HDR scanline width")
        p += 4
        for k in range(4):
            i = 0
            while i < w:
                if p >= len(buf):
                    raise ValueError("truncated HDR data")
                count = int(buf[p])
                p += 1
                if count > 128:
                    count -= 128
                    if count > w - i:
                        raise ValueError("bad RLE data in HDR")
                    out[j, i:i + count, k] = buf[p]
                    p += 1
                else:
                    if count == 0 or count > w - i:
                        raise ValueError("bad RLE data in HDR")
                    out[j, i:i + count, k] = buf[p:p + count]
                    p += count
                i += count
    return _decode(out)


def read_hdr(path):
    with open(path, "rb") as f:
        return parse_hdr(f.read())


def encode_rgbe(img):
    """(h, w, >=3) float -> (h, w, 4) uint8 RGBE."""
    rgb = np.asarray(img, np.float64)[..., :3]
    v = rgb.max(axis=-1)
    m, e = np.frexp(v)
    scale = np.where(v > 1e-32, m * 256.0 / np.where(v > 1e-32, v, 1.0), 0.0)
    out = np.zeros(rgb.shape[:-1] + (4,), np.uint8)
    out[..., :3] = np.clip(np.floor(rgb * scale[..., None]), 0, 255).astype(np.uint8)
    out[..., 3] = np.where(v > 1e-32, np.clip(e + 128, 0, 255), 0).astype(np.uint8)
    return out


def _rle_channel(row):
    """Run-length encodes one scanline channel (runs of >= 4 as runs)."""
    out = bytearray()
    n = len(row)
    i = 0
    while i < n:
        # a run starting here?
        j = i
        while j < n and j - i < 127 and row[j] == row[i]:
            j += 1
        if j - i >= 4:
            out += bytes([128 + (j - i), int(row[i])])
            i = j
            continue
        # literal up to the next run of 4 (at most 128)
        k = i
        while k < n and k - i < 128:
            if k + 3 < n and row[k] == row[k + 1] == row[k + 2] == row[k + 3]:
                break
            k += 1
        out += bytes([k - i]) + bytes(row[i:k].tolist())
        i = k
    return bytes(out)


def write_hdr(path, img, rle=True):
    rgbe = encode_rgbe(img)
    h, w = rgbe.shape[:2]
    with open(path, "wb") as f:
        f.write(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n")
        f.write(("-Y %d +X %d\n" % (h, w)).encode())
        if not rle or w < 8 or w >= 32768:
            f.write(rgbe.tobytes())
            return
        for j in range(h):
            f.write(bytes([2, 2, w >> 8, w & 255]))
            for k in range(4):
                f.write(_rle_channel(rgbe[j, :, k]))
