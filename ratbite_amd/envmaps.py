"""Seeded synthetic lat-long environment maps (RGB FP32)."""
import numpy as np


def lognormal_envmap(width, height, sigma=2.0, mu=0.0, seed=0, sun=True, sun_pixels=16, sun_intensity=1e4,
                     channels=3):
    """Per-texel log-normal radiance (independent RGB), plus an optional sun:
    a compact disc of `sun_pixels` texels at intensity `sun_intensity`."""
    rng = np.random.default_rng(seed)
    img = rng.lognormal(mu, sigma, size=(height, width, 3)).astype(np.float32)
    if sun:
        cy = int(rng.integers(height // 8, height // 2))
        cx = int(rng.integers(0, width))
        r = max(1.0, np.sqrt(sun_pixels / np.pi))
        ys, xs = np.mgrid[-int(np.ceil(r)):int(np.ceil(r)) + 1, -int(np.ceil(r)):int(np.ceil(r)) + 1]
        pts = [(y, x) for y, x in zip(ys.ravel(), xs.ravel())]
        pts.sort(key=lambda q: q[0] * q[0] + q[1] * q[1])
        for y, x in pts[:sun_pixels]:
            img[(cy + y) % height, (cx + x) % width, :] = sun_intensity
    if channels == 4:
        img = np.concatenate([img, np.ones((height, width, 1), np.float32)], axis=2)
    return np.ascontiguousarray(img)
