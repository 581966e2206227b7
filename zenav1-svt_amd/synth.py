"""Seeded synthetic still-image content for the benchmark workloads.

All content is generated once on the host (integer / deterministic double
math) and uploaded, so every backend sees identical samples.  Seeds are
config_index * 1000 + image_index.  Small parameter draws use
splitmix64 -> xorshift64* (multiplier 0x2545F4914F6CDD1D, the generator the
reference's differential tests use); bulk noise uses numpy's PCG64 seeded
from the same splitmix64 stream.
"""
import numpy as np

M64 = (1 << 64) - 1


def splitmix64(x):
    x = (x + 0x9E3779B97F4A7C15) & M64
    z = x
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return x, z ^ (z >> 31)


class XorShift64Star:
    def __init__(self, seed):
        _, s = splitmix64(seed)
        self.s = s or 1

    def next(self):
        x = self.s
        x ^= x >> 12
        x ^= (x << 25) & M64
        x ^= x >> 27
        self.s = x
        return (x * 0x2545F4914F6CDD1D) & M64

    def uniform(self):
        return (self.next() >> 11) / float(1 << 53)

    def randint(self, lo, hi):
        """Uniform integer in [lo, hi]."""
        return lo + int(self.uniform() * (hi - lo + 1))


def _np_rng(seed):
    _, s = splitmix64(seed ^ 0xA5A5A5A5)
    return np.random.Generator(np.random.PCG64(s))


def spectral_noise(h, w, alpha, seed):
    """Real field with power spectrum ~ 1/f^alpha (seeded random phases)."""
    rng = _np_rng(seed)
    fy = np.fft.fftfreq(h)[:, None]
    fx = np.fft.rfftfreq(w)[None, :]
    f = np.sqrt(fx * fx + fy * fy)
    f[0, 0] = 1.0
    amp = f ** (-alpha / 2.0)
    amp[0, 0] = 0.0
    phase = rng.uniform(0, 2 * np.pi, amp.shape)
    spec = amp * np.exp(1j * phase)
    field = np.fft.irfft2(spec, s=(h, w))
    field -= field.min()
    mx = field.max()
    return field / (mx if mx > 0 else 1.0)


def _draw_lines(img, n, wmin, wmax, lo, hi, rnd):
    h, w = img.shape
    yy, xx = np.mgrid[0:h, 0:w]
    for _ in range(n):
        cx, cy = rnd.uniform() * w, rnd.uniform() * h
        th = rnd.uniform() * np.pi
        width = rnd.randint(wmin, wmax)
        val = rnd.randint(lo, hi)
        d = np.abs((xx - cx) * np.sin(th) - (yy - cy) * np.cos(th))
        img[d < width / 2.0] = val


def config2_luma(h=2048, w=2048, seed=2000):
    """1/f^2 spectral noise scaled to [64, 940] plus 30 random-orientation
    hard lines of width 1-4 px (10-bit)."""
    f = spectral_noise(h, w, 2.0, seed)
    img = np.round(64 + f * (940 - 64)).astype(np.int32)
    _draw_lines(img, 30, 1, 4, 64, 940, XorShift64Star(seed))
    return np.clip(img, 0, 1023).astype(np.uint16)


def config1_luma(h=512, w=512, seed=1000):
    """Sum of 3 seeded linear gradients + 20 rectangles U[16,235] + N(0, 4^2)."""
    rnd = XorShift64Star(seed)
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    img = np.zeros((h, w))
    for _ in range(3):
        gx, gy = rnd.uniform() - 0.5, rnd.uniform() - 0.5
        img += (gx * xx / w + gy * yy / h) * 120 + 40
    img /= 3.0
    img += 60
    for _ in range(20):
        x0, y0 = rnd.randint(0, w - 8), rnd.randint(0, h - 8)
        x1, y1 = min(w, x0 + rnd.randint(8, w // 3)), min(h, y0 + rnd.randint(8, h // 3))
        img[y0:y1, x0:x1] = rnd.randint(16, 235)
    img += _np_rng(seed).normal(0, 4, (h, w))
    return np.clip(np.round(img), 0, 255).astype(np.uint8)


def config3_recon(h=4096, w=4096, seed=3000):
    """Loop-filter input: 8-bit 1/f noise source, and a reconstruction with
    8x8 mean-flattening on 30% of the 8x8 blocks (blocky artifacts)."""
    f = spectral_noise(h, w, 1.0, seed)
    src = np.clip(np.round(16 + f * 219), 0, 255).astype(np.uint8)
    rng = _np_rng(seed + 1)
    rec = src.astype(np.int32).copy()
    bh, bw = h // 8, w // 8
    mask = rng.uniform(size=(bh, bw)) < 0.30
    blocks = rec.reshape(bh, 8, bw, 8)
    means = np.round(blocks.mean(axis=(1, 3))).astype(np.int32)
    flat = np.repeat(np.repeat(means, 8, 0), 8, 1)
    m = np.repeat(np.repeat(mask, 8, 0), 8, 1)
    rec[m] = flat[m]
    rec = rec + rng.integers(-2, 3, rec.shape)
    return src, np.clip(rec, 0, 255).astype(np.uint8)


def qp_to_qindex(qp):
    """CQP qp -> qindex (rust/crates/svtav1-encoder/src/rate_control.rs:129)."""
    if qp < 62:
        return 4 * qp
    return 249 if qp == 62 else 255
